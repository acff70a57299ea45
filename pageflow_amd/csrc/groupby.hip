// Exact int64 SUM and COUNT per group, and the serving tier's fused
// fact->dimension join + group-by.
//
// The reference's group-by (operators/aggregate.py:1511-1563 pass 1 ->
// operators/operator_utils.py:342-399) sums through float64 bincount below
// 2^16 rows a group and a two-step int64 scatter_reduce above; its result
// equals the exact int64 sum while every partial stays below 2^53.  Here the
// sums are exact int64 throughout (integer atomics are order-free, so the
// result is bit-identical whatever the schedule), and COUNT counts the
// values that are not the null sentinel (-64).
//
// gfx950 design: each workgroup accumulates into LDS copies of the group
// table (R replicas, lane % R picks one, laid out replica-minor so that a
// wave's adds spread over the banks whatever their groups), reduces the
// replicas and adds
// its partials to the output with one global int64 atomic per group.
//
// Serving: a dimension table (int32 key -> int32 group in [0, G)) is built
// once and kept resident; each request streams its fact rows (int32 foreign
// key, int64 amount) through one kernel that looks the key up and
// aggregates: 12 bytes a fact row, the algorithmic minimum.  Two layouts:
//   dense   when the keys span at most 4n values (and G <= 255): a direct
//           array of one byte a key value (the group, 0xFF = no row), so the
//           table of a 2M-row dimension is 2 MB and stays in an XCD's L2
//           beside the others (dense surrogate keys are the common case);
//   hashed  otherwise: open addressing over packed 64-bit slots (key << 32
//           | group, empty = all ones, which no valid slot equals because a
//           group is never -1).
// info (int64[4]): [0] groups out of range, [1] repeated keys, [2] the
// smallest key, [3] the dense key range (0: hashed).
#include <hip/hip_runtime.h>

#include <string>

#include "pf_common.hpp"
#include "../../include/pageflow_join.h"

namespace pf {

size_t radix_partition_ws(int64_t n, int bits);
int radix_partition(const void* keys, int key_bytes, const int32_t* vals, int64_t n, int bits,
                    int shift, int mode, void* keys_out, int32_t* vals_out,
                    int64_t* bucket_offsets, void* ws, size_t ws_bytes, hipStream_t st,
                    int who, int shape);

namespace {

constexpr int GB_NT = 512;
constexpr int GB_LDS_GROUPS = 1024;   // groups accumulated in LDS (with replicas)
constexpr uint64_t kEmptySlot = ~0ULL;

// replicas of the LDS table for G groups: as many as fit 32 KiB of int64 pairs
__host__ __device__ inline int replicas_for(int G) {
  int r = 32;
  while (r > 1 && (int64_t)r * G * 16 > 32 * 1024) r >>= 1;
  return r;
}

// LDS group accumulators: R replicas of (SUM, COUNT) for each of G groups,
// laid out [group][sum | count][replica], lane l adding into replica l % R.
// With R a power of two (32 for G <= 64) the adds of one wave instruction
// fall on R consecutive 8-byte words of a row of the table whatever the
// lanes' groups, so two lanes share an LDS bank only when they share a
// replica (a [replica][group] layout puts lanes of different groups on the
// same banks).
__device__ __forceinline__ int64_t* acc_at(int64_t* acc, int R, int rep, int g, int c) {
  return acc + ((int64_t)(2 * g + c) * R + rep);
}

__device__ __forceinline__ void lds_add(int64_t* s, int64_t v) {
  atomicAdd((unsigned long long*)s, (unsigned long long)v);
}

__device__ __forceinline__ void global_add(int64_t* s, int64_t v) {
  atomicAdd((unsigned long long*)s, (unsigned long long)v);
}

// SUM/COUNT of val by gid.  Rows with gid outside [0, G) are skipped (and
// counted into *bad when bad is not null).
template <bool LDS>
__global__ __launch_bounds__(GB_NT) void k_group_sum_count(const int32_t* __restrict__ gid,
                                                           const int64_t* __restrict__ val,
                                                           int64_t n, int G, int R,
                                                           int64_t* sum, int64_t* cnt,
                                                           int64_t* bad) {
  __shared__ int64_t acc[LDS ? 32 * 1024 / 8 : 1];  // [G][2][R]
  const int tid = threadIdx.x;
  if (LDS) {
    for (int x = tid; x < R * G * 2; x += GB_NT) acc[x] = 0;
    __syncthreads();
  }
  const int rep = (int)(lane_id() % R);
  int64_t nbad = 0;
  const int64_t stride = (int64_t)gridDim.x * GB_NT;
  for (int64_t i = blockIdx.x * (int64_t)GB_NT + tid; i < n; i += stride) {
    const int g = gid[i];
    const int64_t v = val[i];
    if (g < 0 || g >= G) {
      ++nbad;
      continue;
    }
    if (LDS) {
      lds_add(acc_at(acc, R, rep, g, 0), v);
      if (v != kNull) lds_add(acc_at(acc, R, rep, g, 1), 1);
    } else {
      global_add(&sum[g], v);
      if (v != kNull) global_add(&cnt[g], 1);
    }
  }
  if (bad) {
    nbad = wave_sum(nbad);
    if (lane_id() == 0 && nbad) global_add(bad, nbad);
  }
  if (LDS) {
    __syncthreads();
    for (int g = tid; g < G; g += GB_NT) {
      int64_t s = 0, c = 0;
      for (int r = 0; r < R; ++r) {
        s += *acc_at(acc, R, r, g, 0);
        c += *acc_at(acc, R, r, g, 1);
      }
      if (s) global_add(&sum[g], s);
      if (c) global_add(&cnt[g], c);
    }
  }
}

// ------------------------------------------------------------ serving --

__device__ __forceinline__ uint32_t dim_slot(int32_t key, uint32_t mask) {
  return (uint32_t)mix64((uint64_t)(int64_t)key) & mask;
}

__global__ void k_dim_init(int64_t* info) {
  if (threadIdx.x < 4) info[threadIdx.x] = threadIdx.x == 2 ? INT32_MAX : (threadIdx.x == 3 ? INT32_MIN : 0);
}

__global__ void k_dim_minmax(const int32_t* __restrict__ key, int64_t n, int64_t* info) {
  int32_t lo = INT32_MAX, hi = INT32_MIN;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) {
    const int32_t k = key[i];
    lo = k < lo ? k : lo;
    hi = k > hi ? k : hi;
  }
  for (int o = kWave / 2; o > 0; o >>= 1) {
    const int32_t a = __shfl_xor(lo, o, kWave), b = __shfl_xor(hi, o, kWave);
    lo = a < lo ? a : lo;
    hi = b > hi ? b : hi;
  }
  if (lane_id() == 0) {
    atomicMin((long long*)&info[2], (long long)lo);
    atomicMax((long long*)&info[3], (long long)hi);
  }
}

// info[3] := the dense key range, or 0 for the hashed layout.
__global__ void k_dim_layout(int64_t n, int64_t slots, int G, int64_t* info) {
  const int64_t range = n > 0 ? info[3] - info[2] + 1 : 0;
  info[3] = (n > 0 && G <= 255 && range <= 4 * n && range <= slots * 8) ? range : 0;
}

// Inserts each dimension row's (key, group).  A repeated key keeps its
// first insertion (dimension keys are unique; a repeat counts in info[1]).
__global__ void k_dim_build(const int32_t* __restrict__ key, const int32_t* __restrict__ grp,
                            int64_t n, uint64_t* table, uint32_t mask, int G, int64_t* info) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t k = key[i], g = grp[i];
  if (g < 0 || g >= G) {
    atomicAdd((unsigned long long*)&info[0], 1ULL);
    return;
  }
  const int64_t range = info[3];
  if (range > 0) {  // dense: byte k - min of the array, set by a CAS on its word
    const uint64_t off = (uint64_t)((int64_t)k - info[2]);
    uint32_t* wp = (uint32_t*)table + (off >> 2);
    const int sh = (int)(off & 3) * 8;
    uint32_t old = *wp;
    for (;;) {
      if (((old >> sh) & 0xFFu) != 0xFFu) {
        atomicAdd((unsigned long long*)&info[1], 1ULL);
        return;
      }
      const uint32_t nw = (old & ~(0xFFu << sh)) | ((uint32_t)g << sh);
      const uint32_t prev = atomicCAS(wp, old, nw);
      if (prev == old) return;
      old = prev;
    }
  }
  const uint64_t v = ((uint64_t)(uint32_t)k << 32) | (uint32_t)g;
  uint32_t s = dim_slot(k, mask);
  for (;;) {
    uint64_t prev = table[s];
    if (prev == kEmptySlot)
      prev = atomicCAS((unsigned long long*)&table[s], (unsigned long long)kEmptySlot,
                       (unsigned long long)v);
    if (prev == kEmptySlot) return;
    if ((int32_t)(prev >> 32) == k) {
      atomicAdd((unsigned long long*)&info[1], 1ULL);
      return;
    }
    s = (s + 1) & mask;
  }
}

// Accumulates fact rows [i0, i1) (stride `step` from i0) into the LDS
// group table: U rows a thread with their first table reads in flight
// together (the rare longer probe sequences resolve afterwards).
__device__ __forceinline__ void fact_rows_into(const int32_t* __restrict__ fk,
                                               const int64_t* __restrict__ amt, int64_t i0,
                                               int64_t i1, int64_t step,
                                               const uint64_t* __restrict__ table, uint32_t mask,
                                               int64_t kmin, int64_t range, int64_t* acc,
                                               int R, int rep) {
  constexpr int U = 8;
  if (range > 0) {  // dense: one byte a key value
    const uint8_t* __restrict__ arr = (const uint8_t*)table;
    for (int64_t b = i0; b < i1; b += step * U) {
      int32_t k[U];
      int64_t v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int64_t i = b + u * step;
        k[u] = i < i1 ? fk[i] : 0;
        v[u] = i < i1 ? amt[i] : 0;
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const uint64_t off = (uint64_t)((int64_t)k[u] - kmin);
        const uint32_t g = b + u * step < i1 && off < (uint64_t)range ? arr[off] : 0xFFu;
        if (g == 0xFFu) continue;
        lds_add(acc_at(acc, R, rep, (int)g, 0), v[u]);
        if (v[u] != kNull) lds_add(acc_at(acc, R, rep, (int)g, 1), 1);
      }
    }
    return;
  }
  for (int64_t b = i0; b < i1; b += step * U) {
    int32_t k[U];
    int64_t v[U];
    uint32_t s[U];
    uint64_t t[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t i = b + u * step;
      k[u] = i < i1 ? fk[i] : 0;
      v[u] = i < i1 ? amt[i] : 0;
      s[u] = dim_slot(k[u], mask);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) t[u] = b + u * step < i1 ? table[s[u]] : kEmptySlot;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      while (t[u] != kEmptySlot && (int32_t)(t[u] >> 32) != k[u]) {
        s[u] = (s[u] + 1) & mask;
        t[u] = table[s[u]];
      }
      if (t[u] == kEmptySlot) continue;  // no dimension row (or past the end)
      const int g = (int)(uint32_t)t[u];
      lds_add(acc_at(acc, R, rep, g, 0), v[u]);
      if (v[u] != kNull) lds_add(acc_at(acc, R, rep, g, 1), 1);
    }
  }
}

__device__ __forceinline__ void flush_groups(int64_t* acc, int G, int R, int64_t* sum,
                                             int64_t* cnt, int nt) {
  for (int g = threadIdx.x; g < G; g += nt) {
    int64_t s = 0, c = 0;
    for (int r = 0; r < R; ++r) {
      s += *acc_at(acc, R, r, g, 0);
      c += *acc_at(acc, R, r, g, 1);
    }
    if (s) global_add(&sum[g], s);
    if (c) global_add(&cnt[g], c);
  }
}

// SUM(amount), COUNT(amount) by the dimension group of each fact row's key
// (inner join: rows without a dimension match are dropped).
__global__ __launch_bounds__(GB_NT) void k_fact_join_groupby(
    const int32_t* __restrict__ fk, const int64_t* __restrict__ amt, int64_t n,
    const uint64_t* __restrict__ table, uint32_t mask, const int64_t* __restrict__ info, int G,
    int R, int64_t* sum, int64_t* cnt) {
  __shared__ int64_t acc[32 * 1024 / 8];
  for (int x = threadIdx.x; x < R * G * 2; x += GB_NT) acc[x] = 0;
  __syncthreads();
  const int rep = (int)(lane_id() % R);
  const int64_t stride = (int64_t)gridDim.x * GB_NT;
  fact_rows_into(fk, amt, blockIdx.x * (int64_t)GB_NT + threadIdx.x, n, stride, table, mask,
                 info[2], info[3], acc, R, rep);
  __syncthreads();
  flush_groups(acc, G, R, sum, cnt, GB_NT);
}

// One step of the serving engine: up to kServeItems (request, row range)
// items passed by value.  Workgroups [wg_begin[i], wg_begin[i+1]) take item
// i's rows, `rows_per_wg` each, and add their partial sums to the item's
// request accumulators (integer atomics: the sums are exact and do not
// depend on how a request's rows are split over steps and workgroups).
struct ServeArgs {
  int32_t n_items;
  int32_t rows_per_wg;
  pf_serve_item it[PF_SERVE_MAX_ITEMS];
  int32_t wg_begin[PF_SERVE_MAX_ITEMS + 1];
};

// A serving step's workgroup stages a dense dimension of at most SV_DIM key
// values (one byte each) in LDS before its fact rows: its lookups then stay
// in LDS instead of being random byte reads from L2, which are bound by the
// L2's request rate (the step went from 1.3 to 2.0 TB/s of fact rows).
constexpr int SV_NT = 1024;
constexpr int SV_DIM = 65536;

__global__ __launch_bounds__(SV_NT) void k_serve_step(const ServeArgs a) {
  __shared__ int64_t acc[32 * 1024 / 8];
  __shared__ uint32_t dimw[SV_DIM / 4];
  const int wg = blockIdx.x;
  int i = 0;
  while (i + 1 < a.n_items && a.wg_begin[i + 1] <= wg) ++i;
  const pf_serve_item& it = a.it[i];
  const int G = it.G;
  const int R = replicas_for(G);
  for (int x = threadIdx.x; x < R * G * 2; x += SV_NT) acc[x] = 0;
  const int64_t kmin = it.info[2], range = it.info[3];
  const bool stage = range > 0 && range <= SV_DIM;
  if (stage) {
    const uint32_t* src = (const uint32_t*)it.table;
    for (int x = threadIdx.x; x < (int)((range + 3) / 4); x += SV_NT) dimw[x] = src[x];
  }
  __syncthreads();
  const int64_t r0 = (int64_t)(wg - a.wg_begin[i]) * a.rows_per_wg;
  const int64_t r1 = r0 + a.rows_per_wg < it.n ? r0 + a.rows_per_wg : it.n;
  const int rep = (int)(lane_id() % R);
  if (stage) {
    const uint8_t* arr = (const uint8_t*)dimw;
    constexpr int U = 8;
    for (int64_t b = r0 + threadIdx.x; b < r1; b += (int64_t)SV_NT * U) {
      int32_t k[U];
      int64_t v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int64_t ii = b + (int64_t)u * SV_NT;
        k[u] = ii < r1 ? it.fk[ii] : 0;
        v[u] = ii < r1 ? it.amt[ii] : 0;
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const uint64_t off = (uint64_t)((int64_t)k[u] - kmin);
        const uint32_t g = b + (int64_t)u * SV_NT < r1 && off < (uint64_t)range ? arr[off] : 0xFFu;
        if (g == 0xFFu) continue;
        lds_add(acc_at(acc, R, rep, (int)g, 0), v[u]);
        if (v[u] != kNull) lds_add(acc_at(acc, R, rep, (int)g, 1), 1);
      }
    }
  } else {
    fact_rows_into(it.fk, it.amt, r0 + threadIdx.x, r1, SV_NT, it.table,
                   (uint32_t)(it.slots - 1), kmin, range, acc, R, rep);
  }
  __syncthreads();
  flush_groups(acc, G, R, it.sum, it.cnt, SV_NT);
}

unsigned groupby_grid(int64_t n) {
  int64_t g = (n + GB_NT * 8 - 1) / (GB_NT * 8);
  if (g < 1) g = 1;
  if (g > 2048) g = 2048;
  return (unsigned)g;
}

// ------------------------------------------------------ keyed group-by --
//
// Rows grouped by an exact tuple of one or two int64 key columns: an
// open-addressing table of one slot a group (2n slots at most, CAS
// inserts), every row's slot kept; the occupied slots numbered densely in
// slot order; SUM/COUNT by those ids; then the groups ordered by the
// reference's linear grouping hash (sum of key x coefficient, wrapping
// int64) ascending, by stable LSD radix passes over its bits.

constexpr int GK_NT = 256;
constexpr int GK_IT = 16;
constexpr int GK_TILE = GK_NT * GK_IT;
constexpr int GK_W = GK_NT / kWave;
static_assert(GK_IT * GK_W == kWave, "one wave scans a tile's (round, wave) counts");
constexpr uint64_t kSignBit = 1ULL << 63;

// 1 key: the slot holds the key (kEmptyKey: empty); the key int64 min
// itself is slot `cap`, one past the table.
__global__ void k_gk_insert1(const int64_t* __restrict__ k0, int64_t n, int64_t* tkey,
                             uint32_t cap, int32_t* __restrict__ rowslot, int32_t* special) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t k = k0[i];
  if (k == kEmptyKey) {
    rowslot[i] = (int32_t)cap;
    atomicOr(special, 1);
    return;
  }
  uint32_t s = slot_of(mix64((uint64_t)k), cap);
  for (;;) {
    int64_t prev = tkey[s];
    if (prev == kEmptyKey)
      prev = (int64_t)atomicCAS((unsigned long long*)&tkey[s], (unsigned long long)kEmptyKey,
                                (unsigned long long)k);
    if (prev == kEmptyKey || prev == k) {
      rowslot[i] = (int32_t)s;
      return;
    }
    s = next_slot(s, cap);
  }
}

// 2 keys: the slot holds the row that claimed it (-1: empty); tuples are
// compared through the key columns.
__global__ void k_gk_insert2(const int64_t* __restrict__ k0, const int64_t* __restrict__ k1,
                             int64_t n, int32_t* trow, uint32_t cap,
                             int32_t* __restrict__ rowslot) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t a = k0[i], b = k1[i];
  uint32_t s = slot_of(tuple_hash(a, b, true), cap);
  for (;;) {
    int32_t r = trow[s];
    if (r == -1) r = atomicCAS(&trow[s], -1, (int32_t)i);
    if (r == -1 || (k0[r] == a && k1[r] == b)) {
      rowslot[i] = (int32_t)s;
      return;
    }
    s = next_slot(s, cap);
  }
}

template <bool TWO>
__device__ __forceinline__ bool gk_occupied(int64_t s, uint32_t cap, const int64_t* tkey,
                                            const int32_t* trow, const int32_t* special) {
  if (s > (int64_t)cap) return false;
  if (TWO) return s < (int64_t)cap && trow[s] != -1;
  return s == (int64_t)cap ? *special != 0 : tkey[s] != kEmptyKey;
}

// occupied slots a tile of GK_TILE slots (slots 0..cap, the extra one too)
template <bool TWO>
__global__ __launch_bounds__(GK_NT) void k_gk_count(const int64_t* __restrict__ tkey,
                                                    const int32_t* __restrict__ trow,
                                                    const int32_t* __restrict__ special,
                                                    uint32_t cap, int32_t* __restrict__ tilecnt) {
  __shared__ int red[GK_W];
  const int64_t base = (int64_t)blockIdx.x * GK_TILE;
  int c = 0;
#pragma unroll
  for (int j = 0; j < GK_IT; ++j)
    c += gk_occupied<TWO>(base + j * GK_NT + threadIdx.x, cap, tkey, trow, special);
  c = wave_sum(c);
  if (lane_id() == 0) red[threadIdx.x / kWave] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < GK_W; ++w) t += red[w];
    tilecnt[blockIdx.x] = t;
  }
}

// The occupied slots numbered in slot order: each group's dense id, its key
// tuple, and its sort key (the linear hash with the sign bit flipped, so
// that unsigned order is signed order) beside its id.
template <bool TWO>
__global__ __launch_bounds__(GK_NT) void k_gk_number(
    const int64_t* __restrict__ tkey, const int32_t* __restrict__ trow,
    const int32_t* __restrict__ special, uint32_t cap, const int64_t* __restrict__ k0,
    const int64_t* __restrict__ k1, int64_t c0, int64_t c1, const int64_t* __restrict__ offs,
    int32_t* __restrict__ slot_gid, int64_t* __restrict__ rk0, int64_t* __restrict__ rk1,
    int64_t* __restrict__ sort_key, int32_t* __restrict__ sort_val) {
  __shared__ int cnt[kWave];
  const int64_t base = (int64_t)blockIdx.x * GK_TILE;
  const int lane = lane_id(), w = threadIdx.x / kWave;
  bool occ[GK_IT];
  int rank[GK_IT];
#pragma unroll
  for (int j = 0; j < GK_IT; ++j) {
    occ[j] = gk_occupied<TWO>(base + j * GK_NT + threadIdx.x, cap, tkey, trow, special);
    const uint64_t bal = ballot(occ[j]);
    rank[j] = mbcnt(bal);
    if (lane == 0) cnt[j * GK_W + w] = __popcll(bal);
  }
  __syncthreads();
  if (w == 0) {
    const int v = cnt[lane];
    cnt[lane] = wave_incl_scan(v) - v;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < GK_IT; ++j) {
    if (!occ[j]) continue;
    const int64_t s = base + j * GK_NT + threadIdx.x;
    const int64_t g = offs[blockIdx.x] + cnt[j * GK_W + w] + rank[j];
    int64_t a, b = 0;
    if (TWO) {
      const int32_t r = trow[s];
      a = k0[r];
      b = k1[r];
      rk1[g] = b;
    } else {
      a = s == (int64_t)cap ? kEmptyKey : tkey[s];
    }
    rk0[g] = a;
    slot_gid[s] = (int32_t)g;
    const uint64_t h = (uint64_t)a * (uint64_t)c0 + (TWO ? (uint64_t)b * (uint64_t)c1 : 0ULL);
    sort_key[g] = (int64_t)(h ^ kSignBit);
    sort_val[g] = (int32_t)g;
  }
}

__global__ void k_gk_row_gid(const int32_t* __restrict__ rowslot, int64_t n,
                             const int32_t* __restrict__ slot_gid, int32_t* __restrict__ gid) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride)
    gid[i] = slot_gid[rowslot[i]];
}

// The groups in hash order: hash, keys, sums, counts and buckets at their
// sorted positions; bucket counts and repeated hashes counted.
__global__ void k_gk_emit(const int64_t* __restrict__ sort_key,
                          const int32_t* __restrict__ sort_val, int64_t G, bool two,
                          const int64_t* __restrict__ rk0, const int64_t* __restrict__ rk1,
                          const int64_t* __restrict__ sum_s, const int64_t* __restrict__ cnt_s,
                          int logB, int64_t* __restrict__ hash_out, int64_t* __restrict__ key0_out,
                          int64_t* __restrict__ key1_out, int64_t* __restrict__ sum_out,
                          int64_t* __restrict__ cnt_out, int32_t* __restrict__ bucket_out,
                          int64_t* bcount, int64_t* repeats) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < G; j += stride) {
    const int32_t g = sort_val[j];
    const uint64_t h = (uint64_t)sort_key[j] ^ kSignBit;
    hash_out[j] = (int64_t)h;
    key0_out[j] = rk0[g];
    if (two) key1_out[j] = rk1[g];
    if (sum_out) sum_out[j] = sum_s[g];
    if (cnt_out) cnt_out[j] = cnt_s[g];
    const int b = logB ? (int)(mix64(h) >> (64 - logB)) : 0;
    if (bucket_out) bucket_out[j] = b;
    atomicAdd((unsigned long long*)&bcount[b], 1ULL);
    if (j > 0 && sort_key[j - 1] == sort_key[j]) atomicAdd((unsigned long long*)repeats, 1ULL);
  }
}

__global__ void k_gk_readback(const int64_t* __restrict__ repeats,
                              const int64_t* __restrict__ bcount, int B, int64_t* host) {
  for (int b = threadIdx.x; b < B + 1; b += blockDim.x)
    store_host(&host[1 + b], b == 0 ? *repeats : bcount[b - 1]);
}

__global__ void k_gk_read_groups(const int64_t* __restrict__ count, int64_t* host) {
  if (threadIdx.x == 0) store_host(&host[0], *count);
}

struct GkWs {
  int64_t* tkey;
  int32_t* trow;
  int32_t* special;
  int32_t* rowslot;
  int32_t* tilecnt;
  int64_t* offs;
  int64_t* count;
  void* scan;
  int32_t* slot_gid;
  int32_t* gid;
  int64_t *rk0, *rk1, *sum_s, *cnt_s;
  int64_t *sk[2];
  int32_t *sv[2];
  int64_t* sort_off;
  void* rp_ws;
  size_t rp_bytes;
  int64_t *bcount, *repeats;
  uint32_t cap;
  int64_t tiles;
  size_t bytes;
};

GkWs carve_gk(void* ws, int64_t n, bool two, int B) {
  Carve cv(ws);
  GkWs w;
  uint32_t cap = 16;
  while ((int64_t)cap < 2 * n) cap <<= 1;
  w.cap = cap;
  const int64_t m = n > 0 ? n : 1;
  w.tiles = ((int64_t)cap + 1 + GK_TILE - 1) / GK_TILE;
  w.tkey = cv.take<int64_t>(two ? 1 : cap);
  w.trow = cv.take<int32_t>(two ? cap : 1);
  w.special = cv.take<int32_t>(1);
  w.rowslot = cv.take<int32_t>(m);
  w.tilecnt = cv.take<int32_t>(w.tiles);
  w.offs = cv.take<int64_t>(w.tiles);
  w.count = cv.take<int64_t>(2);
  w.scan = cv.take<char>(scan_i32_scratch(w.tiles));
  w.slot_gid = cv.take<int32_t>((int64_t)cap + 1);
  w.gid = cv.take<int32_t>(m);
  w.rk0 = cv.take<int64_t>(m);
  w.rk1 = cv.take<int64_t>(two ? m : 1);
  w.sum_s = cv.take<int64_t>(m);
  w.cnt_s = cv.take<int64_t>(m);
  for (int c = 0; c < 2; ++c) {
    w.sk[c] = cv.take<int64_t>(m);
    w.sv[c] = cv.take<int32_t>(m);
  }
  w.sort_off = cv.take<int64_t>(257);
  w.rp_bytes = radix_partition_ws(m, 8);
  w.rp_ws = cv.take<char>(w.rp_bytes);
  w.bcount = cv.take<int64_t>(B > 0 ? B : 1);
  w.repeats = cv.take<int64_t>(1);
  w.bytes = cv.bytes();
  return w;
}

}  // namespace
}  // namespace pf

extern "C" {

int pf_group_sum_count(const int32_t* gid, const int64_t* val, int64_t n, int32_t G,
                       int64_t* sum, int64_t* cnt, int32_t accumulate, int64_t* bad,
                       void* stream) {
  using namespace pf;
  hipStream_t st = (hipStream_t)stream;
  if (n < 0 || G < 1) return fail("group_sum_count: need n >= 0 and G >= 1");
  if (!accumulate) {
    if (fill_i64(sum, G, 0, st) || fill_i64(cnt, G, 0, st)) return 1;
  }
  if (n == 0) return ok();
  const bool lds = G <= GB_LDS_GROUPS;
  const int R = lds ? replicas_for(G) : 1;
  {
    KTimer _kt("group_sum_count", st);
    if (lds)
      k_group_sum_count<true><<<groupby_grid(n), GB_NT, 0, st>>>(gid, val, n, G, R, sum, cnt,
                                                                 bad);
    else
      k_group_sum_count<false><<<groupby_grid(n), GB_NT, 0, st>>>(gid, val, n, G, R, sum, cnt,
                                                                  bad);
  }
  if (launch_check("group_sum_count")) return 1;
  return ok();
}

size_t pf_group_by_keys_workspace_bytes(int64_t n, int32_t nkeys, int32_t B) {
  return pf::carve_gk(nullptr, n, nkeys == 2, B).bytes;
}

int pf_group_by_keys(const int64_t* k0, const int64_t* k1, int32_t nkeys, int64_t n,
                     const int64_t* coeffs, const int64_t* val, int32_t B, int64_t* hash_out,
                     int64_t* key0_out, int64_t* key1_out, int64_t* sum_out, int64_t* cnt_out,
                     int32_t* bucket_out, int64_t* host_out, void* workspace,
                     size_t workspace_bytes, void* stream) {
  using namespace pf;
  hipStream_t st = (hipStream_t)stream;
  if (nkeys != 1 && nkeys != 2) return fail("group_by_keys: one or two key columns");
  if (n < 0 || n >= (1LL << 30)) return fail("group_by_keys: n must be in [0, 2^30)");
  if (B < 1 || B > 4096 || (B & (B - 1)))
    return fail("group_by_keys: B must be a power of two in [1, 4096]");
  if (!coeffs) return fail("group_by_keys: the hash coefficients are required");
  if (val && (!sum_out || !cnt_out)) return fail("group_by_keys: a value column needs sum and count outputs");
  const bool two = nkeys == 2;
  GkWs w = carve_gk(workspace, n, two, B);
  if (workspace_bytes < w.bytes) return fail("group_by_keys: workspace too small");
  void* hdev = nullptr;
  if (host_alias(host_out, &hdev)) return 1;
  int logB = 0;
  while ((1 << logB) < B) ++logB;
  const uint32_t cap = w.cap;
  if (n == 0) {
    for (int b = 0; b < 2 + B; ++b) host_out[b] = 0;
    return ok();
  }
  // 1. every row's slot
  if (two) {
    if (fill_i32(w.trow, cap, -1, st)) return 1;
    k_gk_insert2<<<grid_for(n, 256), 256, 0, st>>>(k0, k1, n, w.trow, cap, w.rowslot);
  } else {
    if (fill_i64(w.tkey, cap, kEmptyKey, st) || fill_i32(w.special, 1, 0, st)) return 1;
    k_gk_insert1<<<grid_for(n, 256), 256, 0, st>>>(k0, n, w.tkey, cap, w.rowslot, w.special);
  }
  // 2. the groups numbered in slot order, with their keys and sort keys
  if (two)
    k_gk_count<true><<<(unsigned)w.tiles, GK_NT, 0, st>>>(w.tkey, w.trow, w.special, cap,
                                                          w.tilecnt);
  else
    k_gk_count<false><<<(unsigned)w.tiles, GK_NT, 0, st>>>(w.tkey, w.trow, w.special, cap,
                                                           w.tilecnt);
  if (launch_check("group_by_keys")) return 1;
  if (scan_i32_to_i64_excl(w.tilecnt, w.offs, w.tiles, w.count, w.scan, st)) return 1;
  const int64_t c1 = two ? coeffs[1] : 0;
  if (two)
    k_gk_number<true><<<(unsigned)w.tiles, GK_NT, 0, st>>>(
        w.tkey, w.trow, w.special, cap, k0, k1, coeffs[0], c1, w.offs, w.slot_gid, w.rk0,
        w.rk1, w.sk[0], w.sv[0]);
  else
    k_gk_number<false><<<(unsigned)w.tiles, GK_NT, 0, st>>>(
        w.tkey, w.trow, w.special, cap, k0, k0, coeffs[0], 0, w.offs, w.slot_gid, w.rk0,
        w.rk1, w.sk[0], w.sv[0]);
  k_gk_read_groups<<<1, 64, 0, st>>>(w.count, (int64_t*)hdev);
  if (launch_check("group_by_keys")) return 1;
  if (wait_stream(st, "group_by_keys")) return 1;
  const int64_t G = host_out[0];
  if (G < 1 || G > n) return fail("group_by_keys: counted " + std::to_string(G) + " groups");
  // 3. SUM / COUNT by dense id
  if (val) {
    k_gk_row_gid<<<groupby_grid(n), GB_NT, 0, st>>>(w.rowslot, n, w.slot_gid, w.gid);
    if (fill_i64(w.sum_s, G, 0, st) || fill_i64(w.cnt_s, G, 0, st)) return 1;
    const bool lds = G <= GB_LDS_GROUPS;
    const int R = lds ? replicas_for((int)G) : 1;
    KTimer _kt("group_by_keys_sum", st);
    if (lds)
      k_group_sum_count<true><<<groupby_grid(n), GB_NT, 0, st>>>(w.gid, val, n, (int)G, R,
                                                                 w.sum_s, w.cnt_s, nullptr);
    else
      k_group_sum_count<false><<<groupby_grid(n), GB_NT, 0, st>>>(w.gid, val, n, (int)G, R,
                                                                  w.sum_s, w.cnt_s, nullptr);
  }
  // 4. the groups in hash order: 8 stable passes of 8 bits, ending in sk[0]
  if (G > 1) {
    for (int pass = 0; pass < 8; ++pass) {
      const int c = pass & 1;
      if (radix_partition(w.sk[c], 8, w.sv[c], G, 8, pass * 8, 2, w.sk[c ^ 1], w.sv[c ^ 1],
                          w.sort_off, w.rp_ws, w.rp_bytes, st, 5, -1))
        return 1;
    }
  }
  if (fill_i64(w.bcount, B, 0, st) || fill_i64(w.repeats, 1, 0, st)) return 1;
  k_gk_emit<<<grid_for(G, 256) < 4096 ? grid_for(G, 256) : 4096, 256, 0, st>>>(
      w.sk[0], w.sv[0], G, two, w.rk0, w.rk1, w.sum_s, w.cnt_s, logB, hash_out, key0_out,
      key1_out, val ? sum_out : nullptr, val ? cnt_out : nullptr, bucket_out, w.bcount,
      w.repeats);
  k_gk_readback<<<1, 256, 0, st>>>(w.repeats, w.bcount, B, (int64_t*)hdev);
  if (launch_check("group_by_keys")) return 1;
  if (wait_stream(st, "group_by_keys")) return 1;
  return ok();
}

int64_t pf_dim_table_slots(int64_t n_dim) {
  int64_t cap = 16;
  while (cap < 2 * n_dim) cap <<= 1;
  return cap;
}

int pf_dim_table_build(const int32_t* key, const int32_t* grp, int64_t n, int32_t G,
                       uint64_t* table, int64_t slots, int64_t* info, void* stream) {
  using namespace pf;
  hipStream_t st = (hipStream_t)stream;
  if (n < 0 || G < 1 || G > GB_LDS_GROUPS)
    return fail("dim_table_build: need n >= 0 and 1 <= G <= 1024");
  if (slots < 2 * n || (slots & (slots - 1)) || slots > (1LL << 32))
    return fail("dim_table_build: slots must be a power of two >= 2n");
  if (fill_i64((int64_t*)table, slots, -1, st)) return 1;  // empty slots, absent bytes
  KTimer _kt("dim_table_build", st);
  k_dim_init<<<1, 64, 0, st>>>(info);
  if (n > 0) {
    const unsigned g = (unsigned)(grid_for(n, 256) < 1024 ? grid_for(n, 256) : 1024);
    k_dim_minmax<<<g, 256, 0, st>>>(key, n, info);
  }
  k_dim_layout<<<1, 1, 0, st>>>(n, slots, G, info);
  if (n > 0)
    k_dim_build<<<grid_for(n, 256), 256, 0, st>>>(key, grp, n, table, (uint32_t)(slots - 1), G,
                                                  info);
  if (launch_check("dim_table_build")) return 1;
  return ok();
}

int pf_fact_join_groupby(const int32_t* fk, const int64_t* amt, int64_t n,
                         const uint64_t* table, int64_t slots, const int64_t* info, int32_t G,
                         int64_t* sum, int64_t* cnt, void* stream) {
  using namespace pf;
  hipStream_t st = (hipStream_t)stream;
  if (n < 0 || G < 1 || G > GB_LDS_GROUPS)
    return fail("fact_join_groupby: need n >= 0 and 1 <= G <= 1024");
  if (slots < 16 || (slots & (slots - 1))) return fail("fact_join_groupby: bad table size");
  if (fill_i64(sum, G, 0, st) || fill_i64(cnt, G, 0, st)) return 1;
  if (n > 0) {
    KTimer _kt("fact_join_groupby", st);
    k_fact_join_groupby<<<groupby_grid(n), GB_NT, 0, st>>>(fk, amt, n, table,
                                                           (uint32_t)(slots - 1), info, G,
                                                           replicas_for(G), sum, cnt);
  }
  if (launch_check("fact_join_groupby")) return 1;
  return ok();
}

int pf_serve_step(const pf_serve_item* items, int32_t n_items, int32_t rows_per_wg,
                  void* stream) {
  using namespace pf;
  hipStream_t st = (hipStream_t)stream;
  if (n_items < 0 || n_items > PF_SERVE_MAX_ITEMS)
    return fail("serve_step: 0 to PF_SERVE_MAX_ITEMS items");
  if (rows_per_wg < GB_NT || rows_per_wg > (1 << 24))
    return fail("serve_step: rows_per_wg in [512, 2^24]");
  ServeArgs a{};
  a.n_items = n_items;
  a.rows_per_wg = rows_per_wg;
  int64_t wgs = 0;
  for (int i = 0; i < n_items; ++i) {
    const pf_serve_item& it = items[i];
    if (it.n < 0 || it.G < 1 || it.G > GB_LDS_GROUPS || it.slots < 16 ||
        (it.slots & (it.slots - 1)) || it.slots > (1LL << 32))
      return fail("serve_step: item " + std::to_string(i) +
                  ": need n >= 0, 1 <= G <= 1024 and a power-of-two table");
    if (it.n > 0 && (!it.fk || !it.amt || !it.table || !it.info || !it.sum || !it.cnt))
      return fail("serve_step: item " + std::to_string(i) + " has a null pointer");
    a.it[i] = it;
    a.wg_begin[i] = (int32_t)wgs;
    wgs += (it.n + rows_per_wg - 1) / rows_per_wg;
    if (wgs > (1LL << 30)) return fail("serve_step: too many workgroups");
  }
  a.wg_begin[n_items] = (int32_t)wgs;
  if (wgs == 0) return ok();
  {
    KTimer _kt("serve_step", st);
    k_serve_step<<<(unsigned)wgs, SV_NT, 0, st>>>(a);
  }
  if (launch_check("serve_step")) return 1;
  return ok();
}

}  // extern "C"
