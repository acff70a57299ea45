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
// table (R replicas, lane % R picks one, so lanes of a wave that share a
// group hit different banks and addresses), reduces the replicas and adds
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
  __shared__ int64_t acc[LDS ? 32 * 1024 / 8 : 1];  // [R][G][2]
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
      int64_t* a = acc + ((int64_t)rep * G + g) * 2;
      lds_add(a, v);
      if (v != kNull) lds_add(a + 1, 1);
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
        s += acc[((int64_t)r * G + g) * 2];
        c += acc[((int64_t)r * G + g) * 2 + 1];
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
                                               int64_t kmin, int64_t range, int64_t* acc_rep) {
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
        int64_t* a = acc_rep + (int64_t)g * 2;
        lds_add(a, v[u]);
        if (v[u] != kNull) lds_add(a + 1, 1);
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
      int64_t* a = acc_rep + (int64_t)(uint32_t)t[u] * 2;
      lds_add(a, v[u]);
      if (v[u] != kNull) lds_add(a + 1, 1);
    }
  }
}

__device__ __forceinline__ void flush_groups(const int64_t* acc, int G, int R, int64_t* sum,
                                             int64_t* cnt, int nt) {
  for (int g = threadIdx.x; g < G; g += nt) {
    int64_t s = 0, c = 0;
    for (int r = 0; r < R; ++r) {
      s += acc[((int64_t)r * G + g) * 2];
      c += acc[((int64_t)r * G + g) * 2 + 1];
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
                 info[2], info[3], acc + (int64_t)rep * G * 2);
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
  int64_t* accr = acc + (int64_t)rep * G * 2;
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
        int64_t* ap = accr + (int64_t)g * 2;
        lds_add(ap, v[u]);
        if (v[u] != kNull) lds_add(ap + 1, 1);
      }
    }
  } else {
    fact_rows_into(it.fk, it.amt, r0 + threadIdx.x, r1, SV_NT, it.table,
                   (uint32_t)(it.slots - 1), kmin, range, accr);
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
