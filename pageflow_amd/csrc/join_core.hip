// The join core's per-call kernels: drop-ins for the reference's
// join_unique_hash, join_semi_anti and join_expand entry points
// (operators/kernels/join_core.cu:263-400; Python contract
// operators/join_core.py:54-143).
//
//   unique hash  inner join on a build tuple (1-2 int64 keys) proven unique:
//                an open-addressing table at load <= 0.5, one partner per
//                probe row (the lowest build row with the tuple), the matched
//                probe rows compacted stably (probe-row order)
//   semi / anti  a hash set of the build keys with an optional per-key
//                min/max of a build value, the probe rows with (semi) or
//                without (anti) a passing match compacted stably
//   expand       each probe row's run of a sorted build written as (build
//                row, probe row) pairs in probe-row order
//
// gfx950 design: 1-key tables keep the key in the slot (one 8-B probe load
// per step, no dependent load of the build key), slots come from the low 32
// bits of mix64 by multiply-shift, inserts read before their 64-bit CAS.
// Compaction is two passes over 4096-row tiles (predicate + tile counts,
// device scan, ballot-ranked writes), so the order is deterministic; no CUB.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <type_traits>

#include "pf_common.hpp"
#include "pf_residual.hpp"
#include "../../include/pageflow_join.h"

namespace pf {
namespace {

constexpr int CT_NT = 256;          // threads a compaction tile
constexpr int CT_IT = 16;           // rounds
constexpr int CT_TILE = CT_NT * CT_IT;
constexpr int CT_W = CT_NT / kWave;
constexpr int64_t MAX_ROWS = (1LL << 31) - (1LL << 16);  // operators/join_core.py:30

enum Op : int { OP_NONE = 0, OP_NE = 1, OP_LT = 2, OP_LE = 3, OP_GT = 4, OP_GE = 5 };

// Block-wide sum (every thread gets it).
__device__ __forceinline__ int block_sum_ct(int v, int* red) {
  v = wave_sum(v);
  if (lane_id() == 0) red[threadIdx.x / kWave] = v;
  __syncthreads();
  int t = 0;
#pragma unroll
  for (int w = 0; w < CT_W; ++w) t += red[w];
  __syncthreads();
  return t;
}

// -------------------------------------------------------- unique hash --

// 1 key: the slot holds the key (empty = int64 min) and the lowest build
// row with it.  The key int64 min itself cannot be stored; its lowest row
// goes to *special.
__global__ void k_uh_build1(const int64_t* __restrict__ b0, int64_t nb, int64_t* hkey,
                            int32_t* hrow, uint32_t cap, int32_t* special) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nb) return;
  const int64_t k = b0[i];
  if (k == kEmptyKey) {
    atomicMin(special, (int32_t)i);
    return;
  }
  uint32_t s = slot_of(mix64((uint64_t)k), cap);
  for (;;) {
    int64_t prev = hkey[s];
    if (prev == kEmptyKey)
      prev = (int64_t)atomicCAS((unsigned long long*)&hkey[s], (unsigned long long)kEmptyKey,
                                (unsigned long long)k);
    if (prev == kEmptyKey || prev == k) {
      atomicMin(&hrow[s], (int32_t)i);
      return;
    }
    s = next_slot(s, cap);
  }
}

// 2 keys: the slot holds a build row (-1 empty); the tuple is compared
// through the build columns.  A repeated tuple keeps its lowest row.
__global__ void k_uh_build2(const int64_t* __restrict__ b0, const int64_t* __restrict__ b1,
                            int64_t nb, int32_t* hrow, uint32_t cap) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nb) return;
  const int64_t k0 = b0[i], k1 = b1[i];
  uint32_t s = slot_of(tuple_hash(k0, k1, true), cap);
  for (;;) {
    int32_t r = hrow[s];
    if (r == -1) r = atomicCAS(&hrow[s], -1, (int32_t)i);
    if (r == -1) return;
    if (b0[r] == k0 && b1[r] == k1) {
      atomicMin(&hrow[s], (int32_t)i);
      return;
    }
    s = next_slot(s, cap);
  }
}

// partner[i]: probe row i's build row or -1; tilecnt[t]: tile t's matches.
template <bool TWO>
__global__ __launch_bounds__(CT_NT) void k_uh_probe(
    const int64_t* __restrict__ b0, const int64_t* __restrict__ b1,
    const int64_t* __restrict__ p0, const int64_t* __restrict__ p1, int64_t np,
    const int64_t* __restrict__ hkey, const int32_t* __restrict__ hrow, uint32_t cap,
    const int32_t* __restrict__ special, int32_t* __restrict__ partner,
    int32_t* __restrict__ tilecnt) {
  __shared__ int red[CT_W];
  const int64_t base = (int64_t)blockIdx.x * CT_TILE;
  const int32_t sp = TWO ? -1 : *special;
  int c = 0;
#pragma unroll 4
  for (int j = 0; j < CT_IT; ++j) {
    const int64_t i = base + j * CT_NT + threadIdx.x;
    if (i >= np) break;
    const int64_t k0 = p0[i];
    int32_t r = -1;
    if (TWO) {
      const int64_t k1 = p1[i];
      uint32_t s = slot_of(tuple_hash(k0, k1, true), cap);
      for (;;) {
        const int32_t b = hrow[s];
        if (b == -1) break;
        if (b0[b] == k0 && b1[b] == k1) {
          r = b;
          break;
        }
        s = next_slot(s, cap);
      }
    } else if (k0 == kEmptyKey) {
      r = sp == INT32_MAX ? -1 : sp;
    } else {
      uint32_t s = slot_of(mix64((uint64_t)k0), cap);
      for (;;) {
        const int64_t t = hkey[s];
        if (t == k0) {
          r = hrow[s];
          break;
        }
        if (t == kEmptyKey) break;
        s = next_slot(s, cap);
      }
    }
    partner[i] = r;
    c += r >= 0;
  }
  c = block_sum_ct(c, red);
  if (threadIdx.x == 0) tilecnt[blockIdx.x] = c;
}

// --------------------------------------------------- stable compaction --

// Positions of a tile's kept rows (row order): rounds j-major, lanes in
// order.  keep[j]: row (base + j*CT_NT + tid) is kept.  Returns, per round,
// the output index of the row (only meaningful where kept).
__device__ __forceinline__ void tile_positions(const bool (&keep)[CT_IT], int64_t off,
                                               int64_t (&pos)[CT_IT], int* cnt) {
  const int lane = lane_id(), w = threadIdx.x / kWave;
  int rank[CT_IT];
#pragma unroll
  for (int j = 0; j < CT_IT; ++j) {
    const uint64_t bal = ballot(keep[j]);
    rank[j] = mbcnt(bal);
    if (lane == 0) cnt[j * CT_W + w] = __popcll(bal);
  }
  __syncthreads();
  if (w == 0) {  // CT_IT * CT_W = 64 counts, (round, wave) order
    const int v = cnt[lane];
    const int incl = wave_incl_scan(v);
    cnt[lane] = incl - v;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < CT_IT; ++j) pos[j] = off + cnt[j * CT_W + w] + rank[j];
}
static_assert(CT_IT * CT_W == kWave, "one wave scans a tile's (round, wave) counts");

// probe_idx / build_idx of the matched probe rows at their scanned tile
// offsets.
__global__ __launch_bounds__(CT_NT) void k_uh_compact(const int32_t* __restrict__ partner,
                                                      int64_t np,
                                                      const int64_t* __restrict__ offs,
                                                      int64_t* __restrict__ probe_idx,
                                                      int64_t* __restrict__ build_idx) {
  __shared__ int cnt[kWave];
  const int64_t base = (int64_t)blockIdx.x * CT_TILE;
  bool keep[CT_IT];
  int32_t r[CT_IT];
#pragma unroll
  for (int j = 0; j < CT_IT; ++j) {
    const int64_t i = base + j * CT_NT + threadIdx.x;
    r[j] = i < np ? partner[i] : -1;
    keep[j] = r[j] >= 0;
  }
  int64_t pos[CT_IT];
  tile_positions(keep, offs[blockIdx.x], pos, cnt);
#pragma unroll
  for (int j = 0; j < CT_IT; ++j) {
    if (keep[j]) {
      probe_idx[pos[j]] = base + j * CT_NT + threadIdx.x;
      build_idx[pos[j]] = r[j];
    }
  }
}

// kept probe rows of a uint8 keep array at their scanned tile offsets.
__global__ __launch_bounds__(CT_NT) void k_rows_compact(const uint8_t* __restrict__ keepv,
                                                        int64_t n,
                                                        const int64_t* __restrict__ offs,
                                                        int64_t* __restrict__ out) {
  __shared__ int cnt[kWave];
  const int64_t base = (int64_t)blockIdx.x * CT_TILE;
  bool keep[CT_IT];
#pragma unroll
  for (int j = 0; j < CT_IT; ++j) {
    const int64_t i = base + j * CT_NT + threadIdx.x;
    keep[j] = i < n && keepv[i];
  }
  int64_t pos[CT_IT];
  tile_positions(keep, offs[blockIdx.x], pos, cnt);
#pragma unroll
  for (int j = 0; j < CT_IT; ++j)
    if (keep[j]) out[pos[j]] = base + j * CT_NT + threadIdx.x;
}

// rows whose flag equals `want`: per-tile counts, then compaction.
__global__ __launch_bounds__(CT_NT) void k_flag_count(const uint8_t* __restrict__ flags,
                                                      int64_t n, bool want,
                                                      int32_t* __restrict__ tilecnt) {
  __shared__ int red[CT_W];
  const int64_t base = (int64_t)blockIdx.x * CT_TILE;
  int c = 0;
#pragma unroll
  for (int j = 0; j < CT_IT; ++j) {
    const int64_t i = base + j * CT_NT + threadIdx.x;
    c += (i < n && ((flags[i] != 0) == want)) ? 1 : 0;
  }
  c = block_sum_ct(c, red);
  if (threadIdx.x == 0) tilecnt[blockIdx.x] = c;
}

__global__ __launch_bounds__(CT_NT) void k_flag_compact(const uint8_t* __restrict__ flags,
                                                        int64_t n, bool want,
                                                        const int64_t* __restrict__ offs,
                                                        int64_t* __restrict__ out) {
  __shared__ int cnt[kWave];
  const int64_t base = (int64_t)blockIdx.x * CT_TILE;
  bool keep[CT_IT];
#pragma unroll
  for (int j = 0; j < CT_IT; ++j) {
    const int64_t i = base + j * CT_NT + threadIdx.x;
    keep[j] = i < n && ((flags[i] != 0) == want);
  }
  int64_t pos[CT_IT];
  tile_positions(keep, offs[blockIdx.x], pos, cnt);
#pragma unroll
  for (int j = 0; j < CT_IT; ++j)
    if (keep[j]) out[pos[j]] = base + j * CT_NT + threadIdx.x;
}

__global__ void k_readback(const int64_t* __restrict__ dev, int n, int64_t* host) {
  if ((int)threadIdx.x < n) store_host(&host[threadIdx.x], dev[threadIdx.x]);
}

// --------------------------------------------------------- semi / anti --

// Insert the key (read before the CAS, so repeats neither CAS nor contend)
// and fold the build value into the slot's min and/or max.  A key equal to
// the empty-slot value sets *bad (the caller falls back).
__global__ void k_set_build(const int64_t* __restrict__ key, const int64_t* __restrict__ val,
                            int64_t n, int64_t* skey, int64_t* mn, int64_t* mx, uint32_t cap,
                            int64_t* bad) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t k = key[i];
  if (k == kEmptyKey) {
    *bad = 1;
    return;
  }
  uint32_t s = slot_of(mix64((uint64_t)k), cap);
  for (;;) {
    int64_t prev = skey[s];
    if (prev == kEmptyKey)
      prev = (int64_t)atomicCAS((unsigned long long*)&skey[s], (unsigned long long)kEmptyKey,
                                (unsigned long long)k);
    if (prev == kEmptyKey || prev == k) break;
    s = next_slot(s, cap);
  }
  if (mn || mx) {
    const int64_t v = val[i];
    // a stale read is never past the atomic's result: min only falls
    if (mn && v < mn[s]) atomicMin((long long*)&mn[s], (long long)v);
    if (mx && v > mx[s]) atomicMax((long long*)&mx[s], (long long)v);
  }
}

__device__ __forceinline__ bool set_hit(int64_t k, const int64_t* __restrict__ skey,
                                        const int64_t* __restrict__ mn,
                                        const int64_t* __restrict__ mx, uint32_t cap, int op,
                                        int64_t v) {
  if (k == kEmptyKey) return false;
  uint32_t s = slot_of(mix64((uint64_t)k), cap);
  for (;;) {
    const int64_t t = skey[s];
    if (t == k) break;
    if (t == kEmptyKey) return false;
    s = next_slot(s, cap);
  }
  switch (op) {
    case OP_NONE: return true;
    case OP_NE: return mn[s] != v || mx[s] != v;
    case OP_LT: return mn[s] < v;
    case OP_LE: return mn[s] <= v;
    case OP_GT: return mx[s] > v;
    default: return mx[s] >= v;
  }
}

// keep[i] for probe row i, and tile counts (when tilecnt is not null).
__global__ __launch_bounds__(CT_NT) void k_set_probe(
    const int64_t* __restrict__ pk, const int64_t* __restrict__ pv, int64_t np,
    const int64_t* __restrict__ skey, const int64_t* __restrict__ mn,
    const int64_t* __restrict__ mx, uint32_t cap, int op, int anti,
    uint8_t* __restrict__ keepv, int32_t* __restrict__ tilecnt) {
  __shared__ int red[CT_W];
  const int64_t base = (int64_t)blockIdx.x * CT_TILE;
  int c = 0;
#pragma unroll 4
  for (int j = 0; j < CT_IT; ++j) {
    const int64_t i = base + j * CT_NT + threadIdx.x;
    if (i >= np) break;
    // the probe stream is read and the mask written past L2 (nontemporal),
    // which keeps L2 for the set's lines: a skewed probe's hot keys then
    // stay resident
    const int64_t k = __builtin_nontemporal_load(&pk[i]);
    const int64_t v = op ? __builtin_nontemporal_load(&pv[i]) : 0;
    const bool hit = set_hit(k, skey, mn, mx, cap, op, v);
    const bool keep = hit != (anti != 0);
    __builtin_nontemporal_store((uint8_t)keep, &keepv[i]);
    c += keep;
  }
  if (tilecnt) {
    c = block_sum_ct(c, red);
    if (threadIdx.x == 0) tilecnt[blockIdx.x] = c;
  }
}

// -------------------------------------------------------------- expand --

// Pair k belongs to the last probe row i with ex[i] <= k (that row has a
// non-empty run); its build row is sorted_ids[start[i] + k - ex[i]].
constexpr int EX_NT = 256, EX_PER = 8, EX_TILE = EX_NT * EX_PER;

__device__ __forceinline__ int64_t last_le(const int64_t* __restrict__ ex, int64_t lo,
                                           int64_t hi, int64_t k) {
  // largest i in [lo, hi) with ex[i] <= k (ex[lo] <= k assumed)
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (ex[mid] <= k)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(EX_NT) void k_expand(const int64_t* __restrict__ sorted_ids,
                                                  const int64_t* __restrict__ start,
                                                  const int64_t* __restrict__ ex, int64_t m,
                                                  int64_t n_flat, int64_t* __restrict__ cs,
                                                  int64_t* __restrict__ cr) {
  __shared__ int64_t range[2];
  const int64_t k0 = (int64_t)blockIdx.x * EX_TILE;
  const int64_t k1 = (k0 + EX_TILE < n_flat ? k0 + EX_TILE : n_flat) - 1;
  if (threadIdx.x < 2) range[threadIdx.x] = last_le(ex, 0, m, threadIdx.x ? k1 : k0);
  __syncthreads();
  const int64_t r0 = range[0], r1 = range[1] + 1;
#pragma unroll
  for (int q = 0; q < EX_PER; ++q) {
    const int64_t k = k0 + q * EX_NT + threadIdx.x;
    if (k >= n_flat) break;
    const int64_t i = last_le(ex, r0, r1, k);
    cs[k] = sorted_ids[start[i] + (k - ex[i])];
    cr[k] = i;
  }
}

// ----------------------------------------------------------- layouts --

struct UniqueWs {
  int64_t* hkey;
  int32_t* hrow;
  int32_t* special;
  int32_t* partner;
  int32_t* tilecnt;
  int64_t* offs;
  int64_t* count;
  void* scan;
  size_t bytes;
};

UniqueWs carve_unique(void* ws, int64_t nb, int64_t np, bool two) {
  Carve cv(ws);
  UniqueWs u;
  const int64_t cap = 2 * (nb > 0 ? nb : 1);
  const int64_t tiles = (np + CT_TILE - 1) / CT_TILE;
  u.hkey = cv.take<int64_t>(two ? 1 : cap);
  u.hrow = cv.take<int32_t>(cap);
  u.special = cv.take<int32_t>(1);
  u.partner = cv.take<int32_t>(np > 0 ? np : 1);
  u.tilecnt = cv.take<int32_t>(tiles > 0 ? tiles : 1);
  u.offs = cv.take<int64_t>(tiles > 0 ? tiles : 1);
  u.count = cv.take<int64_t>(2);
  u.scan = cv.take<char>(scan_i32_scratch(tiles));
  u.bytes = cv.bytes();
  return u;
}

struct SetWs {
  int64_t* skey;
  int64_t* mn;
  int64_t* mx;
  uint8_t* keep;
  int32_t* tilecnt;
  int64_t* offs;
  int64_t* dev;  // [0] kept rows, [1] bad
  void* scan;
  size_t bytes;
};

SetWs carve_set(void* ws, int64_t nb, int64_t np, int op, bool compact) {
  Carve cv(ws);
  SetWs s;
  const int64_t cap = 2 * (nb > 0 ? nb : 1);
  const int64_t tiles = (np + CT_TILE - 1) / CT_TILE;
  s.skey = cv.take<int64_t>(cap);
  const bool need_mn = op == OP_NE || op == OP_LT || op == OP_LE;
  const bool need_mx = op == OP_NE || op == OP_GT || op == OP_GE;
  s.mn = cv.take<int64_t>(need_mn ? cap : 1);
  s.mx = cv.take<int64_t>(need_mx ? cap : 1);
  if (!need_mn) s.mn = nullptr;
  if (!need_mx) s.mx = nullptr;
  s.keep = cv.take<uint8_t>(compact ? (np > 0 ? np : 1) : 1);
  s.tilecnt = cv.take<int32_t>(tiles > 0 ? tiles : 1);
  s.offs = cv.take<int64_t>(tiles > 0 ? tiles : 1);
  s.dev = cv.take<int64_t>(2);
  s.scan = cv.take<char>(scan_i32_scratch(tiles));
  s.bytes = cv.bytes();
  return s;
}

// Builds a semi/anti set into s (fills, inserts); shared by the compacted
// and the mask entry points.
int build_set(const SetWs& s, const int64_t* bk, const int64_t* bv, int64_t nb, int op,
              hipStream_t st) {
  const int64_t cap = 2 * nb;
  if (fill_i64(s.skey, cap, kEmptyKey, st)) return 1;
  if (s.mn && fill_i64(s.mn, cap, INT64_MAX, st)) return 1;
  if (s.mx && fill_i64(s.mx, cap, INT64_MIN, st)) return 1;
  if (fill_i64(s.dev, 2, 0, st)) return 1;
  k_set_build<<<grid_for(nb, 256), 256, 0, st>>>(bk, op ? bv : nullptr, nb, s.skey, s.mn,
                                                   s.mx, (uint32_t)cap, s.dev + 1);
  return launch_check("semi_anti build");
}

// ---------------------------------------------------- residual filter --
//
// The residual conditions on a join's pairs in one pass: every thread
// interprets the same postfix program (its words are kernel arguments, read
// through scalar loads, so the dispatch on opcode, relation and column type
// is uniform) for RF_G pairs at a time, the operand gathers of those pairs
// in flight together; a per-pair bit stack holds the partial results.  The
// pass writes one 64-bit ballot word per (round, wave) of a 4096-pair tile
// and the tile's count; after the device scan the compaction pass reads the
// words and writes the kept pairs at their scanned positions, reading only
// the kept pairs' indexes.

constexpr int RF_G = 4;  // pairs a thread evaluates together
static_assert(CT_IT % RF_G == 0, "rounds split into groups");

// Column values at idx: integers sign- (uint8 zero-) extended, float64 as
// its bits; F64 converts integers to float64 bits.  An index outside the
// column reads 0 and counts in bad.
template <typename T, bool F64>
__device__ __forceinline__ void rf_fetch(const void* p, int64_t rows,
                                         const int64_t (&idx)[RF_G], const bool (&live)[RF_G],
                                         int64_t (&v)[RF_G], int& bad) {
  const T* c = (const T*)p;
#pragma unroll
  for (int j = 0; j < RF_G; ++j) {
    const bool in = live[j] && (uint64_t)idx[j] < (uint64_t)rows;
    bad += live[j] && !in;
    const T x = in ? c[idx[j]] : (T)0;
    if constexpr (sizeof(T) == 8 && !std::is_integral<T>::value)
      v[j] = __double_as_longlong(x);
    else if constexpr (F64)
      v[j] = __double_as_longlong((double)x);
    else
      v[j] = (int64_t)x;
  }
}

template <bool F64>
__device__ __forceinline__ void rf_operand(const RfProgram& P, bool is_col, int64_t w,
                                           const int64_t (&li)[RF_G],
                                           const int64_t (&ri)[RF_G],
                                           const bool (&live)[RF_G], int64_t (&v)[RF_G],
                                           int& bad) {
  if (!is_col) {
#pragma unroll
    for (int j = 0; j < RF_G; ++j) v[j] = w;
    return;
  }
  const int s = (int)w;
  const bool right = P.side[s] != 0;
  int64_t idx[RF_G];
#pragma unroll
  for (int j = 0; j < RF_G; ++j) idx[j] = right ? ri[j] : li[j];
  const void* p = P.col[s];
  const int64_t rows = P.rows[s];
  switch (P.type[s]) {
    case PF_COL_I64: rf_fetch<int64_t, F64>(p, rows, idx, live, v, bad); break;
    case PF_COL_I32: rf_fetch<int32_t, F64>(p, rows, idx, live, v, bad); break;
    case PF_COL_I16: rf_fetch<int16_t, F64>(p, rows, idx, live, v, bad); break;
    case PF_COL_I8: rf_fetch<int8_t, F64>(p, rows, idx, live, v, bad); break;
    case PF_COL_U8: rf_fetch<uint8_t, F64>(p, rows, idx, live, v, bad); break;
    default: rf_fetch<double, F64>(p, rows, idx, live, v, bad); break;
  }
}

__global__ __launch_bounds__(CT_NT) void k_rf_eval(const int64_t* __restrict__ left,
                                                   const int64_t* __restrict__ right, int64_t n,
                                                   const RfProgram P,
                                                   uint64_t* __restrict__ words,
                                                   uint8_t* __restrict__ mask_out,
                                                   int32_t* __restrict__ tilecnt,
                                                   int64_t* __restrict__ bad_out) {
  __shared__ int red[CT_W];
  const int64_t base = (int64_t)blockIdx.x * CT_TILE;
  const int w = threadIdx.x / kWave;
  int kept = 0, bad = 0;
#pragma unroll 1
  for (int g = 0; g < CT_IT; g += RF_G) {
    int64_t li[RF_G], ri[RF_G];
    bool live[RF_G];
    uint32_t stk[RF_G];
#pragma unroll
    for (int j = 0; j < RF_G; ++j) {
      const int64_t i = base + (int64_t)(g + j) * CT_NT + threadIdx.x;
      live[j] = i < n;
      li[j] = live[j] ? left[i] : 0;
      ri[j] = live[j] ? right[i] : 0;
      stk[j] = 0;
    }
#pragma unroll 1
    for (int pc = 0; pc < P.n_ins; ++pc) {
      const int64_t w0 = P.ins[pc][0];
      const int opc = (int)(w0 & 0xff);
      if (opc == PF_RES_CMP) {
        const int rel = (int)((w0 >> 8) & 0xff);
        int64_t a[RF_G], b[RF_G];
        if (w0 & PF_RES_F64) {
          rf_operand<true>(P, (w0 & PF_RES_A_COL) != 0, P.ins[pc][1], li, ri, live, a, bad);
          rf_operand<true>(P, (w0 & PF_RES_B_COL) != 0, P.ins[pc][2], li, ri, live, b, bad);
#pragma unroll
          for (int j = 0; j < RF_G; ++j)
            stk[j] = (stk[j] << 1) |
                     (uint32_t)rf_rel_f(rel, __longlong_as_double(a[j]),
                                        __longlong_as_double(b[j]));
        } else {
          rf_operand<false>(P, (w0 & PF_RES_A_COL) != 0, P.ins[pc][1], li, ri, live, a, bad);
          rf_operand<false>(P, (w0 & PF_RES_B_COL) != 0, P.ins[pc][2], li, ri, live, b, bad);
#pragma unroll
          for (int j = 0; j < RF_G; ++j) stk[j] = (stk[j] << 1) | (uint32_t)rf_rel_i(rel, a[j], b[j]);
        }
      } else if (opc == PF_RES_AND) {
#pragma unroll
        for (int j = 0; j < RF_G; ++j) {
          const uint32_t t = stk[j] & 1u;
          stk[j] >>= 1;
          stk[j] &= ~1u | t;
        }
      } else if (opc == PF_RES_OR) {
#pragma unroll
        for (int j = 0; j < RF_G; ++j) {
          const uint32_t t = stk[j] & 1u;
          stk[j] = (stk[j] >> 1) | t;
        }
      } else if (opc == PF_RES_NOT) {
#pragma unroll
        for (int j = 0; j < RF_G; ++j) stk[j] ^= 1u;
      } else {  // PF_RES_TRUE
#pragma unroll
        for (int j = 0; j < RF_G; ++j) stk[j] = (stk[j] << 1) | 1u;
      }
    }
#pragma unroll
    for (int j = 0; j < RF_G; ++j) {
      const bool keep = live[j] && (stk[j] & 1u);
      const uint64_t bal = ballot(keep);
      if (lane_id() == 0) words[(base + (int64_t)(g + j) * CT_NT) / kWave + w] = bal;
      if (mask_out && live[j]) mask_out[base + (int64_t)(g + j) * CT_NT + threadIdx.x] = keep;
      kept += keep;
    }
  }
  kept = block_sum_ct(kept, red);
  if (threadIdx.x == 0) tilecnt[blockIdx.x] = kept;
  bad = block_sum_ct(bad, red);
  if (threadIdx.x == 0 && bad) atomicAdd((unsigned long long*)bad_out, (unsigned long long)bad);
}

// The kept pairs of a tile at its scanned offset, in pair order.
__global__ __launch_bounds__(CT_NT) void k_rf_compact(const int64_t* __restrict__ left,
                                                      const int64_t* __restrict__ right,
                                                      const uint64_t* __restrict__ words,
                                                      const int64_t* __restrict__ offs,
                                                      int64_t* __restrict__ left_out,
                                                      int64_t* __restrict__ right_out) {
  __shared__ int cnt[kWave];
  const int64_t base = (int64_t)blockIdx.x * CT_TILE;
  const int w = threadIdx.x / kWave, lane = lane_id();
  bool keep[CT_IT];
#pragma unroll
  for (int j = 0; j < CT_IT; ++j)
    keep[j] = (words[(int64_t)blockIdx.x * (CT_TILE / kWave) + j * CT_W + w] >> lane) & 1u;
  int64_t pos[CT_IT];
  tile_positions(keep, offs[blockIdx.x], pos, cnt);
#pragma unroll
  for (int j = 0; j < CT_IT; ++j) {
    if (keep[j]) {
      const int64_t i = base + j * CT_NT + threadIdx.x;
      left_out[pos[j]] = left[i];
      right_out[pos[j]] = right[i];
    }
  }
}

struct RfWs {
  uint64_t* words;  // tiles * 64 ballot words
  int32_t* tilecnt;
  int64_t* offs;
  int64_t* dev;  // [0] kept pairs, [1] indexes outside their column
  void* scan;
  size_t bytes;
};

RfWs carve_rf(void* ws, int64_t n) {
  Carve cv(ws);
  RfWs s;
  const int64_t tiles = n > 0 ? (n + CT_TILE - 1) / CT_TILE : 1;
  s.words = cv.take<uint64_t>(tiles * (CT_TILE / kWave));
  s.tilecnt = cv.take<int32_t>(tiles);
  s.offs = cv.take<int64_t>(tiles);
  s.dev = cv.take<int64_t>(2);
  s.scan = cv.take<char>(scan_i32_scratch(tiles));
  s.bytes = cv.bytes();
  return s;
}

int unique_hash_run(const int64_t* b0, const int64_t* b1, int64_t nb, const int64_t* p0,
                    const int64_t* p1, int64_t np, int32_t nkeys, int64_t* build_idx,
                    int64_t* probe_idx, int64_t* host_out, void* workspace,
                    size_t workspace_bytes, hipStream_t st) {
  if (nkeys != 1 && nkeys != 2) return fail("join_core: one or two aligned build/probe keys");
  if (nb <= 0 || np <= 0) return fail("join_core: unique hash needs rows on both sides");
  if (nb > MAX_ROWS || np > MAX_ROWS)
    return fail("join_core: a call takes at most " + std::to_string(MAX_ROWS) + " rows a side");
  const bool two = nkeys == 2;
  UniqueWs u = carve_unique(workspace, nb, np, two);
  if (workspace_bytes < u.bytes) return fail("join_core: workspace too small");
  void* hdev = nullptr;
  if (host_alias(host_out, &hdev)) return 1;
  const uint32_t cap = (uint32_t)(2 * nb);
  const int64_t tiles = (np + CT_TILE - 1) / CT_TILE;
  if (two) {
    if (fill_i32(u.hrow, cap, -1, st)) return 1;
    k_uh_build2<<<grid_for(nb, 256), 256, 0, st>>>(b0, b1, nb, u.hrow, cap);
    k_uh_probe<true><<<(unsigned)tiles, CT_NT, 0, st>>>(b0, b1, p0, p1, np, u.hkey, u.hrow,
                                                        cap, u.special, u.partner, u.tilecnt);
  } else {
    if (fill_i64(u.hkey, cap, kEmptyKey, st)) return 1;
    if (fill_i32(u.hrow, cap, INT32_MAX, st)) return 1;
    if (fill_i32(u.special, 1, INT32_MAX, st)) return 1;
    k_uh_build1<<<grid_for(nb, 256), 256, 0, st>>>(b0, nb, u.hkey, u.hrow, cap, u.special);
    {
    KTimer _kt("uh_probe", st);
    k_uh_probe<false><<<(unsigned)tiles, CT_NT, 0, st>>>(b0, b0, p0, p0, np, u.hkey, u.hrow,
                                                         cap, u.special, u.partner,
                                                         u.tilecnt);
  }
  }
  if (launch_check("join_unique_hash")) return 1;
  if (scan_i32_to_i64_excl(u.tilecnt, u.offs, tiles, u.count, u.scan, st)) return 1;
  k_readback<<<1, 64, 0, st>>>(u.count, 1, (int64_t*)hdev);
  {
    KTimer _kt("uh_compact", st);
    k_uh_compact<<<(unsigned)tiles, CT_NT, 0, st>>>(u.partner, np, u.offs, probe_idx, build_idx);
  }
  if (launch_check("join_unique_hash")) return 1;
  if (wait_stream(st, "join_unique_hash")) return 1;
  return ok();
}

// The allocation-free unique-hash call's scratch: one fixed array a device
// holds from its first use until the process ends.  Host code never passes
// `scratch` itself to a kernel (on the host it names a stand-in, not device
// memory): each device's address comes from hipGetSymbolAddress, taken once
// per device with that device current.
__device__ __align__(256) unsigned char scratch[PF_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The current device's address of `scratch`.
int noalloc_scratch(void** out) {
  int dev = 0;
  if (hip_check(hipGetDevice(&dev), "join_unique_hash_noalloc")) return 1;
  if (dev < 0 || dev >= kMaxDevices)
    return fail("join_unique_hash_noalloc: device " + std::to_string(dev) + " out of range");
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)),
                  "join_unique_hash_noalloc: scratch address"))
      return 1;
    if (!p) return fail("join_unique_hash_noalloc: no scratch address on this device");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

}  // namespace
}  // namespace pf

extern "C" {

size_t pf_join_unique_hash_workspace_bytes(int64_t nb, int64_t np, int32_t nkeys) {
  return pf::carve_unique(nullptr, nb, np, nkeys == 2).bytes;
}

int pf_join_unique_hash(const int64_t* b0, const int64_t* b1, int64_t nb, const int64_t* p0,
                        const int64_t* p1, int64_t np, int32_t nkeys, int64_t* build_idx,
                        int64_t* probe_idx, int64_t* host_out, void* workspace,
                        size_t workspace_bytes, void* stream) {
  return pf::unique_hash_run(b0, b1, nb, p0, p1, np, nkeys, build_idx, probe_idx, host_out,
                             workspace, workspace_bytes, (hipStream_t)stream);
}

int pf_join_unique_hash_noalloc(const int64_t* b0, const int64_t* b1, int64_t nb,
                                const int64_t* p0, const int64_t* p1, int64_t np,
                                int32_t nkeys, int64_t* build_idx, int64_t* probe_idx,
                                int64_t* host_out, void* stream) {
  using namespace pf;
  if (nb > 0 && np > 0 && (nkeys == 1 || nkeys == 2)) {
    const size_t need = carve_unique(nullptr, nb, np, nkeys == 2).bytes;
    if (need > (size_t)PF_NOALLOC_SCRATCH_BYTES)
      return fail("join_unique_hash_noalloc: " + std::to_string(nb) + " build and " +
                  std::to_string(np) + " probe rows need " + std::to_string(need) +
                  " bytes of scratch, over the fixed " +
                  std::to_string((size_t)PF_NOALLOC_SCRATCH_BYTES));
  }
  void* ws = nullptr;
  if (noalloc_scratch(&ws)) return 1;
  return unique_hash_run(b0, b1, nb, p0, p1, np, nkeys, build_idx, probe_idx, host_out, ws,
                         (size_t)PF_NOALLOC_SCRATCH_BYTES, (hipStream_t)stream);
}

size_t pf_join_semi_anti_workspace_bytes(int64_t nb, int64_t np, int32_t op) {
  return pf::carve_set(nullptr, nb, np, op, true).bytes;
}

int pf_join_semi_anti(const int64_t* bk, int64_t nb, const int64_t* pk, int64_t np,
                      const int64_t* bv, const int64_t* pv, int32_t op, int32_t anti,
                      int64_t* out, int64_t* host_out, void* workspace, size_t workspace_bytes,
                      void* stream) {
  using namespace pf;
  hipStream_t st = (hipStream_t)stream;
  if (nb <= 0 || np <= 0) return fail("join_core: semi/anti needs rows on both sides");
  if (np > MAX_ROWS || nb > MAX_ROWS)
    return fail("join_core: a call takes at most " + std::to_string(MAX_ROWS) + " rows a side");
  if (op < OP_NONE || op > OP_GE)
    return fail("join_core: unknown residual operator " + std::to_string(op));
  if (op != OP_NONE && (!bv || !pv)) return fail("join_core: a residual needs both value columns");
  SetWs s = carve_set(workspace, nb, np, op, true);
  if (workspace_bytes < s.bytes) return fail("join_core: workspace too small");
  void* hdev = nullptr;
  if (host_alias(host_out, &hdev)) return 1;
  if (build_set(s, bk, bv, nb, op, st)) return 1;
  const int64_t tiles = (np + CT_TILE - 1) / CT_TILE;
  k_set_probe<<<(unsigned)tiles, CT_NT, 0, st>>>(pk, pv, np, s.skey, s.mn, s.mx,
                                                 (uint32_t)(2 * nb), op, anti, s.keep,
                                                 s.tilecnt);
  if (launch_check("join_semi_anti")) return 1;
  if (scan_i32_to_i64_excl(s.tilecnt, s.offs, tiles, s.dev, s.scan, st)) return 1;
  k_readback<<<1, 64, 0, st>>>(s.dev, 2, (int64_t*)hdev);
  k_rows_compact<<<(unsigned)tiles, CT_NT, 0, st>>>(s.keep, np, s.offs, out);
  if (launch_check("join_semi_anti")) return 1;
  if (wait_stream(st, "join_semi_anti")) return 1;
  return ok();
}

size_t pf_semi_anti_mask_workspace_bytes(int64_t nb, int64_t np, int32_t op) {
  return pf::carve_set(nullptr, nb, np, op, false).bytes;
}

int pf_semi_anti_mask(const int64_t* bk, int64_t nb, const int64_t* pk, int64_t np,
                      const int64_t* bv, const int64_t* pv, int32_t op, int32_t anti,
                      uint8_t* mask_out, int64_t* bad_out, void* workspace,
                      size_t workspace_bytes, void* stream) {
  using namespace pf;
  hipStream_t st = (hipStream_t)stream;
  if (nb <= 0 || np < 0) return fail("semi_anti_mask: needs build rows");
  if (nb > MAX_ROWS) return fail("semi_anti_mask: too many build rows");
  if (op < OP_NONE || op > OP_GE)
    return fail("semi_anti_mask: unknown residual operator " + std::to_string(op));
  if (op != OP_NONE && (!bv || (np > 0 && !pv)))
    return fail("semi_anti_mask: a residual needs both columns");
  SetWs s = carve_set(workspace, nb, np, op, false);
  if (workspace_bytes < s.bytes) return fail("semi_anti_mask: workspace too small");
  if (build_set(s, bk, bv, nb, op, st)) return 1;
  if (np > 0) {
    const int64_t tiles = (np + CT_TILE - 1) / CT_TILE;
    k_set_probe<<<(unsigned)tiles, CT_NT, 0, st>>>(pk, pv, np, s.skey, s.mn, s.mx,
                                                   (uint32_t)(2 * nb), op, anti, mask_out,
                                                   nullptr);
    if (launch_check("semi_anti_mask")) return 1;
  }
  if (bad_out) {
    k_readback<<<1, 64, 0, st>>>(s.dev + 1, 1, bad_out);
    if (launch_check("semi_anti_mask")) return 1;
  }
  return ok();
}

size_t pf_select_rows_workspace_bytes(int64_t n) {
  pf::Carve cv(nullptr);
  const int64_t tiles = (n + pf::CT_TILE - 1) / pf::CT_TILE;
  cv.take<int32_t>(tiles > 0 ? tiles : 1);
  cv.take<int64_t>(tiles > 0 ? tiles : 1);
  cv.take<int64_t>(2);
  cv.take<char>(pf::scan_i32_scratch(tiles));
  return cv.bytes();
}

int pf_select_rows(const uint8_t* flags, int64_t n, int32_t want, int64_t* out,
                   int64_t* host_out, void* workspace, size_t workspace_bytes, void* stream) {
  using namespace pf;
  hipStream_t st = (hipStream_t)stream;
  if (n < 0) return fail("select_rows: negative rows");
  if (workspace_bytes < pf_select_rows_workspace_bytes(n))
    return fail("select_rows: workspace too small");
  void* hdev = nullptr;
  if (host_alias(host_out, &hdev)) return 1;
  Carve cv(workspace);
  const int64_t tiles = (n + CT_TILE - 1) / CT_TILE;
  int32_t* tilecnt = cv.take<int32_t>(tiles > 0 ? tiles : 1);
  int64_t* offs = cv.take<int64_t>(tiles > 0 ? tiles : 1);
  int64_t* total = cv.take<int64_t>(2);
  void* scan = cv.take<char>(scan_i32_scratch(tiles));
  if (n > 0) {
    k_flag_count<<<(unsigned)tiles, CT_NT, 0, st>>>(flags, n, want != 0, tilecnt);
    if (launch_check("select_rows")) return 1;
  }
  if (scan_i32_to_i64_excl(tilecnt, offs, tiles, total, scan, st)) return 1;
  k_readback<<<1, 64, 0, st>>>(total, 1, (int64_t*)hdev);
  if (n > 0) k_flag_compact<<<(unsigned)tiles, CT_NT, 0, st>>>(flags, n, want != 0, offs, out);
  if (launch_check("select_rows")) return 1;
  if (wait_stream(st, "select_rows")) return 1;
  return ok();
}

size_t pf_join_expand_workspace_bytes(int64_t m) {
  pf::Carve cv(nullptr);
  cv.take<int64_t>(m > 0 ? m : 1);
  cv.take<int64_t>(2);
  cv.take<char>(pf::scan_i64_scratch(m));
  return cv.bytes();
}

int pf_join_expand_count(const int64_t* len, int64_t m, int64_t* host_out, void* workspace,
                         size_t workspace_bytes, void* stream) {
  using namespace pf;
  hipStream_t st = (hipStream_t)stream;
  if (m < 0 || m > MAX_ROWS) return fail("join_core: start/len rows out of range");
  if (workspace_bytes < pf_join_expand_workspace_bytes(m))
    return fail("join_core: workspace too small");
  void* hdev = nullptr;
  if (host_alias(host_out, &hdev)) return 1;
  Carve cv(workspace);
  int64_t* ex = cv.take<int64_t>(m > 0 ? m : 1);
  int64_t* total = cv.take<int64_t>(2);
  void* scan = cv.take<char>(scan_i64_scratch(m));
  if (scan_i64_excl_from(len, ex, m, total, scan, st)) return 1;
  k_readback<<<1, 64, 0, st>>>(total, 1, (int64_t*)hdev);
  if (launch_check("join_expand")) return 1;
  // the one wait: the pair count sizes the outputs
  if (wait_stream(st, "join_expand")) return 1;
  return ok();
}

int pf_join_expand_write(const int64_t* sorted_ids, const int64_t* start, int64_t m,
                         int64_t n_flat, int64_t* cs, int64_t* cr, void* workspace,
                         void* stream) {
  using namespace pf;
  hipStream_t st = (hipStream_t)stream;
  if (n_flat <= 0 || m <= 0) return ok();
  const int64_t* ex = (const int64_t*)workspace;  // the count call's exclusive offsets
  k_expand<<<grid_for(n_flat, EX_TILE), EX_NT, 0, st>>>(sorted_ids, start, ex, m, n_flat, cs,
                                                        cr);
  if (launch_check("join_expand")) return 1;
  return ok();
}

size_t pf_pair_filter_workspace_bytes(int64_t n) { return pf::carve_rf(nullptr, n).bytes; }

int pf_pair_filter_count(const int64_t* left_idx, const int64_t* right_idx, int64_t n,
                         const void* const* cols, const int32_t* col_types,
                         const int32_t* col_sides, const int64_t* col_rows, int32_t n_cols,
                         const int64_t* ins, int32_t n_ins, uint8_t* mask_out,
                         int64_t* host_out, void* workspace, size_t workspace_bytes,
                         void* stream) {
  using namespace pf;
  hipStream_t st = (hipStream_t)stream;
  if (n < 0) return fail("pair_filter: negative pairs");
  RfProgram P;
  if (rf_program_of(cols, col_types, col_sides, col_rows, n_cols, ins, n_ins, &P)) return 1;
  if (n > 0 && (!left_idx || !right_idx)) return fail("pair_filter: missing pair indexes");
  if (workspace_bytes < pf_pair_filter_workspace_bytes(n))
    return fail("pair_filter: workspace too small");
  void* hdev = nullptr;
  if (host_alias(host_out, &hdev)) return 1;
  RfWs s = carve_rf(workspace, n);
  const int64_t tiles = (n + CT_TILE - 1) / CT_TILE;
  if (fill_i64(s.dev, 2, 0, st)) return 1;
  if (n > 0) {
    KTimer _kt("rf_eval", st);
    k_rf_eval<<<(unsigned)tiles, CT_NT, 0, st>>>(left_idx, right_idx, n, P, s.words, mask_out,
                                                  s.tilecnt, s.dev + 1);
    if (launch_check("pair_filter eval")) return 1;
    if (scan_i32_to_i64_excl(s.tilecnt, s.offs, tiles, s.dev, s.scan, st)) return 1;
  }
  k_readback<<<1, 64, 0, st>>>(s.dev, 2, (int64_t*)hdev);
  if (launch_check("pair_filter")) return 1;
  if (wait_stream(st, "pair_filter")) return 1;
  return ok();
}

int pf_pair_filter_write(const int64_t* left_idx, const int64_t* right_idx, int64_t n,
                         int64_t* left_out, int64_t* right_out, void* workspace,
                         void* stream) {
  using namespace pf;
  hipStream_t st = (hipStream_t)stream;
  if (n < 0) return fail("pair_filter: negative pairs");
  if (n == 0) return ok();
  RfWs s = carve_rf(workspace, n);
  const int64_t tiles = (n + CT_TILE - 1) / CT_TILE;
  {
    KTimer _kt("rf_compact", st);
    k_rf_compact<<<(unsigned)tiles, CT_NT, 0, st>>>(left_idx, right_idx, s.words, s.offs,
                                                     left_out, right_out);
  }
  if (launch_check("pair_filter write")) return 1;
  return ok();
}

}  // extern "C"
