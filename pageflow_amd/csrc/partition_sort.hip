// Stable hash partition of one tick into the padded per-bucket layout: the
// drop-in for the reference's partition_sort entry point
// (operators/kernels/partition_sort.cu:610-768, contract
// operators/partition_sort.py:68-143).
//
// A row's bucket is the top log2(B) bits of mix64(key) (bucket 0 when B is
// 1); rows keep their source order inside a bucket; each output (every
// column, then the int64 gidx = goff + pre-mask row) starts every non-empty
// bucket on a whole lcm(row bytes, 512)-byte unit, padding rows unwritten.
//
// gfx950 design: a 4096-row tile per 1024-thread workgroup (16 wave64s, four
// 64-row rounds per wave).  The stable rank inside a tile comes from
// ballots (one per bucket bit plus one for validity) and mbcnt, with
// per-wave running counts in LDS as u16; the tile is staged in LDS in
// bucket order and each column leaves as coalesced bucket runs.
//
//   prologue (masks)  ps_mask_count + ps_scan_counts + ps_mask_select
//                     (select-first: kept rows' pre-mask ids), or
//                     ps_tile_count + ps_scan_counts (compacted-input:
//                     each tile's offset into the compacted columns)
//   ps_hist           per-tile bucket counts, hist[b][tile]
//   ps_bucket_scan    per-bucket exclusive scan over the tiles
//   ps_finish         counts and padded totals into host_out, padded bucket
//                     starts per row-width class, the overflow check
//   ps_scatter        rank, stage, write
//
// The host waits once, for ps_finish's read-back, while ps_scatter runs.
#include <hip/hip_runtime.h>

#include <numeric>
#include <string>

#include "pf_common.hpp"
#include "../../include/pageflow_join.h"

namespace pf {
namespace {

constexpr int MAXC = 6;     // data columns a call
constexpr int MAXCLS = 4;   // row-width classes, the gidx's 8 B included
constexpr int BMAX = 512;   // buckets
constexpr int LOGBMAX = 9;
constexpr int NT = 1024;    // threads a tile (4096-row tiles: 3.59 -> 3.31 ms
                            // against 512, aggregation pass 1 into 256 buckets)
constexpr int NW = NT / kWave;
constexpr int IT = 4;       // rounds a wave
constexpr int TILE = NT * IT;
constexpr int WROWS = kWave * IT;  // rows of one wave's segment
constexpr int MS_NT = 256;         // select-first prologue: 16 mask bytes a thread
constexpr int MS_ROWS = MS_NT * 16;
constexpr int SC_NT = 1024;        // one-block count scan

struct ColumnSet {
  const uint8_t* in[MAXC];
  uint8_t* out[MAXC];
  int rb[MAXC];
  int comp[MAXC];
  int cls[MAXC];
  int ncols;
};

struct PsArgs {
  ColumnSet c;
  int64_t* gout;          // null: no gidx output
  const int32_t* bucket;  // the rows' buckets given (pf_partition_by_bucket), or null
  int gcls;
  int64_t goff;
  const int64_t* key;
  int key_comp;
  const uint8_t* mask;    // null after select-first
  int64_t ndom;           // rows the tiles walk
  const int32_t* blk_coff;
  int needs_j;            // compacted columns read at their compact index
  const int32_t* src;     // select-first: kept row i's pre-mask row
  int B, logB, ncls;
  int64_t nblk;
  int32_t* hist;          // [B][nblk], scanned in place
  const int64_t* bases;   // [MAXCLS][B]
  int32_t* flags;         // [1] prologue count != n_valid, [2] counts > n_valid, [3] prologue
                          // total, [4] given buckets out of range
  int check;              // bit 0: flags[1] applies, bit 1: flags[2] applies
};

__device__ __forceinline__ int bucket_of(int64_t key, int logB) {
  return logB ? (int)(mix64((uint64_t)key) >> (64 - logB)) : 0;
}

// The row of a column that holds domain row i: a compacted column's is its
// position among the kept rows, a full column's the row before the mask
// (after select-first, the id the prologue recorded for kept row i).
template <bool SRC>
__device__ __forceinline__ int64_t domain_row(const PsArgs& p, bool comp, int64_t i, int jl,
                                             int64_t coff) {
  if (SRC) return comp ? i : (int64_t)p.src[i];
  return comp ? coff + jl : i;
}

// Lanes of the wave holding the same bucket (b < 0: not a row).
__device__ __forceinline__ uint64_t match_bucket(int b, int logB) {
  const uint64_t v = ballot(b >= 0);
  uint64_t m = (b >= 0) ? v : ~v;
#pragma unroll
  for (int bit = 0; bit < LOGBMAX; ++bit) {
    if (bit < logB) {
      const bool on = (b >> bit) & 1;
      const uint64_t bal = ballot(on);
      m &= on ? bal : ~bal;
    }
  }
  return m;
}

template <bool SRC>
__device__ __forceinline__ void load_valid(const PsArgs& p, int64_t seg, bool (&valid)[IT]) {
  const int lane = lane_id();
#pragma unroll
  for (int k = 0; k < IT; ++k) {
    const int64_t i = seg + k * kWave + lane;
    bool v = i < p.ndom;
    if (!SRC && v && p.mask) v = p.mask[i] != 0;
    valid[k] = v;
  }
}

// Tile-local compact index of every valid row (row order); contains a
// __syncthreads.
__device__ __forceinline__ void tile_compact(const bool (&valid)[IT], int (&jl)[IT],
                                             int32_t* wv) {
  const int lane = lane_id(), w = threadIdx.x / kWave;
  int run = 0;
#pragma unroll
  for (int k = 0; k < IT; ++k) {
    const uint64_t bal = ballot(valid[k]);
    jl[k] = run + mbcnt(bal);
    run += __popcll(bal);
  }
  if (lane == 0) wv[w] = run;
  __syncthreads();
  int pre = 0;
#pragma unroll
  for (int q = 0; q < NW; ++q) pre += q < w ? wv[q] : 0;
#pragma unroll
  for (int k = 0; k < IT; ++k) jl[k] += pre;
}

// ---------------------------------------------------------- prologues --

__device__ __forceinline__ void load16(const uint8_t* __restrict__ m, int64_t r0, int64_t n,
                                       uint8_t (&v)[16]) {
  if (((uintptr_t)(m + r0) & 15) == 0 && r0 + 16 <= n) {
    const uint4 x = *reinterpret_cast<const uint4*>(m + r0);
    const uint8_t* b = reinterpret_cast<const uint8_t*>(&x);
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = b[q];
  } else {
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = (r0 + q < n) ? m[r0 + q] : 0;
  }
}

// Kept rows of each block's MS_ROWS mask rows.
__global__ __launch_bounds__(MS_NT) void ps_mask_count(const uint8_t* __restrict__ mask,
                                                       int64_t n, int32_t* __restrict__ cnt) {
  __shared__ int32_t red[MS_NT / kWave];
  const int64_t r0 = ((int64_t)blockIdx.x * MS_NT + threadIdx.x) * 16;
  uint8_t v[16];
  load16(mask, r0, n, v);
  int c = 0;
#pragma unroll
  for (int q = 0; q < 16; ++q) c += v[q] != 0;
  c = wave_sum(c);
  if (lane_id() == 0) red[threadIdx.x / kWave] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
#pragma unroll
    for (int w = 0; w < MS_NT / kWave; ++w) t += red[w];
    cnt[blockIdx.x] = t;
  }
}

// Kept rows of each tile.
__global__ __launch_bounds__(NT) void ps_tile_count(const uint8_t* __restrict__ mask, int64_t n,
                                                    int32_t* __restrict__ cnt) {
  __shared__ int32_t red[NW];
  const int64_t base = (int64_t)blockIdx.x * TILE;
  int c = 0;
#pragma unroll
  for (int k = 0; k < IT; ++k) {
    const int64_t i = base + k * NT + threadIdx.x;
    c += (i < n && mask[i]) ? 1 : 0;
  }
  c = wave_sum(c);
  if (lane_id() == 0) red[threadIdx.x / kWave] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) t += red[w];
    cnt[blockIdx.x] = t;
  }
}

// One block: exclusive scan in place over a prologue's n counts; flags[3]
// gets the total and flags[1] is set when it is not `expect`.
__global__ __launch_bounds__(SC_NT) void ps_scan_counts(int32_t* a, int64_t n, int32_t* flags,
                                                        int64_t expect) {
  __shared__ int64_t tmp[SC_NT / kWave + 1];
  constexpr int PER = 4, CH = SC_NT * PER;
  int64_t running = 0;
  for (int64_t base = 0; base < n; base += CH) {
    int64_t v[PER], s = 0;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int64_t i = base + threadIdx.x * PER + q;
      v[q] = i < n ? a[i] : 0;
      s += v[q];
    }
    int64_t agg;
    int64_t ex = block_excl_scan<SC_NT>(s, tmp, &agg) + running;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int64_t i = base + threadIdx.x * PER + q;
      if (i < n) a[i] = (int32_t)ex;
      ex += v[q];
    }
    running += agg;
  }
  if (threadIdx.x == 0) {
    flags[3] = (int32_t)running;
    if (running != expect) flags[1] = 1;
  }
}

// src[j] = pre-mask row of kept row j (at most cap written).
__global__ __launch_bounds__(MS_NT) void ps_mask_select(const uint8_t* __restrict__ mask,
                                                        int64_t n,
                                                        const int32_t* __restrict__ coff,
                                                        int32_t* __restrict__ src, int64_t cap) {
  __shared__ int32_t tmp[MS_NT / kWave + 1];
  const int64_t r0 = ((int64_t)blockIdx.x * MS_NT + threadIdx.x) * 16;
  uint8_t v[16];
  load16(mask, r0, n, v);
  int c = 0;
#pragma unroll
  for (int q = 0; q < 16; ++q) c += v[q] != 0;
  int tot;
  int ex = block_excl_scan<MS_NT>(c, tmp, &tot);
  int64_t o = (int64_t)coff[blockIdx.x] + ex;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    if (v[q]) {
      if (o < cap) src[o] = (int32_t)(r0 + q);
      ++o;
    }
  }
}

// ---------------------------------------------------------------- K1 --

template <bool SRC>
__global__ __launch_bounds__(NT) void ps_hist(PsArgs p) {
  __shared__ int32_t h[BMAX];
  __shared__ int32_t wv[NW];
  const int tid = threadIdx.x, w = tid / kWave;
  const int64_t blk = blockIdx.x;
  for (int b = tid; b < p.B; b += NT) h[b] = 0;
  if ((p.check & 1) && p.flags[1]) {  // a wrong count: the host raises
    for (int b = tid; b < p.B; b += NT) p.hist[(int64_t)b * p.nblk + blk] = 0;
    return;
  }
  const int64_t seg = blk * TILE + (int64_t)w * WROWS;
  bool valid[IT];
  load_valid<SRC>(p, seg, valid);
  int jl[IT] = {};
  int64_t coff = 0;
  if (!SRC && p.key_comp) {
    tile_compact(valid, jl, wv);  // its barrier also ends the zeroing
    coff = p.blk_coff[blk];
  } else {
    __syncthreads();
  }
  const int lane = lane_id();
#pragma unroll
  for (int k = 0; k < IT; ++k) {
    if (valid[k]) {
      const int64_t i = seg + k * kWave + lane;
      int b;
      if (p.bucket) {  // given: no mask, so row i itself
        b = p.bucket[i];
        if ((unsigned)b >= (unsigned)p.B) {  // dropped, and reported
          atomicAdd(&p.flags[4], 1);
          continue;
        }
      } else {
        b = bucket_of(p.key[domain_row<SRC>(p, p.key_comp, i, jl[k], coff)], p.logB);
      }
      atomicAdd(&h[b], 1);
    }
  }
  __syncthreads();
  for (int b = tid; b < p.B; b += NT) p.hist[(int64_t)b * p.nblk + blk] = h[b];
}

// ---------------------------------------------------------------- K2 --

// Block b: exclusive scan of hist row b in place; counts[b] = its total.
__global__ __launch_bounds__(NT) void ps_bucket_scan(int32_t* __restrict__ hist, int64_t nblk,
                                                     int64_t* __restrict__ counts) {
  __shared__ int64_t tmp[NW + 1];
  constexpr int PER = 4, CH = NT * PER;
  int32_t* row = hist + (int64_t)blockIdx.x * nblk;
  int64_t running = 0;
  for (int64_t base = 0; base < nblk; base += CH) {
    int64_t v[PER], s = 0;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int64_t i = base + threadIdx.x * PER + q;
      v[q] = i < nblk ? row[i] : 0;
      s += v[q];
    }
    int64_t agg;
    int64_t ex = block_excl_scan<NT>(s, tmp, &agg) + running;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int64_t i = base + threadIdx.x * PER + q;
      if (i < nblk) row[i] = (int32_t)ex;
      ex += v[q];
    }
    running += agg;
  }
  if (threadIdx.x == 0) counts[blockIdx.x] = running;
}

struct Rpa {
  int64_t v[MAXCLS];
};

// One block: host_out[0, B) the counts, host_out[B + c] class c's padded
// total, host_out[B + 4] the kept rows a prologue found (-1: none ran);
// bases[c][b] the padded start of bucket b in class c; flags[2] when the
// counts total more than the n_valid the outputs were sized for.
__global__ __launch_bounds__(NT) void ps_finish(const int64_t* __restrict__ counts, int B,
                                                Rpa rpa, int ncls, int64_t* __restrict__ bases,
                                                int64_t* host_out, int32_t* flags,
                                                int64_t n_valid, int prologue, int given) {
  __shared__ int64_t tmp[NW + 1];
  const int tid = threadIdx.x;
  // B <= 512 = NT: one bucket a thread
  const int64_t n = tid < B ? counts[tid] : 0;
  if (tid < B) store_host(&host_out[tid], n);
  int64_t total;
  (void)block_excl_scan<NT>(n, tmp, &total);
  for (int c = 0; c < ncls; ++c) {
    const int64_t r = rpa.v[c];
    const int64_t padded = n ? ((n + r - 1) / r) * r : 0;
    int64_t ctot;
    const int64_t ex = block_excl_scan<NT>(padded, tmp, &ctot);
    if (tid < B) bases[(int64_t)c * B + tid] = ex;
    if (tid == 0) store_host(&host_out[B + c], ctot);
  }
  if (tid == 0) {
    if (total > n_valid) flags[2] = 1;
    store_host(&host_out[B + MAXCLS], prologue ? (int64_t)flags[3] : -1);
    if (given) store_host(&host_out[B + MAXCLS + 1], (int64_t)flags[4]);
  }
}

// ---------------------------------------------------------------- K3 --

template <int ES>
struct WordOf;
template <>
struct WordOf<1> { typedef uint8_t T; };
template <>
struct WordOf<2> { typedef uint16_t T; };
template <>
struct WordOf<4> { typedef uint32_t T; };
template <>
struct WordOf<8> { typedef uint64_t T; };

template <bool SRC>
__global__ __launch_bounds__(NT) void ps_scatter(PsArgs p) {
  if (((p.check & 1) && p.flags[1]) || ((p.check & 2) && p.flags[2])) return;
  __shared__ uint64_t stage[TILE];
  __shared__ int64_t gb[MAXCLS * BMAX];
  __shared__ uint16_t wh[NW * BMAX];
  __shared__ int32_t tstart[BMAX];
  __shared__ int32_t tot[BMAX];
  __shared__ uint16_t sb[TILE];
  __shared__ uint16_t sidx[TILE];
  __shared__ uint16_t sjl[TILE];
  __shared__ int32_t wv[NW];
  __shared__ int32_t misc[2];

  const int tid = threadIdx.x, lane = lane_id(), w = tid / kWave;
  const int B = p.B;
  const int64_t blk = blockIdx.x;
  const int64_t tile_base = blk * TILE;
  const int segl = w * WROWS;
  const int64_t seg = tile_base + segl;
  for (int x = tid; x < NW * B; x += NT) wh[x] = 0;

  bool valid[IT];
  load_valid<SRC>(p, seg, valid);
  int jl[IT] = {};
  int64_t coff = 0;
  if (!SRC && p.needs_j) {
    tile_compact(valid, jl, wv);
    coff = p.blk_coff[blk];
  } else {
    __syncthreads();
  }
  int bk[IT];
#pragma unroll
  for (int k = 0; k < IT; ++k) {
    const int64_t i = seg + k * kWave + lane;
    if (!valid[k]) {
      bk[k] = -1;
    } else if (p.bucket) {
      const int b = p.bucket[i];
      bk[k] = (unsigned)b < (unsigned)B ? b : -1;
    } else {
      bk[k] = bucket_of(p.key[domain_row<SRC>(p, p.key_comp, i, jl[k], coff)], p.logB);
    }
  }
  // stable rank inside (wave, bucket): rounds in row order, lanes in order
  int rk[IT];
#pragma unroll
  for (int k = 0; k < IT; ++k) {
    const uint64_t peers = match_bucket(bk[k], p.logB);
    const int below = mbcnt(peers);
    const int base = bk[k] >= 0 ? (int)wh[w * B + bk[k]] : 0;
    rk[k] = base + below;
    if (bk[k] >= 0 && below == 0) wh[w * B + bk[k]] = (uint16_t)(base + __popcll(peers));
  }
  __syncthreads();
  // per bucket: each wave's start inside the tile's run of the bucket
  for (int b = tid; b < B; b += NT) {
    int run = 0;
#pragma unroll
    for (int q = 0; q < NW; ++q) {
      const int t = wh[q * B + b];
      wh[q * B + b] = (uint16_t)run;
      run += t;
    }
    tot[b] = run;
  }
  __syncthreads();
  // the tile's bucket runs in bucket order: one wave, 8 buckets a lane
  if (w == 0) {
    constexpr int BPL = BMAX / kWave;
    int v[BPL], s = 0;
#pragma unroll
    for (int q = 0; q < BPL; ++q) {
      const int b = lane * BPL + q;
      v[q] = b < B ? tot[b] : 0;
      s += v[q];
    }
    const int incl = wave_incl_scan(s);
    int ex = incl - s;
#pragma unroll
    for (int q = 0; q < BPL; ++q) {
      const int b = lane * BPL + q;
      if (b < B) tstart[b] = ex;
      ex += v[q];
    }
    if (lane == kWave - 1) misc[0] = incl;
  }
  __syncthreads();
  // sorted slot s of bucket b lands on row gb[c][b] + s of class c's output
  for (int x = tid; x < p.ncls * B; x += NT) {
    const int c = x / B, b = x - c * B;
    gb[x] = p.bases[(int64_t)c * B + b] + p.hist[(int64_t)b * p.nblk + blk] - tstart[b];
  }
  int pos[IT];
#pragma unroll
  for (int k = 0; k < IT; ++k) {
    pos[k] = -1;
    if (bk[k] >= 0) {
      pos[k] = tstart[bk[k]] + wh[w * B + bk[k]] + rk[k];
      sb[pos[k]] = (uint16_t)bk[k];
      sidx[pos[k]] = (uint16_t)(segl + k * kWave + lane);
      sjl[pos[k]] = (uint16_t)jl[k];
    }
  }
  __syncthreads();
  const int tvalid = misc[0];

  for (int c = 0; c < p.c.ncols; ++c) {
    const int rb = p.c.rb[c];
    const uint8_t* in = p.c.in[c];
    uint8_t* out = p.c.out[c];
    const bool comp = p.c.comp[c] != 0;
    const int64_t* gbc = gb + p.c.cls[c] * B;
#define PF_STAGE(ES)                                                                   \
  {                                                                                    \
    typedef WordOf<ES>::T T;                                                               \
    _Pragma("unroll") for (int k = 0; k < IT; ++k) {                                   \
      if (pos[k] >= 0) {                                                               \
        const int64_t i = domain_row<SRC>(p, comp, seg + k * kWave + lane, jl[k], coff); \
        ((T*)stage)[pos[k]] = ((const T*)in)[i];                                       \
      }                                                                                \
    }                                                                                  \
    __syncthreads();                                                                   \
    for (int s = tid; s < tvalid; s += NT) ((T*)out)[gbc[sb[s]] + s] = ((T*)stage)[s]; \
    __syncthreads();                                                                   \
  }
    if (rb == 8) PF_STAGE(8)
    else if (rb == 4) PF_STAGE(4)
    else if (rb == 2) PF_STAGE(2)
    else if (rb == 1) PF_STAGE(1)
    else {
#undef PF_STAGE
      // wide or odd rows: consecutive threads on consecutive output words
      // (4-byte words when both sides allow, else bytes)
      const bool words = (rb % 4 == 0) && (((uintptr_t)in | (uintptr_t)out) % 4 == 0);
      const int unit = words ? 4 : 1;
      const int per = rb / unit;
      const int64_t total = (int64_t)tvalid * per;
      for (int64_t e = tid; e < total; e += NT) {
        const int s = (int)(e / per), q = (int)(e - (int64_t)s * per);
        const int64_t dom = tile_base + sidx[s];
        const int64_t i = domain_row<SRC>(p, comp, dom, sjl[s], coff);
        const int64_t o = gbc[sb[s]] + s;
        if (words)
          ((uint32_t*)out)[o * per + q] = ((const uint32_t*)in)[i * per + q];
        else
          out[o * rb + q] = in[i * rb + q];
      }
    }
  }
  if (p.gout) {
    const int64_t* gbg = gb + p.gcls * B;
    for (int s = tid; s < tvalid; s += NT) {
      const int64_t i = tile_base + sidx[s];
      p.gout[gbg[sb[s]] + s] = p.goff + (SRC ? (int64_t)p.src[i] : i);
    }
  }
}

int64_t rows_per_unit(int64_t rb) { return 512 / std::gcd(rb, (int64_t)512); }

struct Layout {
  int32_t* flags;
  int64_t* counts;
  int64_t* bases;
  int32_t* blk_cnt;
  int32_t* src;
  int32_t* hist;
  size_t bytes;
};

Layout carve_layout(void* ws, int64_t ndom, int64_t n_valid, int B, bool use_sel,
                    bool needs_j) {
  Carve cv(ws);
  Layout l;
  l.flags = cv.take<int32_t>(8);
  l.counts = cv.take<int64_t>(B);
  l.bases = cv.take<int64_t>((size_t)MAXCLS * B);
  const int64_t dom = use_sel ? n_valid : ndom;
  const int64_t nblk = dom > 0 ? (dom + TILE - 1) / TILE : 1;
  const int64_t mblk = ndom > 0 ? (ndom + MS_ROWS - 1) / MS_ROWS : 1;
  l.blk_cnt = cv.take<int32_t>(use_sel ? mblk : (needs_j ? nblk : 1));
  l.src = cv.take<int32_t>(use_sel ? (n_valid > 0 ? n_valid : 1) : 1);
  l.hist = cv.take<int32_t>((size_t)B * nblk);
  l.bytes = cv.bytes();
  return l;
}

}  // namespace
}  // namespace pf

extern "C" {

size_t pf_partition_sort_workspace_bytes(int64_t ndom, int64_t n_valid, int32_t B,
                                         int32_t has_mask, int32_t select_first,
                                         int32_t any_comp) {
  if (B < 1) B = 1;
  const bool use_sel = select_first && has_mask;
  const bool needs_j = any_comp && !use_sel;
  return pf::carve_layout(nullptr, ndom, n_valid, B, use_sel, needs_j).bytes;
}

}  // extern "C"

namespace pf {
namespace {

// The partition of both entry points: buckets from the key column's mix64,
// or (given) from the bucket column, with no key, mask or gidx.
int partition_impl(const void* const* cols_in, void* const* cols_out, const int32_t* row_bytes,
                   const int32_t* comp, int32_t ncols, int32_t key_idx, bool given,
                   const int32_t* bucket, const uint8_t* mask, int64_t ndom, int64_t n_valid, int64_t goff, int32_t B,
                   int32_t select_first, int64_t* gidx_out, int64_t* host_out, void* workspace,
                   size_t workspace_bytes, hipStream_t st) {
  if (ncols < 1 || ncols > MAXC)
    return fail("partition_sort: 1-" + std::to_string(MAXC) + " columns, got " +
                std::to_string(ncols));
  if (B < 1 || B > BMAX || (B & (B - 1)))
    return fail("partition_sort: B must be a power of two in [1, " + std::to_string(BMAX) +
                "], got " + std::to_string(B));
  if (!given) {
    if (key_idx < 0 || key_idx >= ncols)
      return fail("partition_sort: key index out of range");
    if (row_bytes[key_idx] != 8) return fail("partition_sort: the key must be 8-byte int64");
  }
  if (ndom < 0 || ndom >= (1LL << 31))
    return fail("partition_sort: a tick must hold fewer than 2^31 rows");
  if (n_valid < 0 || n_valid > ndom)
    return fail("partition_sort: n_valid " + std::to_string(n_valid) + " outside [0, " +
                std::to_string(ndom) + "]");
  const bool has_mask = mask != nullptr;
  PsArgs p{};
  int rb_cls[MAXCLS];
  int ncls = 0;
  bool bad_cls = false;
  auto cls_of = [&](int rb) {
    for (int c = 0; c < ncls; ++c)
      if (rb_cls[c] == rb) return c;
    if (ncls >= MAXCLS) {
      bad_cls = true;
      return 0;
    }
    rb_cls[ncls] = rb;
    return ncls++;
  };
  bool any_comp = false;
  for (int c = 0; c < ncols; ++c) {
    if (row_bytes[c] < 1) return fail("partition_sort: empty rows");
    if (comp[c] && !has_mask) return fail("partition_sort: a compacted column needs the mask");
    p.c.in[c] = (const uint8_t*)cols_in[c];
    p.c.out[c] = (uint8_t*)cols_out[c];
    p.c.rb[c] = row_bytes[c];
    p.c.comp[c] = comp[c] ? 1 : 0;
    p.c.cls[c] = cls_of(row_bytes[c]);
    any_comp |= comp[c] != 0;
  }
  p.c.ncols = ncols;
  if (gidx_out) p.gcls = cls_of(8);
  if (bad_cls)
    return fail("partition_sort: more than " + std::to_string(MAXCLS) + " row widths");
  const bool use_sel = select_first && has_mask;
  p.needs_j = any_comp && !use_sel;
  Layout l = carve_layout(workspace, ndom, n_valid, B, use_sel, p.needs_j);
  if (workspace_bytes < l.bytes)
    return fail("partition_sort: workspace of " + std::to_string(workspace_bytes) +
                " bytes, needs " + std::to_string(l.bytes));
  void* hdev = nullptr;
  if (host_alias(host_out, &hdev)) return 1;

  p.gout = gidx_out;
  p.bucket = bucket;
  p.goff = goff;
  p.key = given ? nullptr : (const int64_t*)cols_in[key_idx];
  p.key_comp = !given && comp[key_idx] ? 1 : 0;
  p.mask = mask;
  p.ndom = ndom;
  p.B = B;
  p.ncls = ncls;
  int logB = 0;
  while ((1 << logB) < B) ++logB;
  p.logB = logB;
  p.flags = l.flags;
  p.check = ((use_sel || p.needs_j) ? 1 : 0) | (has_mask ? 2 : 0);
  if (fill_i32(l.flags, 8, 0, st)) return 1;

  if (use_sel) {
    const int64_t mblk = ndom > 0 ? (ndom + MS_ROWS - 1) / MS_ROWS : 1;
    ps_mask_count<<<(unsigned)mblk, MS_NT, 0, st>>>(mask, ndom, l.blk_cnt);
    ps_scan_counts<<<1, SC_NT, 0, st>>>(l.blk_cnt, mblk, l.flags, n_valid);
    ps_mask_select<<<(unsigned)mblk, MS_NT, 0, st>>>(mask, ndom, l.blk_cnt, l.src, n_valid);
    if (launch_check("partition_sort prologue")) return 1;
    p.src = l.src;
    p.mask = nullptr;
    p.ndom = n_valid;
  }
  const int64_t nblk = p.ndom > 0 ? (p.ndom + TILE - 1) / TILE : 1;
  p.nblk = nblk;
  if (p.needs_j) {
    ps_tile_count<<<(unsigned)nblk, NT, 0, st>>>(mask, ndom, l.blk_cnt);
    ps_scan_counts<<<1, SC_NT, 0, st>>>(l.blk_cnt, nblk, l.flags, n_valid);
    if (launch_check("partition_sort prologue")) return 1;
    p.blk_coff = l.blk_cnt;
  }
  p.hist = l.hist;
  p.bases = l.bases;
  {
    KTimer _kt("ps_hist", st);
    if (use_sel)
    ps_hist<true><<<(unsigned)nblk, NT, 0, st>>>(p);
  else
    ps_hist<false><<<(unsigned)nblk, NT, 0, st>>>(p);
  }
  ps_bucket_scan<<<(unsigned)B, NT, 0, st>>>(l.hist, nblk, l.counts);
  Rpa rpa;
  for (int c = 0; c < MAXCLS; ++c) rpa.v[c] = 1;
  for (int c = 0; c < ncls; ++c) rpa.v[c] = rows_per_unit(rb_cls[c]);
  ps_finish<<<1, NT, 0, st>>>(l.counts, B, rpa, ncls, l.bases, (int64_t*)hdev, l.flags,
                              n_valid, (use_sel || p.needs_j) ? 1 : 0, given ? 1 : 0);
  if (launch_check("partition_sort")) return 1;
  hipEvent_t ev = stream_event(st);
  if (!ev) return fail("partition_sort: cannot create an event");
  if (hip_check(hipEventRecord(ev, st), "partition_sort")) return 1;
  {
    KTimer _kt("ps_scatter", st);
    if (use_sel)
    ps_scatter<true><<<(unsigned)nblk, NT, 0, st>>>(p);
  else
    ps_scatter<false><<<(unsigned)nblk, NT, 0, st>>>(p);
  }
  if (launch_check("partition_sort scatter")) return 1;
  // the call's one host wait: the counts, while the scatter runs
  if (hip_check(hipEventSynchronize(ev), "partition_sort")) return 1;
  return ok();
}

}  // namespace
}  // namespace pf

extern "C" {

int pf_partition_sort(const void* const* cols_in, void* const* cols_out,
                      const int32_t* row_bytes, const int32_t* comp, int32_t ncols,
                      int32_t key_idx, const uint8_t* mask, int64_t ndom, int64_t n_valid,
                      int64_t goff, int32_t B, int32_t select_first, int64_t* gidx_out,
                      int64_t* host_out, void* workspace, size_t workspace_bytes,
                      void* stream) {
  using namespace pf;
  if (!gidx_out) return fail("partition_sort: no gidx output");
  return partition_impl(cols_in, cols_out, row_bytes, comp, ncols, key_idx, false, nullptr,
                        mask, ndom, n_valid, goff, B, select_first, gidx_out, host_out,
                        workspace, workspace_bytes, (hipStream_t)stream);
}

size_t pf_partition_by_bucket_workspace_bytes(int64_t n, int32_t B) {
  if (B < 1) B = 1;
  if (B > pf::BMAX) B = pf::BMAX;
  return pf::carve_layout(nullptr, n, n, B, false, false).bytes;
}

int pf_partition_by_bucket(const void* const* cols_in, void* const* cols_out,
                           const int32_t* row_bytes, int32_t ncols, const int32_t* bucket,
                           int64_t n, int32_t B, int64_t* host_out, void* workspace,
                           size_t workspace_bytes, void* stream) {
  using namespace pf;
  if (!bucket && n > 0) return fail("partition_by_bucket: no bucket column");
  const int32_t comp[MAXC] = {0, 0, 0, 0, 0, 0};
  return partition_impl(cols_in, cols_out, row_bytes, comp, ncols, -1, true, bucket, nullptr, n,
                        n, 0, B, 0, nullptr, host_out, workspace, workspace_bytes,
                        (hipStream_t)stream);
}

}  // extern "C"
