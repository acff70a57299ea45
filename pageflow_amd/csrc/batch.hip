// Batched equi-joins: many independent inner joins of uneven size run as
// one pipeline of persistent kernels that pull work items from device
// queues, so a batch of 256 joins costs a handful of launches and no host
// wait; the steady state is captured once as a hipGraph and replayed.
//
// Each join j: int32 build keys (unique) and int32 probe keys, stored back
// to back in two buffers (offsets b_off / p_off).  Output: join j's matched
// (probe row, build row) pairs, join-local int64 row ids, in probe-row
// order, in out_probe/out_build at p_off[j] (the upper bound of a unique
// build's matches), their count in counts[j].
//
//   bj_build    items = (join, 4096 build rows): insert (key << 32 | row)
//               into the join's slice of one table arena with a 64-bit CAS
//               (empty = all ones, never a valid slot: a row id is < 2^31);
//               a repeated key keeps its lowest row (atomicMin)
//   bj_probe    items = (join, probe chunk): each probe row's build row or
//               -1 into partner, the item's matches into item_cnt
//   scan        item_cnt -> item_pos (items are join-major, rows ascending)
//   bj_compact  items again: ballot-ranked stable compaction of matches
//   bj_counts   per-join match counts
// A join whose table exceeds the probe slice size has no table in HBM: both
// sides are partitioned by the top bits of the table slot, and bj_probe_lds
// builds each slice's table in LDS and probes it there.
//
// Tunables per size class of the build (small < 32K rows, medium < 512K,
// large): the table's slots per build row (2 or 4, rounded to a power of
// two), the probe chunk (a power of two), the workgroup size of the probe
// and compaction kernels (256 or 512), and the probe slice size.
// pf_batch_plan_create takes them (the autotuner in pageflow_amd/batch.py
// measures and picks them).
//
// The plain baseline (mode 1) runs the same kernels one join at a time,
// with the reference's one host wait per join (kernel_utils.cuh:37-40).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <type_traits>
#include <vector>

#include "pf_common.hpp"
#include "../../include/pageflow_join.h"

namespace pf {

namespace {

constexpr uint64_t kEmpty = ~0ULL;
constexpr int BUILD_ROWS = 4096;
constexpr int NCLASS = 3;

struct Tune {
  int load;      // slots per build row (2 or 4)
  int chunk;     // probe rows an item
  int nt;        // probe / compact workgroup size (256 or 512)
  int slice_kb;  // 0, or the table bytes (KiB) of a probe slice (sliced probe in LDS)
};

int class_of(int64_t nb) { return nb < 32768 ? 0 : (nb < 524288 ? 1 : 2); }

constexpr int NXQ = 8;            // queues (XCDs)

struct SliceQueues {
  int32_t lo[NXQ], hi[NXQ];   // each queue's (join, slice) pairs in the pair lists
};

struct Plan {
  int nj = 0;
  std::vector<int64_t> nb, np, b_off, p_off, t_off, cap;
  std::vector<int> cls;
  Tune tune[NCLASS];
  // items: build items all joins; probe items grouped by workgroup size
  int64_t n_bitems = 0;
  int64_t n_pitems[2] = {0, 0};      // [0]: nt 256, [1]: nt 512
  int64_t pitem_base[2] = {0, 0};    // first item of each group in the item arrays
  int64_t n_pitems_plain[2] = {0, 0};  // each group's items of joins probed whole (first)
  int64_t n_pitems_all = 0;
  int64_t total_slots = 0, total_b = 0, total_p = 0;
  int64_t plain_slots = 0;  // the tables of joins built whole come first in the arena
  std::vector<int64_t> first_item, last_item;  // per join, in the probe item arrays
  // device
  void* dmem = nullptr;
  size_t dbytes = 0;
  int64_t *d_b_off, *d_p_off, *d_t_off;
  uint32_t* d_mask;
  int32_t* d_bitem_join;
  int64_t* d_bitem_start;
  int32_t *d_pitem_join, *d_pitem_len;
  int64_t* d_pitem_start;
  int64_t *d_item_cnt, *d_item_pos;
  int64_t *d_first, *d_last;
  int32_t* d_partner;
  uint64_t* d_table;
  int32_t* d_queue;  // [0] build, [1] probe 256, [2] probe 512, [3] compact 256, [4] compact 512
  void* d_scan;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int persist = 1024;  // workgroups of a persistent launch
  // sliced probe: slice bits per join (0: probed whole), the (join, slice)
  // pairs of each queue, each join's pairs in each queue (per-join runs)
  std::vector<int> sbits, log2cap;
  std::vector<int64_t> soff_base;
  int64_t soff_total = 0;
  bool any_sliced = false;
  std::vector<int32_t> pair_join, pair_slice;
  SliceQueues qall{};
  std::vector<SliceQueues> qjoin;
  int32_t *d_sliced = nullptr, *d_chunk = nullptr, *d_pair_join = nullptr,
          *d_pair_slice = nullptr, *d_xstart = nullptr, *d_pk_part = nullptr,
          *d_pv_part = nullptr;
  int64_t *d_soff_base = nullptr, *d_soff = nullptr;
  // the sliced joins' probe keys as segments of one partition: all of them
  // (batched runs) and each alone (per-join runs)
  std::vector<RpSeg> segs, segs1;
  std::vector<int32_t> seg_join;
  std::vector<int> seg_of;  // per join: its segment, or -1
  int64_t seg_blocks = 0, seg_entries = 0;
  RpSeg *d_segs = nullptr, *d_segs1 = nullptr;
  int32_t *d_seg_join = nullptr, *d_shist = nullptr;
  int64_t* d_soffs = nullptr;
  void* d_sscan = nullptr;
  // the sliced joins' build rows as segments, partitioned by the same
  // slices as their probe rows (each probe item builds its LDS table from
  // its slice's build rows), and the build slice offsets
  std::vector<RpSeg> bsegs, bsegs1;
  int64_t bseg_blocks = 0, bseg_entries = 0;
  RpSeg *d_bsegs = nullptr, *d_bsegs1 = nullptr;
  int32_t *d_bk_part = nullptr, *d_bv_part = nullptr;
  int64_t* d_bsoff = nullptr;
  std::vector<int64_t> bsoff_base;
  int64_t bsoff_total = 0;
  int64_t* d_bsoff_base = nullptr;
  int64_t* d_overflow = nullptr;  // LDS tables that took more than one pass
  int64_t n_bitems_plain = 0;    // build items of joins built whole (listed first)
  std::vector<int64_t> bitem_first, bitem_n;  // per join: its build items
};

// A work item taken from the queue, as a wave-uniform (scalar) value.  The
// persistent loops below exit on it and hold barriers; the compiler must see
// the exit as uniform, otherwise it may nest the loop so that the barriers
// and the read of the next item run in an inner loop that the lane fetching
// items (thread 0) never re-enters, and the other lanes spin on a stale item.
__device__ __forceinline__ int64_t uniform_item(int64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)v >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// A key's home slot in a join's table of mask + 1 slots: the low bits of
// fmix32(key) (the slice partition takes the top bits of the same slot).
__device__ __forceinline__ uint32_t bslot(int32_t key, uint32_t mask) {
  return fmix32((uint32_t)key) & mask;
}

__global__ void bj_fill_tables(uint64_t* table, int64_t n, int32_t* queue, int nq) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t t0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  for (int64_t i = t0; i < n; i += stride) table[i] = kEmpty;
  if (t0 < nq) queue[t0] = 0;
}

// Build items: (join, first row) with BUILD_ROWS rows (fewer at a join's end).
__global__ __launch_bounds__(256) void bj_build(const int32_t* __restrict__ bkeys,
                                                const int64_t* __restrict__ b_off,
                                                const int64_t* __restrict__ t_off,
                                                const uint32_t* __restrict__ mask,
                                                const int32_t* __restrict__ item_join,
                                                const int64_t* __restrict__ item_start,
                                                int64_t n_items, int64_t item_lo,
                                                uint64_t* table, int32_t* queue) {
  __shared__ int64_t s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = item_lo + atomicAdd(queue, 1);
    __syncthreads();
    const int64_t item = uniform_item(s_item);
    __syncthreads();
    if (item >= n_items) return;
    const int j = item_join[item];
    const int64_t r0 = item_start[item];
    const int64_t nrows = b_off[j + 1] - b_off[j];
    const int64_t r1 = r0 + BUILD_ROWS < nrows ? r0 + BUILD_ROWS : nrows;
    uint64_t* t = table + t_off[j];
    const uint32_t m = mask[j];
    // U rows a thread with their CASes in flight together (one at a time is
    // latency-bound).  No read before the CAS: in a fresh table of unique
    // keys the first slot is usually empty and the CAS returns what a read
    // would have.
    constexpr int U = 4;
    for (int64_t rr = r0 + threadIdx.x; rr < r1; rr += 256 * U) {
      int32_t k[U];
      uint32_t s[U];
      bool todo[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int64_t r = rr + u * 256;
        todo[u] = r < r1;
        k[u] = todo[u] ? bkeys[b_off[j] + r] : 0;
        s[u] = bslot(k[u], m);
      }
      bool any = true;
      while (any) {
        uint64_t prev[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const uint64_t v = ((uint64_t)(uint32_t)k[u] << 32) | (uint32_t)(rr + u * 256);
          prev[u] = todo[u] ? atomicCAS((unsigned long long*)&t[s[u]], (unsigned long long)kEmpty,
                                        (unsigned long long)v)
                            : kEmpty;
        }
        any = false;
#pragma unroll
        for (int u = 0; u < U; ++u) {
          if (!todo[u]) continue;
          const uint64_t v = ((uint64_t)(uint32_t)k[u] << 32) | (uint32_t)(rr + u * 256);
          if (prev[u] == kEmpty) {
            todo[u] = false;
          } else if ((uint32_t)(prev[u] >> 32) == (uint32_t)k[u]) {
            atomicMin((unsigned long long*)&t[s[u]], (unsigned long long)v);
            todo[u] = false;
          } else {
            s[u] = (s[u] + 1) & m;
            any = true;
          }
        }
      }
    }
  }
}

template <int NT>
__global__ __launch_bounds__(NT) void bj_probe(const int32_t* __restrict__ pkeys,
                                               const int64_t* __restrict__ p_off,
                                               const int64_t* __restrict__ t_off,
                                               const uint32_t* __restrict__ mask,
                                               const int32_t* __restrict__ item_join,
                                               const int64_t* __restrict__ item_start,
                                               const int32_t* __restrict__ item_len,
                                               int64_t n_items, int64_t item_lo,
                                               const uint64_t* __restrict__ table,
                                               const int32_t* __restrict__ sliced,
                                               int32_t* __restrict__ partner,
                                               int64_t* __restrict__ item_cnt, int32_t* queue) {
  __shared__ int64_t s_item;
  __shared__ int red[NT / kWave];
  // one thread-0 region an iteration (the previous item's count, then the
  // next item), as in the other persistent loops
  int64_t prev = -1;
  int prev_tot = 0;
  for (;;) {
    if (threadIdx.x == 0) {
      if (prev >= 0) item_cnt[prev] = prev_tot;
      s_item = item_lo + atomicAdd(queue, 1);
    }
    __syncthreads();
    const int64_t item = uniform_item(s_item);
    __syncthreads();
    if (item >= item_lo + n_items) return;
    const int j = item_join[item];
    if (sliced[j]) {  // probed by bj_probe_lds, which also counts the item
      prev = -1;
      continue;
    }
    const int64_t base = p_off[j] + item_start[item];
    const int len = item_len[item];
    const uint64_t* t = table + t_off[j];
    const uint32_t m = mask[j];
    int c = 0;
    // U probe rows a thread with their first-slot reads in flight together;
    // the rare longer probe sequences are resolved afterwards
    constexpr int U = 8;
    for (int r0 = threadIdx.x; r0 < len; r0 += NT * U) {
      int32_t k[U];
      uint32_t s[U];
      uint64_t v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int r = r0 + u * NT;
        k[u] = r < len ? pkeys[base + r] : 0;
        s[u] = bslot(k[u], m);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) v[u] = r0 + u * NT < len ? t[s[u]] : kEmpty;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int r = r0 + u * NT;
        if (r >= len) break;
        while (v[u] != kEmpty && (uint32_t)(v[u] >> 32) != (uint32_t)k[u]) {
          s[u] = (s[u] + 1) & m;
          v[u] = t[s[u]];
        }
        const int32_t hit = v[u] == kEmpty ? -1 : (int32_t)(uint32_t)v[u];
        partner[base + r] = hit;
        c += hit >= 0;
      }
    }
    c = wave_sum(c);
    if (lane_id() == 0) red[threadIdx.x / kWave] = c;
    __syncthreads();
    prev_tot = 0;
#pragma unroll
    for (int w = 0; w < NT / kWave; ++w) prev_tot += red[w];
    prev = item;
  }
}

template <int NT>
__global__ __launch_bounds__(NT) void bj_compact(const int64_t* __restrict__ p_off,
                                                 const int32_t* __restrict__ item_join,
                                                 const int64_t* __restrict__ item_start,
                                                 const int32_t* __restrict__ item_len,
                                                 int64_t n_items, int64_t item_lo,
                                                 const int64_t* __restrict__ first,
                                                 const int32_t* __restrict__ partner,
                                                 const int32_t* __restrict__ sliced,
                                                 const int32_t* __restrict__ rows_part,
                                                 const int64_t* __restrict__ item_pos,
                                                 int64_t* __restrict__ out_probe,
                                                 int64_t* __restrict__ out_build, int32_t* queue) {
  // U rows a thread a round (row r0 + u * NT + thread): one barrier pair
  // a round of NT * U rows
  constexpr int W = NT / kWave, U = 4;
  __shared__ int64_t s_item;
  __shared__ int cnt[U][W];
  const int w = threadIdx.x / kWave, lane = lane_id();
  for (;;) {
    if (threadIdx.x == 0) s_item = item_lo + atomicAdd(queue, 1);
    __syncthreads();
    const int64_t item = uniform_item(s_item);
    __syncthreads();
    if (item >= item_lo + n_items) return;
    const int j = item_join[item];
    const int64_t r00 = item_start[item];
    const int64_t base = p_off[j] + r00;
    const int len = item_len[item];
    // a sliced join's partner is in partition order: its probe rows come
    // from the partition's row ids
    const int32_t* rmap = sliced[j] ? rows_part + base : nullptr;
    int64_t o = p_off[j] + item_pos[item] - item_pos[first[j]];
    for (int r0 = 0; r0 < len; r0 += NT * U) {
      int32_t pr[U];
      uint64_t bal[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int r = r0 + u * NT + (int)threadIdx.x;
        pr[u] = r < len ? partner[base + r] : -1;
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        bal[u] = ballot(pr[u] >= 0);
        if (lane == 0) cnt[u][w] = __popcll(bal[u]);
      }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < U; ++u) {
        int pre = 0, tot = 0;
#pragma unroll
        for (int q = 0; q < W; ++q) {
          const int c = cnt[u][q];
          pre += q < w ? c : 0;
          tot += c;
        }
        if (pr[u] >= 0) {
          const int r = r0 + u * NT + (int)threadIdx.x;
          const int64_t pos = o + pre + mbcnt(bal[u]);
          out_probe[pos] = rmap ? (int64_t)rmap[r] : r00 + r;
          out_build[pos] = (int64_t)pr[u];
        }
        o += tot;
      }
      __syncthreads();
    }
  }
}

// ------------------------------------------------------ sliced probe --
//
// A large join's table spans far more than an XCD's L2, so random probes
// of a global table would be bound by 64-byte memory requests.  Both sides
// of such a join are instead partitioned (stably, one pass) by the top bits
// of the key's table slot into slices, and each work item builds the table
// of one sub-range of 2^14 slots of its slice in LDS, from the slice's
// build rows whose slot falls in it, then answers its probe rows' lookups
// there.  Items are (join, slice, probe chunk, sub-range): a slice larger
// than 2^14 slots is taken by several items, each probing the chunk's rows
// whose slot falls in its sub-range.  An LDS table holds 7/8 of its slots;
// a sub-range with more build rows (keys clustered by their hash) is joined
// in several passes of build rows in slice order, a probe row keeping the
// first (lowest) build row that answers it.  The partition writes each key
// as fmix32(key), a bijection: a slot is the low bits of that value and the
// LDS tables hold and compare it in place of the key, so no lookup hashes.
//
// Items of slice s of join j go to queue (s + j) % 8, and workgroup b starts
// on queue b % 8: workgroups are dealt to the eight XCDs round robin, so the
// items of one slice (its chunks and sub-ranges) share an XCD's L2.  A
// workgroup whose queue runs dry takes items from the next queues, so every
// item is probed once whatever the placement.

constexpr int PL_NT = 1024;
constexpr int PL_LG = 14;                // log2 of the slots of an item's LDS table
constexpr int PL_SR = 1 << PL_LG;
constexpr int PL_ROWS = 1 << 17;         // probe rows a sliced item (one chunk)

// log2 of the sub-ranges of join j's slices
__device__ __forceinline__ int sub_log(uint32_t m, int sbits) {
  const int slog = 32 - __clz(m) - sbits;  // log2(slice slots); m = cap - 1
  return slog > PL_LG ? slog - PL_LG : 0;
}

// Block x: each pair's items and their exclusive scan over the queue's
// pairs [lo, hi) into xstart[lo..hi] (xstart[hi] = the queue's items).
__global__ __launch_bounds__(1024) void bj_slice_scan(const int32_t* __restrict__ pair_join,
                                                      const int32_t* __restrict__ pair_slice,
                                                      const int64_t* __restrict__ soff_base,
                                                      const int64_t* __restrict__ soff,
                                                      const uint32_t* __restrict__ mask,
                                                      const int32_t* __restrict__ sbits,
                                                      const SliceQueues q,
                                                      int32_t* __restrict__ xstart) {
  __shared__ int32_t wsum[1024 / kWave];
  __shared__ int32_t carry;
  const int x = blockIdx.x;
  const int lo = q.lo[x], hi = q.hi[x];
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int p0 = lo; p0 < hi; p0 += 1024) {
    const int p = p0 + (int)threadIdx.x;
    int nit = 0;
    if (p < hi) {
      const int j = pair_join[p];
      const int64_t* so = soff + soff_base[j];
      const int sl = pair_slice[p];
      const int nch = (int)((so[sl + 1] - so[sl] + PL_ROWS - 1) / PL_ROWS);
      nit = nch << sub_log(mask[j], sbits[j]);
    }
    const int incl = wave_incl_scan(nit);
    if (lane_id() == kWave - 1) wsum[threadIdx.x / kWave] = incl;
    __syncthreads();
    int before = carry;
    for (int w = 0; w < (int)threadIdx.x / kWave; ++w) before += wsum[w];
    if (p < hi) xstart[p] = before + incl - nit;
    __syncthreads();
    if (threadIdx.x == 1023) carry = before + incl;
    __syncthreads();
  }
  if (threadIdx.x == 0) xstart[hi] = carry;
}

// (key, row) into an LDS table of msk + 1 slots from slot pos, linear
// probing; a repeated key keeps its lowest row.  The table has an empty slot.
__device__ __forceinline__ void lds_insert(uint64_t* T, uint32_t pos, uint32_t msk, int32_t k,
                                           int32_t row) {
  const uint64_t v = ((uint64_t)(uint32_t)k << 32) | (uint32_t)row;
  for (;;) {
    const uint64_t prev = atomicCAS((unsigned long long*)&T[pos], (unsigned long long)kEmpty,
                                    (unsigned long long)v);
    if (prev == kEmpty) return;
    if ((uint32_t)(prev >> 32) == (uint32_t)k) {
      atomicMin((unsigned long long*)&T[pos], (unsigned long long)v);
      return;
    }
    pos = (pos + 1) & msk;
  }
}

// A lookup of key k in the aligned slot pair e of an LDS table, starting at
// its second slot when `odd`: the matching slot's row, -1 when an empty slot
// ends the probe, or kOpen when both slots are taken by other keys.  An
// empty slot's row half is all ones (rows are non-negative); a slot before
// the key's home cannot hold it (the table is never full, so no probe wraps
// all the way round).
constexpr int kOpen = -2;
__device__ __forceinline__ int pair_lookup(uint64_t e0, uint64_t e1, uint32_t odd, int32_t k) {
  // branch-free: both slots are tested, the first that decides wins
  const uint32_t r0 = (uint32_t)e0, r1 = (uint32_t)e1;
  const bool z0 = !odd && r0 == ~0u, z1 = r1 == ~0u;
  const bool m0 = !odd && r0 != ~0u && (uint32_t)(e0 >> 32) == (uint32_t)k;
  const bool m1 = !z1 && (uint32_t)(e1 >> 32) == (uint32_t)k;
  int r = z1 ? -1 : kOpen;
  r = m1 ? (int)r1 : r;
  r = z0 ? -1 : r;
  return m0 ? (int)r0 : r;
}

__global__ __launch_bounds__(PL_NT) void bj_probe_lds(
    const int32_t* __restrict__ pk_part, const int64_t* __restrict__ p_off,
    const int32_t* __restrict__ bk_part, const int32_t* __restrict__ bv_part,
    const int64_t* __restrict__ b_off, const int64_t* __restrict__ bsoff_base,
    const int64_t* __restrict__ bsoff, const uint32_t* __restrict__ mask,
    const int32_t* __restrict__ sbits, const int32_t* __restrict__ pair_join,
    const int32_t* __restrict__ pair_slice, const int64_t* __restrict__ soff_base,
    const int64_t* __restrict__ soff, const int32_t* __restrict__ xstart, const SliceQueues q,
    const int64_t* __restrict__ first, const int32_t* __restrict__ lgchunk,
    int32_t* __restrict__ partner, int64_t* __restrict__ item_cnt, int64_t* overflow,
    int32_t* xqueue) {
  __shared__ __align__(16) uint64_t T[PL_SR];
  ulonglong2* const T2 = reinterpret_cast<ulonglong2*>(T);  // aligned slot pairs
  // matches per compaction item (chunks of 2^lgchunk partitioned rows, at
  // least 256) that this item's rows fall in
  constexpr int NC = PL_ROWS / 256 + 1;
  constexpr int W = PL_NT / kWave;
  __shared__ int32_t ccnt[NC];
  __shared__ int32_t s_red[W];
  __shared__ int64_t s_w[2];  // item, pair (or -1: this queue is done)
  const int tid = threadIdx.x, lane = lane_id(), w = tid / kWave;
  int qi = 0;
  int x = (int)(blockIdx.x % NXQ);
  int pre = -1;  // thread 0: the next item of queue x, taken during this item's lookups
  for (;;) {
    if (w == 0) {  // wave 0 fetches the next item of queue x
      const int lo = q.lo[x], hi = q.hi[x];
      const int n = xstart[hi];
      int it = 0;
      if (lane == 0) it = pre >= 0 ? pre : atomicAdd(&xqueue[x], 1);
      pre = -1;
      it = __shfl(it, 0, kWave);
      int p = -1;
      if (it < n) {
        // the pair holding item `it`: the last p in [lo, hi) with
        // xstart[p] <= it (xstart ascends; xstart[lo] <= it < xstart[hi]),
        // narrowed 64 ways a step: lane l tests a + l * step
        int a = lo, b = hi;
        while (b - a > 1) {
          const int step = (b - a + kWave - 1) / kWave;
          const int c = a + lane * step;
          const uint64_t le = ballot(c < b && xstart[c] <= it);
          a += (63 - (int)__clzll(le)) * step;  // lane 0 (c = a) is always in
          b = b < a + step ? b : a + step;
        }
        p = a;
      }
      if (lane == 0) {
        s_w[0] = it;
        s_w[1] = p;
      }
    }
    __syncthreads();  // also: the previous item's lookups and counts are done
    const int64_t it = uniform_item(s_w[0]);
    const int64_t p = uniform_item(s_w[1]);
    __syncthreads();
    if (p < 0) {  // this queue is done: the next one, until all eight are
      if (++qi == NXQ) return;
      x = (x + 1) % NXQ;
      continue;
    }
    const int j = pair_join[p];
    const int sl = pair_slice[p];
    const uint32_t m = mask[j];
    const int sb = sbits[j];
    const int lsub = sub_log(m, sb);
    const int slog = 32 - __clz(m) - sb;
    const int SR = 1 << (slog < PL_LG ? slog : PL_LG);
    const int cap = SR - SR / 8;  // build rows a pass holds
    const int idx = (int)(it - xstart[p]);
    const int sub = idx & ((1 << lsub) - 1);
    const int64_t* so = soff + soff_base[j];
    const int64_t r0 = so[sl] + (int64_t)(idx >> lsub) * PL_ROWS;
    const int n = __builtin_amdgcn_readfirstlane(
        (int)(r0 + PL_ROWS < so[sl + 1] ? PL_ROWS : so[sl + 1] - r0));
    const uint32_t base = ((uint32_t)sl << slog) + ((uint32_t)sub << PL_LG);
    // the slice's build rows
    const int64_t* bso = bsoff + bsoff_base[j];
    const int32_t* __restrict__ bk = bk_part + b_off[j] + bso[sl];
    const int32_t* __restrict__ bv = bv_part + b_off[j] + bso[sl];
    const int nbs = __builtin_amdgcn_readfirstlane((int)(bso[sl + 1] - bso[sl]));
    // the build rows of this sub-range (all of the slice's when it is one),
    // counted with BU loads a thread in flight together
    constexpr int BU = 16;
    int nin = nbs;
    if (lsub > 0) {
      int c = 0;
      for (int rb = 0; rb < nbs; rb += PL_NT * BU) {
        int32_t t[BU];
#pragma unroll
        for (int u = 0; u < BU; ++u) {
          const int r = rb + tid + u * PL_NT;
          t[u] = r < nbs ? bk[r] : 0;
        }
#pragma unroll
        for (int u = 0; u < BU; ++u)
          c += rb + tid + u * PL_NT < nbs && (((uint32_t)t[u] & m) - base) < (uint32_t)SR;
      }
      c = wave_sum(c);
      if (lane == 0) s_red[w] = c;
      __syncthreads();
      nin = 0;
#pragma unroll
      for (int v = 0; v < W; ++v) nin += s_red[v];
      nin = __builtin_amdgcn_readfirstlane(nin);
      __syncthreads();
    }
    const int npass = nin <= cap ? 1 : (nin + cap - 1) / cap;
    if (npass > 1 && tid == 0) atomicAdd((unsigned long long*)overflow, 1ULL);
    const int lg = lgchunk[j];
    const int64_t c0 = r0 >> lg;  // the first compaction item this item touches
    const int nc = (int)(((r0 + n - 1) >> lg) - c0) + 1;
    for (int i = tid; i < nc; i += PL_NT) ccnt[i] = 0;
    const int32_t* __restrict__ pk = pk_part + p_off[j] + r0;
    int32_t* __restrict__ pt = partner + p_off[j] + r0;
    const int coff = (int)(r0 & ((1 << lg) - 1));  // r0's offset in its compaction item
    // the next item's index: its atomic returns while this item is probed
    if (tid == 0) pre = atomicAdd(&xqueue[x], 1);
    for (int pass = 0; pass < npass; ++pass) {
      for (int i = tid; i < SR; i += PL_NT) T[i] = kEmpty;
      __syncthreads();
      if (npass == 1) {  // BU rows a thread loaded, then inserted
        for (int rb = 0; rb < nbs; rb += PL_NT * BU) {
          int32_t tk[BU], tv[BU];
#pragma unroll
          for (int u = 0; u < BU; ++u) {
            const int r = rb + tid + u * PL_NT;
            tk[u] = r < nbs ? bk[r] : 0;
            tv[u] = r < nbs ? bv[r] : 0;
          }
#pragma unroll
          for (int u = 0; u < BU; ++u) {
            const uint32_t pos = ((uint32_t)tk[u] & m) - base;
            if (rb + tid + u * PL_NT < nbs && pos < (uint32_t)SR)
              lds_insert(T, pos, SR - 1, tk[u], tv[u]);
          }
        }
      } else {
        // pass `pass` takes the sub-range's build rows of rank [pass * cap,
        // (pass + 1) * cap) in slice order
        int before = 0;
        for (int rb = 0; rb < nbs; rb += PL_NT) {
          const int r = rb + tid;
          const int32_t k = r < nbs ? bk[r] : 0;
          const uint32_t pos = ((uint32_t)k & m) - base;
          const bool in = r < nbs && pos < (uint32_t)SR;
          const uint64_t bal = ballot(in);
          if (lane == 0) s_red[w] = (int)__popcll(bal);
          __syncthreads();
          int below = 0, tot = 0;
#pragma unroll
          for (int v = 0; v < W; ++v) {
            const int c = s_red[v];
            below += v < w ? c : 0;
            tot += c;
          }
          const int rank = before + below + mbcnt(bal);
          if (in && rank >= pass * cap && rank < (pass + 1) * cap)
            lds_insert(T, pos, SR - 1, k, bv[r]);
          before += tot;
          __syncthreads();
        }
      }
      __syncthreads();
      // U rows a thread a round; the next round's keys are loaded while this
      // round's are looked up.  A lookup reads its home slot's aligned pair
      // of slots (16 bytes); the U first reads of a round are in flight
      // together.  Matches are counted per wave into the compaction item
      // of its rows (scalar sums, flushed to ccnt when the item changes).
      constexpr int U = 8;
      int acc_c = -1, acc_n = 0;  // wave-uniform: an item and its uncounted matches
      auto count = [&](int c, int cnt) {
        if (cnt == 0) return;
        if (c != acc_c) {
          if (acc_n > 0 && lane == 0) atomicAdd(&ccnt[acc_c], acc_n);
          acc_c = c;
          acc_n = 0;
        }
        acc_n += cnt;
      };
      auto sweep = [&](auto first_pass) {
        constexpr bool FIRST = decltype(first_pass)::value;
        int32_t k[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int i = tid + u * PL_NT;
          k[u] = pk[i < n ? i : n - 1];
        }
        for (int i0 = tid; i0 < n; i0 += PL_NT * U) {
          int32_t kn[U];
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const int i = i0 + (U + u) * PL_NT;
            kn[u] = pk[i < n ? i : n - 1];
          }
          uint32_t pos[U];
          bool mine[U];
          uint64_t g0[U], g1[U];
#pragma unroll
          for (int u = 0; u < U; ++u) {
            pos[u] = ((uint32_t)k[u] & m) - base;
            mine[u] = i0 + u * PL_NT < n && pos[u] < (uint32_t)SR;
          }
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const ulonglong2 t = T2[(mine[u] ? pos[u] : 0u) >> 1];
            g0[u] = t.x;
            g1[u] = t.y;
          }
          int32_t hit[U];  // a match new in this pass, or -1
#pragma unroll
          for (int u = 0; u < U; ++u) {
            int r = pair_lookup(g0[u], g1[u], pos[u] & 1u, k[u]);
            r = mine[u] ? r : -1;
            uint32_t q = pos[u];
            while (r == kOpen) {
              q = ((q | 1u) + 1u) & (uint32_t)(SR - 1);
              const ulonglong2 t = T2[q >> 1];
              r = pair_lookup(t.x, t.y, 0u, k[u]);
            }
            hit[u] = r;
          }
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const int i = i0 + u * PL_NT;
            if (FIRST) {
              if (mine[u]) pt[i] = hit[u];
            } else if (hit[u] >= 0) {
              if (pt[i] < 0)  // this thread's own earlier write
                pt[i] = hit[u];
              else
                hit[u] = -1;
            }
          }
          // the wave's rows of step u are [first, first + 64): in one
          // compaction item, or two when an item starts at lane `split`
          const int wfirst = __builtin_amdgcn_readfirstlane(i0 - lane) + coff;
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const uint64_t bh = ballot(hit[u] >= 0);
            if (bh) {
              const int first = wfirst + u * PL_NT;
              const int cw = first >> lg;
              const int split = ((cw + 1) << lg) - first;
              if (split >= kWave) {
                count(cw, (int)__popcll(bh));
              } else {
                const uint64_t lo = bh & ((1ULL << split) - 1);
                count(cw, (int)__popcll(lo));
                count(cw + 1, (int)__popcll(bh ^ lo));
              }
            }
          }
#pragma unroll
          for (int u = 0; u < U; ++u) k[u] = kn[u];
        }
      };
      if (pass == 0)
        sweep(std::true_type{});
      else
        sweep(std::false_type{});
      if (acc_n > 0 && lane == 0) atomicAdd(&ccnt[acc_c], acc_n);
      __syncthreads();  // the table is rebuilt by the next pass
    }
    for (int i = tid; i < nc; i += PL_NT)
      if (ccnt[i]) atomicAdd((unsigned long long*)&item_cnt[first[j] + c0 + i],
                             (unsigned long long)ccnt[i]);
  }
}

__global__ void bj_counts(const int64_t* __restrict__ item_pos, const int64_t* __restrict__ first,
                          const int64_t* __restrict__ last, int nj, int64_t* __restrict__ counts) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < nj) counts[j] = first[j] < 0 ? 0 : item_pos[last[j] + 1] - item_pos[first[j]];
}

int64_t pow2_at_least(int64_t x) {
  int64_t c = 16;
  while (c < x) c <<= 1;
  return c;
}

int plan_build(Plan& P, int nj, const int64_t* nb, const int64_t* np, const int32_t* tune) {
  P.nj = nj;
  for (int c = 0; c < NCLASS; ++c) {
    P.tune[c] = Tune{2, c == 0 ? 4096 : 16384, c == 2 ? 512 : 256, c == 2 ? 128 : 0};
    if (tune) {
      P.tune[c].load = tune[c * 4 + 0];
      P.tune[c].chunk = tune[c * 4 + 1];
      P.tune[c].nt = tune[c * 4 + 2];
      P.tune[c].slice_kb = tune[c * 4 + 3];
    }
    const Tune& t = P.tune[c];
    if ((t.load != 2 && t.load != 4) || t.chunk < 256 || t.chunk > (1 << 20) ||
        (t.chunk & (t.chunk - 1)) || (t.nt != 256 && t.nt != 512) ||
        (t.slice_kb != 0 && (t.slice_kb < 64 || t.slice_kb > (1 << 20))))
      return fail("batch plan: tunables are load 2|4, chunk a power of two in [256, 2^20], "
                  "nt 256|512, slice KiB 0 or in [64, 2^20]");
  }
  P.nb.assign(nb, nb + nj);
  P.np.assign(np, np + nj);
  P.b_off.assign(nj + 1, 0);
  P.p_off.assign(nj + 1, 0);
  P.t_off.assign(nj + 1, 0);
  P.cap.assign(nj, 0);
  P.cls.assign(nj, 0);
  for (int j = 0; j < nj; ++j) {
    if (nb[j] < 0 || np[j] < 0 || nb[j] >= (1LL << 31) || np[j] >= (1LL << 31))
      return fail("batch plan: a join side takes fewer than 2^31 rows");
    P.cls[j] = class_of(nb[j]);
    P.cap[j] = pow2_at_least(P.tune[P.cls[j]].load * std::max<int64_t>(nb[j], 1));
    P.b_off[j + 1] = P.b_off[j] + nb[j];
    P.p_off[j + 1] = P.p_off[j] + np[j];
  }
  P.total_b = P.b_off[nj];
  P.total_p = P.p_off[nj];
  // slices: 2^sbits of a table, each about slice_kb KiB of slots
  P.sbits.assign(nj, 0);
  P.log2cap.assign(nj, 0);
  P.soff_base.assign(nj, 0);
  for (int j = 0; j < nj; ++j) {
    int l = 0;
    while ((1LL << l) < P.cap[j]) ++l;
    P.log2cap[j] = l;
    const int64_t kb = P.tune[P.cls[j]].slice_kb;
    int sb = 0;
    if (kb > 0 && np[j] > 0 && nb[j] > 0)
      while (sb < kSegMaxBits && sb < l && (P.cap[j] * 8 >> sb) > kb * 1024) ++sb;
    P.sbits[j] = sb;
    P.soff_base[j] = P.soff_total;
    if (sb > 0) {
      P.soff_total += (1 << sb) + 1;
      P.any_sliced = true;
    }
  }
  // the table arena holds the joins probed whole (filled empty before a
  // run); a sliced join's table lives only in LDS, a sub-range an item
  for (int j = 0; j < nj; ++j) {
    P.t_off[j] = P.total_slots;
    if (P.sbits[j] == 0) P.total_slots += P.cap[j];
  }
  P.plain_slots = P.total_slots;
  P.t_off[nj] = P.total_slots;
  // segments: the sliced joins in join order
  P.seg_of.assign(nj, -1);
  int64_t cum = 0;
  for (int j = 0; j < nj; ++j) {
    if (P.sbits[j] == 0) continue;
    RpSeg g{};
    g.in_off = P.p_off[j];
    g.n = np[j];
    g.nchunks = (np[j] + kSegChunkRows - 1) / kSegChunkRows;
    g.bits = P.sbits[j];
    g.shift = 64 - P.log2cap[j];
    RpSeg one = g;
    g.blk_start = P.seg_blocks;
    g.hist_off = P.seg_entries;
    g.delta = P.p_off[j] - cum;
    one.blk_start = 0;
    one.hist_off = 0;
    one.delta = P.p_off[j];
    P.seg_of[j] = (int)P.segs.size();
    P.segs.push_back(g);
    P.segs1.push_back(one);
    P.seg_join.push_back(j);
    P.seg_blocks += g.nchunks;
    P.seg_entries += g.nchunks << g.bits;
    cum += np[j];
  }
  P.bsoff_base.assign(nj, 0);
  for (int j = 0; j < nj; ++j) {
    if (P.sbits[j] == 0) continue;
    P.bsoff_base[j] = P.bsoff_total;
    P.bsoff_total += (1 << P.sbits[j]) + 1;
  }
  int64_t bcum = 0;
  for (int j = 0; j < nj; ++j) {
    if (P.sbits[j] == 0) continue;
    RpSeg g{};
    g.in_off = P.b_off[j];
    g.n = nb[j];
    g.nchunks = (nb[j] + kSegChunkRows - 1) / kSegChunkRows;
    g.bits = P.sbits[j];
    g.shift = 64 - P.log2cap[j];
    RpSeg one = g;
    g.blk_start = P.bseg_blocks;
    g.hist_off = P.bseg_entries;
    g.delta = P.b_off[j] - bcum;
    one.blk_start = 0;
    one.hist_off = 0;
    one.delta = P.b_off[j];
    P.bsegs.push_back(g);
    P.bsegs1.push_back(one);
    P.bseg_blocks += g.nchunks;
    P.bseg_entries += g.nchunks << g.bits;
    bcum += nb[j];
  }
  // the (join, slice) pairs of each queue, join-major, a one-entry gap after
  // each queue (where the queue's item total is written)
  P.qjoin.assign(nj, SliceQueues{});
  for (int x = 0; x < NXQ; ++x) {
    P.qall.lo[x] = (int32_t)P.pair_join.size();
    for (int j = 0; j < nj; ++j) {
      P.qjoin[j].lo[x] = (int32_t)P.pair_join.size();
      for (int sl = 0; sl < (1 << P.sbits[j]) && P.sbits[j] > 0; ++sl)
        if ((sl + j) % NXQ == x) {
          P.pair_join.push_back(j);
          P.pair_slice.push_back(sl);
        }
      P.qjoin[j].hi[x] = (int32_t)P.pair_join.size();
    }
    P.qall.hi[x] = (int32_t)P.pair_join.size();
    P.pair_join.push_back(-1);  // the gap
    P.pair_slice.push_back(-1);
  }
  return 0;
}

}  // namespace
}  // namespace pf

extern "C" {

void* pf_batch_plan_create(int32_t njoins, const int64_t* nb, const int64_t* np,
                           const int32_t* tune) {
  using namespace pf;
  if (njoins < 1) {
    fail("batch plan: at least one join");
    return nullptr;
  }
  Plan* P = new Plan();
  if (plan_build(*P, njoins, nb, np, tune)) {
    delete P;
    return nullptr;
  }
  // host item lists
  std::vector<int32_t> bj, pj[2], pl[2];
  std::vector<int64_t> bs, ps[2];
  // the sliced joins' build items last: bj_build takes the others (the
  // sliced joins' tables are built in LDS, a slice a workgroup)
  P->bitem_first.assign(njoins, 0);
  P->bitem_n.assign(njoins, 0);
  for (int pass = 0; pass < 2; ++pass) {
    for (int j = 0; j < njoins; ++j) {
      if ((P->sbits[j] > 0) != (pass == 1)) continue;
      P->bitem_first[j] = (int64_t)bj.size();
      for (int64_t r = 0; r < P->nb[j]; r += BUILD_ROWS) {
        bj.push_back(j);
        bs.push_back(r);
      }
      P->bitem_n[j] = (int64_t)bj.size() - P->bitem_first[j];
    }
    if (pass == 0) P->n_bitems_plain = (int64_t)bj.size();
  }
  // in each workgroup-size group the sliced joins' items come last: the
  // probe kernel takes the others, the sliced probe counts those
  for (int pass = 0; pass < 2; ++pass)
    for (int j = 0; j < njoins; ++j) {
      const Tune& t = P->tune[P->cls[j]];
      const int g = t.nt == 512 ? 1 : 0;
      if (P->nb[j] == 0 || (P->sbits[j] > 0) != (pass == 1)) continue;  // no match possible
      for (int64_t r = 0; r < P->np[j]; r += t.chunk) {
        pj[g].push_back(j);
        ps[g].push_back(r);
        pl[g].push_back((int32_t)std::min<int64_t>(t.chunk, P->np[j] - r));
      }
      if (pass == 0) P->n_pitems_plain[g] = (int64_t)pj[g].size();
    }
  // probe items: group 0 then group 1; per-join first/last item
  std::vector<int32_t> ij, il;
  std::vector<int64_t> is;
  for (int g = 0; g < 2; ++g) {
    P->pitem_base[g] = (int64_t)ij.size();
    P->n_pitems[g] = (int64_t)pj[g].size();
    ij.insert(ij.end(), pj[g].begin(), pj[g].end());
    is.insert(is.end(), ps[g].begin(), ps[g].end());
    il.insert(il.end(), pl[g].begin(), pl[g].end());
  }
  P->n_pitems_all = (int64_t)ij.size();
  P->n_bitems = (int64_t)bj.size();
  P->first_item.assign(njoins, -1);
  P->last_item.assign(njoins, -1);
  for (int64_t i = 0; i < P->n_pitems_all; ++i) {
    const int j = ij[i];
    if (P->first_item[j] < 0) P->first_item[j] = i;
    P->last_item[j] = i;
  }
  // device memory: one allocation
  const int64_t nbi = std::max<int64_t>(P->n_bitems, 1), npi = std::max<int64_t>(P->n_pitems_all, 1);
  Carve cv(nullptr);
  auto layout = [&](Carve& c) {
    P->d_b_off = c.take<int64_t>(njoins + 1);
    P->d_p_off = c.take<int64_t>(njoins + 1);
    P->d_t_off = c.take<int64_t>(njoins + 1);
    P->d_mask = c.take<uint32_t>(njoins);
    P->d_bitem_join = c.take<int32_t>(nbi);
    P->d_bitem_start = c.take<int64_t>(nbi);
    P->d_pitem_join = c.take<int32_t>(npi);
    P->d_pitem_len = c.take<int32_t>(npi);
    P->d_pitem_start = c.take<int64_t>(npi);
    P->d_item_cnt = c.take<int64_t>(npi + 1);
    P->d_item_pos = c.take<int64_t>(npi + 1);
    P->d_first = c.take<int64_t>(njoins);
    P->d_last = c.take<int64_t>(njoins);
    P->d_partner = c.take<int32_t>(std::max<int64_t>(P->total_p, 1));
    P->d_table = c.take<uint64_t>(std::max<int64_t>(P->total_slots, 1));
    P->d_queue = c.take<int32_t>(16);  // [0..4] the kernels' queues, [8..15] the slice queues
    P->d_scan = c.take<char>(scan_i64_scratch(npi + 1));
    const int64_t npairs = (int64_t)P->pair_join.size();
    P->d_sliced = c.take<int32_t>(njoins);
    P->d_chunk = c.take<int32_t>(njoins);
    P->d_soff_base = c.take<int64_t>(njoins);
    P->d_pair_join = c.take<int32_t>(npairs);
    P->d_pair_slice = c.take<int32_t>(npairs);
    P->d_xstart = c.take<int32_t>(npairs);
    if (P->any_sliced) {
      P->d_soff = c.take<int64_t>(P->soff_total);
      P->d_pk_part = c.take<int32_t>(std::max<int64_t>(P->total_p, 1));
      P->d_pv_part = c.take<int32_t>(std::max<int64_t>(P->total_p, 1));
      P->d_segs = c.take<RpSeg>(P->segs.size());
      P->d_segs1 = c.take<RpSeg>(P->segs1.size());
      P->d_seg_join = c.take<int32_t>(P->seg_join.size());
      const int64_t ent = std::max(P->seg_entries, P->bseg_entries);
      P->d_shist = c.take<int32_t>(ent);
      P->d_soffs = c.take<int64_t>(ent);
      P->d_sscan = c.take<char>(scan_i32_scratch(ent));
      P->d_bsegs = c.take<RpSeg>(P->bsegs.size());
      P->d_bsegs1 = c.take<RpSeg>(P->bsegs1.size());
      P->d_bk_part = c.take<int32_t>(std::max<int64_t>(P->total_b, 1));
      P->d_bv_part = c.take<int32_t>(std::max<int64_t>(P->total_b, 1));
      P->d_bsoff = c.take<int64_t>(std::max<int64_t>(P->bsoff_total, 1));
      P->d_bsoff_base = c.take<int64_t>(njoins);
      P->d_overflow = c.take<int64_t>(1);
    }
  };
  layout(cv);
  P->dbytes = cv.bytes();
  if (hipMalloc(&P->dmem, P->dbytes) != hipSuccess) {
    fail("batch plan: cannot allocate " + std::to_string(P->dbytes) + " device bytes");
    delete P;
    return nullptr;
  }
  Carve real(P->dmem);
  layout(real);
  std::vector<uint32_t> mask(njoins);
  for (int j = 0; j < njoins; ++j) mask[j] = (uint32_t)(P->cap[j] - 1);
  auto up = [&](void* d, const void* h, size_t bytes) {
    return bytes == 0 || hipMemcpy(d, h, bytes, hipMemcpyHostToDevice) == hipSuccess;
  };
  bool okc = up(P->d_b_off, P->b_off.data(), (njoins + 1) * 8) &&
             up(P->d_p_off, P->p_off.data(), (njoins + 1) * 8) &&
             up(P->d_t_off, P->t_off.data(), (njoins + 1) * 8) &&
             up(P->d_mask, mask.data(), njoins * 4) &&
             up(P->d_bitem_join, bj.data(), bj.size() * 4) &&
             up(P->d_bitem_start, bs.data(), bs.size() * 8) &&
             up(P->d_pitem_join, ij.data(), ij.size() * 4) &&
             up(P->d_pitem_len, il.data(), il.size() * 4) &&
             up(P->d_pitem_start, is.data(), is.size() * 8) &&
             up(P->d_first, P->first_item.data(), njoins * 8) &&
             up(P->d_last, P->last_item.data(), njoins * 8);
  std::vector<int32_t> sliced(njoins), chunk(njoins);
  for (int j = 0; j < njoins; ++j) {
    sliced[j] = P->sbits[j];
    int lg = 0;
    while ((1 << lg) < P->tune[P->cls[j]].chunk) ++lg;
    chunk[j] = lg;  // log2 of the compaction items' rows
  }
  okc = okc && up(P->d_sliced, sliced.data(), njoins * 4) &&
        up(P->d_chunk, chunk.data(), njoins * 4) &&
        up(P->d_soff_base, P->soff_base.data(), njoins * 8) &&
        up(P->d_pair_join, P->pair_join.data(), P->pair_join.size() * 4) &&
        up(P->d_pair_slice, P->pair_slice.data(), P->pair_slice.size() * 4);
  if (P->any_sliced)
    okc = okc && up(P->d_segs, P->segs.data(), P->segs.size() * sizeof(RpSeg)) &&
          up(P->d_segs1, P->segs1.data(), P->segs1.size() * sizeof(RpSeg)) &&
          up(P->d_seg_join, P->seg_join.data(), P->seg_join.size() * 4) &&
          up(P->d_bsegs, P->bsegs.data(), P->bsegs.size() * sizeof(RpSeg)) &&
          up(P->d_bsegs1, P->bsegs1.data(), P->bsegs1.size() * sizeof(RpSeg)) &&
          up(P->d_bsoff_base, P->bsoff_base.data(), njoins * 8);
  if (!okc) {
    fail("batch plan: upload failed");
    (void)hipFree(P->dmem);
    delete P;
    return nullptr;
  }
  int dev = 0, ncu = 256;
  if (hipGetDevice(&dev) == hipSuccess)
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  P->persist = std::max(64, ncu * 4);
  ok();
  return P;
}

int pf_batch_plan_destroy(void* plan) {
  using namespace pf;
  Plan* P = (Plan*)plan;
  if (!P) return ok();
  if (P->exec) (void)hipGraphExecDestroy(P->exec);
  if (P->graph) (void)hipGraphDestroy(P->graph);
  if (P->dmem) (void)hipFree(P->dmem);
  delete P;
  return ok();
}

int64_t pf_batch_plan_info(void* plan, int32_t what) {
  pf::Plan* P = (pf::Plan*)plan;
  if (!P) return -1;
  switch (what) {
    case 0: return P->total_b;
    case 1: return P->total_p;
    case 2: return P->total_slots;
    case 3: return P->n_bitems;
    case 4: return P->n_pitems_all;
    case 5: return (int64_t)P->dbytes;
    case 6: {
      int64_t n = 0;
      for (int b : P->sbits) n += b > 0;
      return n;
    }
    case 7: {  // the last run's LDS tables that took more than one pass (waits)
      int64_t n = 0;
      if (!P->d_overflow || hipDeviceSynchronize() != hipSuccess ||
          hipMemcpy(&n, P->d_overflow, 8, hipMemcpyDeviceToHost) != hipSuccess)
        return 0;
      return n;
    }
    default: return -1;
  }
}

}  // extern "C"

namespace pf {
namespace {

// Block g: segment g's slice offsets, join-local, into soff (2^bits + 1).
__global__ void bj_slice_bounds(const RpSeg* __restrict__ segs, const int32_t* __restrict__ seg_join,
                                const int64_t* __restrict__ offs, const int64_t* __restrict__ p_off,
                                const int64_t* __restrict__ soff_base, int64_t* __restrict__ soff) {
  const RpSeg& g = segs[blockIdx.x];
  const int j = seg_join[blockIdx.x];
  const int B = 1 << g.bits;
  for (int b = threadIdx.x; b <= B; b += blockDim.x)
    soff[soff_base[j] + b] =
        b < B ? offs[g.hist_off + (int64_t)b * g.nchunks] + g.delta - p_off[j] : g.n;
}

// The sliced joins' build rows (all of them, or join `only`'s) partitioned
// by slice, as their probe rows are.
int build_part_stage(Plan& P, const int32_t* bkeys, int only, hipStream_t st) {
  const int nseg = only < 0 ? (int)P.bsegs.size() : 1;
  const int k = only < 0 ? 0 : P.seg_of[only];
  const RpSeg* segs = only < 0 ? P.d_bsegs : P.d_bsegs1 + k;
  const int64_t blocks = only < 0 ? P.bseg_blocks : P.bsegs1[k].nchunks;
  const int64_t entries = only < 0 ? P.bseg_entries : P.bsegs1[k].nchunks << P.bsegs1[k].bits;
  if (fill_i64(P.d_overflow, 1, 0, st)) return 1;
  if (radix_partition_segments(bkeys, segs, nseg, blocks, entries, P.d_shist, P.d_soffs,
                               P.d_sscan, P.d_bk_part, P.d_bv_part, st, "bj_slice_hist",
                               "bj_slice_scatter"))
    return 1;
  bj_slice_bounds<<<nseg, 256, 0, st>>>(segs, P.d_seg_join + k, P.d_soffs, P.d_b_off,
                                        P.d_bsoff_base, P.d_bsoff);
  return launch_check("batch build slices");
}

// The sliced joins' probe rows partitioned by slice (all of them, or join
// `only`), their items counted from zero, and the queues' chunk scans.
int slice_stage(Plan& P, const int32_t* pkeys, const SliceQueues& q, int only, hipStream_t st) {
  if (only < 0) {
    if (fill_i64(P.d_item_cnt, P.n_pitems_all, 0, st)) return 1;
  } else if (P.first_item[only] >= 0) {
    if (fill_i64(P.d_item_cnt + P.first_item[only],
                 P.last_item[only] - P.first_item[only] + 1, 0, st))
      return 1;
  }
  const int nseg = only < 0 ? (int)P.segs.size() : 1;
  const int k = only < 0 ? 0 : P.seg_of[only];
  const RpSeg* segs = only < 0 ? P.d_segs : P.d_segs1 + k;
  const int64_t blocks = only < 0 ? P.seg_blocks : P.segs1[k].nchunks;
  const int64_t entries = only < 0 ? P.seg_entries : P.segs1[k].nchunks << P.segs1[k].bits;
  if (radix_partition_segments(pkeys, segs, nseg, blocks, entries, P.d_shist, P.d_soffs,
                               P.d_sscan, P.d_pk_part, P.d_pv_part, st, "bj_slice_hist",
                               "bj_slice_scatter"))
    return 1;
  bj_slice_bounds<<<nseg, 256, 0, st>>>(segs, P.d_seg_join + k, P.d_soffs, P.d_p_off,
                                        P.d_soff_base, P.d_soff);
  bj_slice_scan<<<NXQ, 1024, 0, st>>>(P.d_pair_join, P.d_pair_slice, P.d_soff_base, P.d_soff,
                                      P.d_mask, P.d_sliced, q, P.d_xstart);
  return launch_check("batch slices");
}

int launch_sliced_probe(Plan& P, const SliceQueues& q, hipStream_t st) {
  KTimer _kt("bj_probe_lds", st);
  // one 1024-thread workgroup a CU (its LDS holds one sub-range)
  bj_probe_lds<<<(unsigned)std::max(NXQ, P.persist / 4), PL_NT, 0, st>>>(
      P.d_pk_part, P.d_p_off, P.d_bk_part, P.d_bv_part, P.d_b_off, P.d_bsoff_base, P.d_bsoff,
      P.d_mask, P.d_sliced, P.d_pair_join, P.d_pair_slice, P.d_soff_base, P.d_soff, P.d_xstart,
      q, P.d_first, P.d_chunk, P.d_partner, P.d_item_cnt, P.d_overflow, P.d_queue + 8);
  return launch_check("batch sliced probe");
}

int run_batched(Plan& P, const int32_t* bkeys, const int32_t* pkeys, int64_t* out_probe,
                int64_t* out_build, int64_t* counts, hipStream_t st) {
  {
    KTimer _kt("bj_fill", st);
    bj_fill_tables<<<2048, 256, 0, st>>>(P.d_table, P.plain_slots, P.d_queue, 16);
  }
  if (P.n_bitems_plain > 0) {
    KTimer _kt("bj_build", st);
    bj_build<<<(unsigned)std::min<int64_t>(P.persist, P.n_bitems_plain), 256, 0, st>>>(
        bkeys, P.d_b_off, P.d_t_off, P.d_mask, P.d_bitem_join, P.d_bitem_start,
        P.n_bitems_plain, 0, P.d_table, P.d_queue + 0);
  }
  if (P.any_sliced && build_part_stage(P, bkeys, -1, st)) return 1;
  if (P.any_sliced && slice_stage(P, pkeys, P.qall, -1, st)) return 1;
  for (int g = 0; g < 2; ++g) {
    if (P.n_pitems_plain[g] == 0) continue;
    KTimer _kt("bj_probe", st);
    const unsigned grid = (unsigned)std::min<int64_t>(P.persist, P.n_pitems_plain[g]);
    if (g == 0)
      bj_probe<256><<<grid, 256, 0, st>>>(pkeys, P.d_p_off, P.d_t_off, P.d_mask, P.d_pitem_join,
                                          P.d_pitem_start, P.d_pitem_len, P.n_pitems_plain[g],
                                          P.pitem_base[g], P.d_table, P.d_sliced, P.d_partner,
                                          P.d_item_cnt, P.d_queue + 1);
    else
      bj_probe<512><<<grid, 512, 0, st>>>(pkeys, P.d_p_off, P.d_t_off, P.d_mask, P.d_pitem_join,
                                          P.d_pitem_start, P.d_pitem_len, P.n_pitems_plain[g],
                                          P.pitem_base[g], P.d_table, P.d_sliced, P.d_partner,
                                          P.d_item_cnt, P.d_queue + 2);
  }
  if (launch_check("batch join")) return 1;
  if (P.any_sliced && launch_sliced_probe(P, P.qall, st)) return 1;
  if (scan_i64_excl_from(P.d_item_cnt, P.d_item_pos, P.n_pitems_all, P.d_item_pos +
                         std::max<int64_t>(P.n_pitems_all, 0), P.d_scan, st))
    return 1;
  for (int g = 0; g < 2; ++g) {
    if (P.n_pitems[g] == 0) continue;
    KTimer _kt("bj_compact", st);
    const unsigned grid = (unsigned)std::min<int64_t>(P.persist, P.n_pitems[g]);
    if (g == 0)
      bj_compact<256><<<grid, 256, 0, st>>>(P.d_p_off, P.d_pitem_join, P.d_pitem_start,
                                            P.d_pitem_len, P.n_pitems[g], P.pitem_base[g],
                                            P.d_first, P.d_partner, P.d_sliced, P.d_pv_part,
                                            P.d_item_pos, out_probe,
                                            out_build, P.d_queue + 3);
    else
      bj_compact<512><<<grid, 512, 0, st>>>(P.d_p_off, P.d_pitem_join, P.d_pitem_start,
                                            P.d_pitem_len, P.n_pitems[g], P.pitem_base[g],
                                            P.d_first, P.d_partner, P.d_sliced, P.d_pv_part,
                                            P.d_item_pos, out_probe,
                                            out_build, P.d_queue + 4);
  }
  bj_counts<<<grid_for(P.nj, 256), 256, 0, st>>>(P.d_item_pos, P.d_first, P.d_last, P.nj,
                                                 counts);
  return launch_check("batch join");
}

// One join at a time through the same kernels, waiting for each join's
// count as a per-call entry point does.
int run_per_join(Plan& P, const int32_t* bkeys, const int32_t* pkeys, int64_t* out_probe,
                 int64_t* out_build, int64_t* counts, hipStream_t st) {
  for (int j = 0; j < P.nj; ++j) {
    // (a sliced join has no table in the arena)
    const int64_t fill = P.sbits[j] > 0 ? 0 : P.cap[j];
    bj_fill_tables<<<std::max<unsigned>(1, std::min<unsigned>(grid_for(fill, 256), 2048)), 256,
                     0, st>>>(P.d_table + P.t_off[j], fill, P.d_queue, 16);
    const int64_t nbi = P.bitem_n[j];
    if (P.sbits[j] > 0) {
      if (build_part_stage(P, bkeys, j, st)) return 1;
    } else if (nbi > 0) {
      bj_build<<<(unsigned)std::min<int64_t>(P.persist, nbi), 256, 0, st>>>(
          bkeys, P.d_b_off, P.d_t_off, P.d_mask, P.d_bitem_join, P.d_bitem_start,
          P.bitem_first[j] + nbi, P.bitem_first[j], P.d_table, P.d_queue + 0);
    }
    const int64_t f = P.first_item[j], l = P.last_item[j];
    if (f >= 0) {
      const int64_t n = l - f + 1;
      const unsigned grid = (unsigned)std::min<int64_t>(P.persist, n);
      const bool big = P.tune[P.cls[j]].nt == 512;
      if (P.sbits[j] > 0) {
        if (slice_stage(P, pkeys, P.qjoin[j], j, st) || launch_sliced_probe(P, P.qjoin[j], st))
          return 1;
      } else if (!big)
        bj_probe<256><<<grid, 256, 0, st>>>(pkeys, P.d_p_off, P.d_t_off, P.d_mask,
                                            P.d_pitem_join, P.d_pitem_start, P.d_pitem_len, n, f,
                                            P.d_table, P.d_sliced, P.d_partner, P.d_item_cnt,
                                            P.d_queue + 1);
      else
        bj_probe<512><<<grid, 512, 0, st>>>(pkeys, P.d_p_off, P.d_t_off, P.d_mask,
                                            P.d_pitem_join, P.d_pitem_start, P.d_pitem_len, n, f,
                                            P.d_table, P.d_sliced, P.d_partner, P.d_item_cnt,
                                            P.d_queue + 1);
      // this join's items scanned on their own; positions relative to f
      if (scan_i64_excl_from(P.d_item_cnt + f, P.d_item_pos + f, n, P.d_item_pos + f + n,
                             P.d_scan, st))
        return 1;
      if (!big)
        bj_compact<256><<<grid, 256, 0, st>>>(P.d_p_off, P.d_pitem_join, P.d_pitem_start,
                                              P.d_pitem_len, n, f, P.d_first, P.d_partner,
                                              P.d_sliced, P.d_pv_part, P.d_item_pos, out_probe,
                                              out_build,
                                              P.d_queue + 3);
      else
        bj_compact<512><<<grid, 512, 0, st>>>(P.d_p_off, P.d_pitem_join, P.d_pitem_start,
                                              P.d_pitem_len, n, f, P.d_first, P.d_partner,
                                              P.d_sliced, P.d_pv_part, P.d_item_pos, out_probe,
                                              out_build,
                                              P.d_queue + 3);
    }
    bj_counts<<<1, 64, 0, st>>>(P.d_item_pos, P.d_first + j, P.d_last + j, 1, counts + j);
    if (launch_check("batch join (per join)")) return 1;
    if (wait_stream(st, "batch join (per join)")) return 1;
  }
  return 0;
}

}  // namespace
}  // namespace pf

extern "C" {

int pf_batch_join(void* plan, const int32_t* bkeys, const int32_t* pkeys, int64_t* out_probe,
                  int64_t* out_build, int64_t* counts, int32_t mode, void* stream) {
  using namespace pf;
  Plan* P = (Plan*)plan;
  if (!P) return fail("batch join: no plan");
  hipStream_t st = (hipStream_t)stream;
  int rc = mode == 1 ? run_per_join(*P, bkeys, pkeys, out_probe, out_build, counts, st)
                     : run_batched(*P, bkeys, pkeys, out_probe, out_build, counts, st);
  if (rc) return rc;
  return ok();
}

int pf_batch_capture(void* plan, const int32_t* bkeys, const int32_t* pkeys, int64_t* out_probe,
                     int64_t* out_build, int64_t* counts, void* stream) {
  using namespace pf;
  Plan* P = (Plan*)plan;
  if (!P) return fail("batch capture: no plan");
  hipStream_t st = (hipStream_t)stream;
  if (P->exec) {
    (void)hipGraphExecDestroy(P->exec);
    P->exec = nullptr;
  }
  if (P->graph) {
    (void)hipGraphDestroy(P->graph);
    P->graph = nullptr;
  }
  if (hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "batch capture"))
    return 1;
  const int rc = run_batched(*P, bkeys, pkeys, out_probe, out_build, counts, st);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(st, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (hip_check(e, "batch capture")) return 1;
  P->graph = g;
  if (hip_check(hipGraphInstantiate(&P->exec, g, nullptr, nullptr, 0), "batch capture"))
    return 1;
  return ok();
}

int pf_batch_replay(void* plan, void* stream) {
  using namespace pf;
  Plan* P = (Plan*)plan;
  if (!P || !P->exec) return fail("batch replay: nothing captured");
  if (hip_check(hipGraphLaunch(P->exec, (hipStream_t)stream), "batch replay")) return 1;
  return ok();
}

}  // extern "C"
