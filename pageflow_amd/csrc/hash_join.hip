// Partitioned equi-join with LDS hash tables: the large-join tier (16M-row
// build, 640M-row probe) and the multi-match join that replaces the
// reference's sort/search core (operators/hashjoin.py:380-441 +
// join_core.expand) for non-unique builds.
//
// Both sides are radix-partitioned by the top `bits` bits of mix64(key)
// (pf_radix_partition in its hash-out mode: each row becomes (mix64(key),
// row id)).  Because mix64 is a bijection on 64 bits, equal hashes are equal
// keys, so the join compares hashes only and never reads a key again.
//
// Work items are (bucket, probe chunk: 256K rows unique, 64K multi).  A
// workgroup loads its bucket's build hashes into LDS (a chunk at a time when
// the bucket is larger than the LDS table), builds an open-addressing table
// over them (16384 u16 slots, two to a word, LDS CAS: at most 5120 rows a
// chunk keeps the load factor near 0.3 and the waves' longest probe chains
// short), and streams the probe chunk against it.
//   unique   each hash's lowest build row answers (a proven-unique build)
//   multi    the build bucket is first sorted by (hash, row) in LDS
//            (hj_sort_buckets), so a key's rows are one run; the table maps
//            a hash to its run's head and the run's length
// Pairs are counted per item (hj_probe<COUNT>), the counts scanned, then
// written (hj_probe<WRITE>) at deterministic positions: items in order,
// probe rows in order inside an item, a probe row's build rows ascending.
// The payload columns of the build side are gathered in the write pass.
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "pf_common.hpp"
#include "pf_residual.hpp"
#include "../../include/pageflow_join.h"

namespace pf {

size_t radix_partition_ws(int64_t n, int bits);
size_t radix_level_ws(int64_t n, int bits);
int radix_partition_level(const int64_t* keys, int64_t n, int bits, uint64_t* tk, int32_t* tv,
                          uint64_t* h_out, int32_t* row_out, int64_t* boff, void* ws,
                          size_t ws_bytes, hipStream_t st, int who);
int radix_partition(const void* keys, int key_bytes, const int32_t* vals, int64_t n, int bits,
                    int shift, int mode, void* keys_out, int32_t* vals_out,
                    int64_t* bucket_offsets, void* ws, size_t ws_bytes, hipStream_t st,
                    int who, int shape);
int radix_partition_carry(const int64_t* keys, int64_t n, int bits, int nv,
                          const void* const* src, const int32_t* type, int64_t* const* dst,
                          uint64_t* tk, int32_t* tv, int64_t* const* tx, uint64_t* h_out,
                          int32_t* row_out, int64_t* boff, void* ws, size_t ws_bytes,
                          hipStream_t st, int who);

namespace {

constexpr int HJ_NT = 1024;
constexpr int HJ_W = HJ_NT / kWave;
constexpr int HJ_IT = 12;                  // write pass: probe rows a thread per round
constexpr int HJ_ROUND = HJ_NT * HJ_IT;
constexpr int CAP_U = 5120;                // unique: build rows an LDS chunk
constexpr int TS_U = 16384;                // unique: table slots (u16 each)
constexpr int CAP_M = 4096;                // multi: build rows an LDS chunk
constexpr int TS_M = 16384;
constexpr int SORT_CAP = 8192;             // multi: largest build bucket sorted in LDS
constexpr int SORT_NT = 512;
// probe rows a work item: a multi-match item writes up to its rows times
// its keys' run lengths, so skewed buckets are cut finer to spread them
// over more workgroups (measured: 8.4 -> 7.2 ms on the Zipf(1.1) join; the
// unique join is fastest at 256K)
constexpr int PCH_U = 262144;
constexpr int PCH_M = 65536;
int pch_of(bool multi) { return multi ? PCH_M : PCH_U; }
constexpr int MAX_PAY = 2;
constexpr int kOnePassBits = 8;            // wider partitions take two passes
constexpr int kSortDigit = 8;              // a bucket sorted in HBM: bits a pass
constexpr int kBigBucketsGlobal = 32;      // more buckets past SORT_CAP: one global sort

// the widest digit one partition pass of partition_side takes
int pass_bits(int bits) { return bits <= kOnePassBits ? bits : bits - bits / 2; }

struct HjParams {
  const uint64_t* bh;     // build (hash, row) partitioned [nb]
  const int32_t* brow;
  const int64_t* boff;    // [P + 1]
  const uint64_t* ph;     // probe (hash, row) partitioned [np]
  const int32_t* prow;
  const int64_t* poff;    // [P + 1]
  const int64_t* item_off;  // [P + 1] each bucket's first work item
  const int32_t* item_bucket;  // [items] each work item's bucket
  int P;
  int pch;                // probe rows a work item
  int64_t n_items;
  int64_t* item_cnt;      // COUNT: pairs per item
  const int64_t* item_pos;  // WRITE: each item's first output row
  int64_t* next_out;        // WRITE in any order: the next free output row
  int64_t* out_probe;
  int64_t* out_build;
  const int64_t* pay_src[MAX_PAY];
  int64_t* pay_out[MAX_PAY];
  int npay;
};

// ------------------------------------------------------------- items --

// One block: each bucket's work items (its probe chunks; none when either
// side is empty), their prefix, the total and the largest build bucket
// into info[0..1].
__global__ __launch_bounds__(1024) void hj_items(const int64_t* __restrict__ boff,
                                                 const int64_t* __restrict__ poff, int P, int pch,
                                                 int64_t max_items, int64_t* __restrict__ item_off,
                                                 int32_t* __restrict__ item_bucket,
                                                 int64_t* __restrict__ info) {
  __shared__ int64_t tmp[1024 / kWave + 1];
  __shared__ int64_t mx[1024 / kWave];
  constexpr int PER = 4;  // P <= 4096
  int64_t v[PER], s = 0, m = 0;
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int b = threadIdx.x * PER + q;
    v[q] = 0;
    if (b < P) {
      const int64_t nb = boff[b + 1] - boff[b], np = poff[b + 1] - poff[b];
      if (nb > 0 && np > 0) v[q] = (np + pch - 1) / pch;
      m = nb > m ? nb : m;
    }
    s += v[q];
  }
  int64_t tot;
  int64_t ex = block_excl_scan<1024>(s, tmp, &tot);
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int b = threadIdx.x * PER + q;
    if (b < P) {
      item_off[b] = ex;
      for (int64_t c = 0; c < v[q] && ex + c < max_items; ++c) item_bucket[ex + c] = b;
    }
    ex += v[q];
  }
  // largest build bucket
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    const int64_t y = __shfl_xor(m, o, kWave);
    m = y > m ? y : m;
  }
  if (lane_id() == 0) mx[threadIdx.x / kWave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t mm = 0;
    for (int w = 0; w < 1024 / kWave; ++w) mm = mx[w] > mm ? mx[w] : mm;
    item_off[P] = tot;
    info[0] = tot;
    info[1] = mm;
  }
}

// ---------------------------------------------------------- LDS sort --

// Block b sorts build bucket b's (hash, row) pairs by hash, then row
// (bitonic network in LDS; the bucket holds at most SORT_CAP rows).  The LDS
// is sized at launch for the largest bucket sorted here (cap rows, a power
// of two): with buckets of a few thousand rows, several workgroups share a
// CU where a fixed SORT_CAP-row allocation would leave one.
__global__ __launch_bounds__(SORT_NT) void hj_sort_buckets(uint64_t* __restrict__ bh,
                                                           int32_t* __restrict__ brow,
                                                           const int64_t* __restrict__ boff,
                                                           int cap, int min_n) {
  extern __shared__ uint64_t sort_lds[];
  uint64_t* sh = sort_lds;
  int32_t* sr = (int32_t*)(sort_lds + cap);
  const int64_t b0 = boff[blockIdx.x], n = boff[blockIdx.x + 1] - b0;
  // larger buckets are sorted in HBM; those up to min_n by another sort
  if (n <= 1 || n > cap || n <= min_n) return;
  int m = 2;
  while (m < n) m <<= 1;
  for (int i = threadIdx.x; i < m; i += SORT_NT) {
    if (i < n) {
      sh[i] = bh[b0 + i];
      sr[i] = brow[b0 + i];
    } else {
      sh[i] = ~0ULL;
      sr[i] = INT32_MAX;
    }
  }
  __syncthreads();
  for (int k = 2; k <= m; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < (m >> 1); i += SORT_NT) {
        const int lo = ((i & ~(j - 1)) << 1) | (i & (j - 1));
        const int hi = lo + j;
        const bool up = (lo & k) == 0;
        const uint64_t a = sh[lo], c = sh[hi];
        const int32_t ra = sr[lo], rc = sr[hi];
        const bool gt = a > c || (a == c && ra > rc);
        if (gt == up) {
          sh[lo] = c;
          sh[hi] = a;
          sr[lo] = rc;
          sr[hi] = ra;
        }
      }
      __syncthreads();
    }
  }
  for (int i = threadIdx.x; i < n; i += SORT_NT) {
    bh[b0 + i] = sh[i];
    brow[b0 + i] = sr[i];
  }
}

// Build bucket b's (hash, key, row) triples sorted by hash, then key, then
// row, for buckets of at most `cap` rows (a power of two <= KEYED_CAP):
// the residual's range form keeps a probe row's kept build rows as one
// sub-run of its key's run.
constexpr int KEYED_CAP = 4096;
__global__ __launch_bounds__(SORT_NT) void hj_sort_keyed(uint64_t* __restrict__ bh,
                                                         int64_t* __restrict__ bkey,
                                                         int32_t* __restrict__ brow,
                                                         const int64_t* __restrict__ boff,
                                                         int cap) {
  extern __shared__ uint64_t keyed_lds[];
  uint64_t* sh = keyed_lds;
  int64_t* sk = (int64_t*)(keyed_lds + cap);
  int32_t* sr = (int32_t*)(keyed_lds + 2 * cap);
  const int64_t b0 = boff[blockIdx.x], n = boff[blockIdx.x + 1] - b0;
  if (n <= 1 || n > cap) return;
  int m = 2;
  while (m < n) m <<= 1;
  for (int i = threadIdx.x; i < m; i += SORT_NT) {
    if (i < n) {
      sh[i] = bh[b0 + i];
      sk[i] = bkey[b0 + i];
      sr[i] = brow[b0 + i];
    } else {
      sh[i] = ~0ULL;
      sk[i] = INT64_MAX;
      sr[i] = INT32_MAX;
    }
  }
  __syncthreads();
  for (int k = 2; k <= m; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < (m >> 1); i += SORT_NT) {
        const int lo = ((i & ~(j - 1)) << 1) | (i & (j - 1));
        const int hi = lo + j;
        const bool up = (lo & k) == 0;
        const uint64_t a = sh[lo], c = sh[hi];
        const int64_t ka = sk[lo], kc = sk[hi];
        const int32_t ra = sr[lo], rc = sr[hi];
        const bool gt = a > c || (a == c && (ka > kc || (ka == kc && ra > rc)));
        if (gt == up) {
          sh[lo] = c;
          sh[hi] = a;
          sk[lo] = kc;
          sk[hi] = ka;
          sr[lo] = rc;
          sr[hi] = ra;
        }
      }
      __syncthreads();
    }
  }
  for (int i = threadIdx.x; i < n; i += SORT_NT) {
    bh[b0 + i] = sh[i];
    bkey[b0 + i] = sk[i];
    brow[b0 + i] = sr[i];
  }
}

// ------------------------------------------------------ probe / write --

struct Item {
  int64_t bs, be, ps, pe;  // the bucket's build rows, the item's probe rows
};

__device__ __forceinline__ Item item_of(const HjParams& p, int64_t item) {
  const int b = p.item_bucket[item];
  const int64_t c = item - p.item_off[b];
  Item it;
  it.bs = p.boff[b];
  it.be = p.boff[b + 1];
  it.ps = p.poff[b] + c * p.pch;
  it.pe = it.ps + p.pch < p.poff[b + 1] ? it.ps + p.pch : p.poff[b + 1];
  return it;
}

template <bool MULTI>
struct Lds;
template <>
struct Lds<false> {
  static constexpr int CAP = CAP_U, TS = TS_U;
};
template <>
struct Lds<true> {
  static constexpr int CAP = CAP_M, TS = TS_M;
};

// The LDS table's entry of hash h: its first row in the chunk, or -1.
// Slots are u16 holding row + 1 (0 = empty), two to a 32-bit word, linear
// probing from the hash's low bits.
static_assert(CAP_U < 65536 && CAP_M < 65536, "row + 1 fits a u16 slot");
template <int TS>
__device__ __forceinline__ int lds_find(const uint64_t* keys, const uint16_t* slot, uint64_t h) {
  uint32_t s = (uint32_t)h & (TS - 1);
  uint32_t v = slot[s];
  while (v != 0u) {
    if (keys[v - 1] == h) return (int)v - 1;
    s = (s + 1) & (TS - 1);
    v = slot[s];
  }
  return -1;
}

// Inserts row i (hash h) into the u16 table held in 32-bit words: CAS on the
// slot's word; a slot holding the same hash keeps the lower row.
template <int TS>
__device__ __forceinline__ void lds_insert(const uint64_t* keys, uint32_t* word, uint64_t h, int i) {
  const uint32_t me = (uint32_t)(i + 1);
  uint32_t s = (uint32_t)h & (TS - 1);
  for (;;) {
    uint32_t* w = word + (s >> 1);
    const int sh = (int)(s & 1u) * 16;
    uint32_t old = atomicCAS(w, 0xffffffffu, 0xffffffffu);  // a read of the word (never all ones)
    uint32_t cur = (old >> sh) & 0xffffu;
    while (cur == 0u) {
      const uint32_t prev = atomicCAS(w, old, old | (me << sh));
      if (prev == old) return;
      old = prev;
      cur = (old >> sh) & 0xffffu;
    }
    if (keys[cur - 1] == h) {  // a repeated hash (unique mode): lowest row
      while (cur > me) {
        const uint32_t prev = atomicCAS(w, old, (old & ~(0xffffu << sh)) | (me << sh));
        if (prev == old) return;
        old = prev;
        cur = (old >> sh) & 0xffffu;
      }
      return;
    }
    s = (s + 1) & (TS - 1);
  }
}

// MODE 0: count pairs per item; MODE 1: write them at the scanned item
// positions; MODE 2 (unique builds): write them without a count pass, each
// round of a workgroup taking its output rows with one atomic add on a
// counter, so rounds land in the order they finish.  Probe rows are taken
// in rounds of HJ_ROUND, item row i0 + k * HJ_NT + tid in slot k of a thread, so a
// wave's lanes touch consecutive rows: loads coalesce and, in the write
// pass, consecutive matches land on consecutive output rows.  A match's
// build row and first payload come from LDS (staged with the chunk).
static_assert(HJ_IT * HJ_W % kWave == 0, "one wave scans a round's (slot, wave) totals");

template <bool MULTI, int MODE>
__global__ __launch_bounds__(HJ_NT) void hj_probe(HjParams p) {
  constexpr int CAP = Lds<MULTI>::CAP, TS = Lds<MULTI>::TS;
  constexpr bool W1 = MODE >= 1;
  constexpr bool ANY = MODE == 2;
  static_assert(!(ANY && MULTI), "the any-order write is for unique builds");
  // rows a thread per round (the count pass stays within 64 VGPRs: two
  // 1024-thread workgroups a CU)
  constexpr int IT = W1 ? HJ_IT : 8;
  constexpr int ROUND = HJ_NT * IT;
  __shared__ uint64_t keys[CAP];
  __shared__ uint32_t slotw[TS / 2];
  const uint16_t* slot = reinterpret_cast<const uint16_t*>(slotw);
  __shared__ int32_t rows[W1 ? CAP : 1];
  __shared__ int64_t pay0[W1 ? CAP : 1];
  __shared__ uint16_t glen[MULTI ? CAP : 1];
  __shared__ int64_t cnt[HJ_IT * HJ_W];
  const int tid = threadIdx.x, lane = lane_id(), w = tid / kWave;
  const int64_t item = blockIdx.x;
  if (item >= p.n_items) return;
  const Item it = item_of(p, item);
  int64_t pos = 0;
  if (W1 && !ANY) pos = p.item_pos[item];
  __shared__ int64_t round_base;
  int64_t count = 0;
  const bool has_pay = W1 && p.npay > 0;
  // a bucket larger than an LDS chunk is joined one build chunk at a time
  for (int64_t cb = it.bs; cb < it.be; cb += CAP) {
    const int nloc = (int)(it.be - cb < CAP ? it.be - cb : CAP);
    __syncthreads();  // the previous chunk's lookups are done
    for (int s = tid; s < TS / 2; s += HJ_NT) slotw[s] = 0;
    for (int i = tid; i < nloc; i += HJ_NT) {
      keys[i] = p.bh[cb + i];
      if (W1) rows[i] = p.brow[cb + i];
      if (has_pay) pay0[i] = p.pay_src[0][cb + i];
    }
    __syncthreads();
    for (int i = tid; i < nloc; i += HJ_NT) {
      const uint64_t h = keys[i];
      if (MULTI && i > 0 && keys[i - 1] == h) continue;  // not a run head
      if (MULTI) {
        int len = 1;
        while (i + len < nloc && keys[i + len] == h) ++len;
        glen[i] = (uint16_t)(len > 65535 ? 65535 : len);
      }
      lds_insert<TS>(keys, slotw, h, i);
    }
    __syncthreads();

    // item-local 32-bit row indices (an item has at most 2^18 rows); the
    // loads clamp to the last row so no lane branches around them
    const uint64_t* __restrict__ ph = p.ph + it.ps;
    const int32_t* __restrict__ prw = p.prow + it.ps;
    const int np_item = (int)(it.pe - it.ps);
    for (int i0 = 0; i0 < np_item; i0 += ROUND) {
      int hidx[IT];
      int m[IT];
      uint64_t hv[IT];
      // the write passes load the round's probe rows here too: a load among
      // the stores below would wait for every store before it
      constexpr bool PR = W1;
      int32_t pr[PR ? IT : 1];
#pragma unroll
      for (int k = 0; k < IT; ++k) {
        const int i = i0 + k * HJ_NT + tid;
        const int ic = i < np_item ? i : np_item - 1;
        hv[k] = ph[ic];
        if (PR) pr[k] = prw[ic];
      }
#pragma unroll
      for (int k = 0; k < IT; ++k) {
        const int i = i0 + k * HJ_NT + tid;
        const int f = lds_find<TS>(keys, slot, hv[k]);
        hidx[k] = i < np_item ? f : -1;
        int len = hidx[k] >= 0 ? 1 : 0;
        if (MULTI && hidx[k] >= 0) {
          len = glen[hidx[k]];
          if (len == 65535)  // a run longer than glen's range, counted to its end
            while (hidx[k] + len < nloc && keys[hidx[k] + len] == hv[k]) ++len;
        }
        m[k] = len;
      }
      if (!W1) {
#pragma unroll
        for (int k = 0; k < IT; ++k) count += m[k];
        continue;
      }
      // output offsets in row order: within a wave by lane, across (round
      // slot, wave) by one wave's scan of the HJ_IT * HJ_W totals
      int inw[HJ_IT];  // a wave's pairs in a round slot: at most 64 * CAP_M
#pragma unroll
      for (int k = 0; k < HJ_IT; ++k) {
        int incl, tot;
        if (MULTI) {
          incl = wave_incl_scan(m[k]);
          tot = __shfl(incl, kWave - 1, kWave);
        } else {
          const uint64_t bal = ballot(m[k] != 0);
          incl = mbcnt(bal) + m[k];
          tot = __popcll(bal);
        }
        inw[k] = incl - m[k];
        if (lane == 0) cnt[k * HJ_W + w] = tot;
      }
      __syncthreads();
      if (w == 0) {  // PER consecutive (slot, wave) totals a lane
        constexpr int PER = HJ_IT * HJ_W / kWave;
        int64_t v[PER], sum = 0;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
          v[q] = cnt[lane * PER + q];
          sum += v[q];
        }
        int64_t ex = wave_incl_scan(sum) - sum;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
          cnt[lane * PER + q] = ex;
          ex += v[q];
        }
        // the last lane holds the round's total: take that many output rows
        if (ANY && lane == kWave - 1)
          round_base = ex ? (int64_t)atomicAdd((unsigned long long*)p.next_out,
                                               (unsigned long long)ex)
                          : 0;
      }
      __syncthreads();
      if (ANY) pos = round_base;
      if (MULTI) {
        // A wave writes its (round slot)'s pairs as one contiguous run: lane
        // j of each store takes output j of the run, found by a binary
        // search of the lanes' offsets staged in LDS, so the stores are
        // coalesced whatever the run lengths (skewed keys fan out).
        __shared__ int32_t s_ex[MULTI ? HJ_NT : 1], s_hi[MULTI ? HJ_NT : 1],
            s_pr[MULTI ? HJ_NT : 1];
        const int wb = w * kWave;
#pragma unroll
        for (int k = 0; k < HJ_IT; ++k) {
          s_ex[tid] = inw[k];
          s_hi[tid] = hidx[k];
          s_pr[tid] = m[k] ? pr[k] : 0;
          const int32_t T = (int32_t)__shfl(inw[k] + m[k], kWave - 1, kWave);
          wave_lds_sync();  // a wave reads only its own lanes' entries
          const int64_t base = pos + cnt[k * HJ_W + w];
          int lo0 = 0;  // the previous window's last source lane: sources only grow
          for (int j = lane; j < T; j += kWave) {
            int lo = lo0, hi = kWave;  // the last lane whose run starts at or before j
            while (hi - lo > 1) {
              const int mid = (lo + hi) >> 1;
              if (s_ex[wb + mid] <= j)
                lo = mid;
              else
                hi = mid;
            }
            const int li = s_hi[wb + lo] + (j - s_ex[wb + lo]);
            const int64_t o = base + j;
            p.out_probe[o] = (int64_t)s_pr[wb + lo];
            p.out_build[o] = (int64_t)rows[li];
            if (has_pay) p.pay_out[0][o] = pay0[li];
            if (p.npay > 1) p.pay_out[1][o] = p.pay_src[1][cb + li];
            // every lane is active here but in the last window, whose value is unused
            lo0 = __shfl(lo, kWave - 1, kWave);
          }
          wave_lds_sync();  // s_* are rewritten for the next slot
        }
      } else {
#pragma unroll
        for (int k = 0; k < HJ_IT; ++k) {
          if (m[k] == 0) continue;
          const int64_t o = pos + cnt[k * HJ_W + w] + inw[k];
          p.out_probe[o] = (int64_t)pr[k];
          p.out_build[o] = (int64_t)rows[hidx[k]];
          if (has_pay) p.pay_out[0][o] = pay0[hidx[k]];
          if (p.npay > 1) p.pay_out[1][o] = p.pay_src[1][cb + hidx[k]];
        }
      }
      // the round's last (slot, wave) total, from the wave that holds it
      __shared__ int64_t last_tot;
      if (tid == HJ_NT - 1) {
        const int k = HJ_IT - 1;
        last_tot = cnt[k * HJ_W + HJ_W - 1] + inw[k] + m[k];
      }
      __syncthreads();
      pos += last_tot;
    }
  }  // build chunks
  if (!W1) {
    const int64_t t = wave_sum(count);
    if (lane == 0) cnt[w] = t;
    __syncthreads();
    if (tid == 0) {
      int64_t s = 0;
      for (int q = 0; q < HJ_W; ++q) s += cnt[q];
      p.item_cnt[item] = s;
    }
  }
}

// ------------------------------------------------- residual in the join --
//
// The join with its residual program evaluated on every candidate pair in
// the count and write passes, so that only kept pairs are counted and
// written (the reference filters the materialised pairs:
// operators/hashjoin.py:428-439).  Operand sources (pf_residual.hpp): the
// first build column staged in LDS with the chunk (gathered into partition
// order once), a second one read in partition order from HBM, up to two
// probe columns carried through the probe side's partition, any other
// column read at the pair's row.
//   unique  a probe row's one candidate is evaluated in its own lane;
//   multi   a wave's candidates are evaluated 64 at a time: lane j of a
//           window takes the wave's candidate j (its source lane found by a
//           binary search over the lanes' candidate offsets staged in LDS),
//           so skewed runs keep every lane busy.  The count pass and the
//           write pass's first sweep count the kept candidates of each
//           (round slot, wave); the second sweep evaluates them again and
//           writes the kept ones compacted by ballot, at positions that
//           depend on nothing but the inputs.
// probe rows a thread a round: multi-match passes keep each slot's lookup
// across two sweeps, so they take fewer rows a round
constexpr int hr_it(bool multi) { return multi ? 4 : 8; }
static_assert(4 * HJ_W % kWave == 0, "one wave scans a round's (slot, wave) totals");

struct HjRes {
  const int64_t* pv[2];  // carried probe values in partition order (nullptr: none)
  const int64_t* bv0;    // the LDS-staged build column in partition order (nullptr: none)
  int64_t* cand;         // count pass: candidate pairs before the residual (one counter)
  int32_t* kt;           // multi: kept totals a (round, slot, wave), KT_ITEM an item
};

// A multi-match item's kept totals: at most PCH_M / (HJ_NT * 4) rounds of
// 4 slots x HJ_W waves.  The count pass stores them for items of one build
// chunk, and the write pass reads them instead of sweeping the candidates
// twice.
constexpr int KT_ITEM = PCH_M / (HJ_NT * 4) * 4 * HJ_W;

// The program on one pair: its fast form (decoded on the host; c.pv0 and
// c.pv1 then hold the row's resolved operands, jr_fast_row), or the
// LDS-decoded interpreter; in the range form, every candidate of a
// narrowed run (KM_ALL) or "build key rrel row key" (KM_KEY: c.bv holds the
// build keys, c.pv0 the row's key).
enum : int { KM_INTERP = 0, KM_FAST = 1, KM_ALL = 2, KM_KEY = 3 };
template <int KM>
__device__ __forceinline__ bool res_keep(const JoinResidual& R, const JrLds* d,
                                         const PairCtx& c) {
  if constexpr (KM == KM_FAST)
    return jr_fast_eval(R.fast, (R.fast.a_bv | R.fast.b_bv) ? c.bv[c.li] : 0, c.pv0, c.pv1);
  else if constexpr (KM == KM_INTERP)
    return jr_eval_lds(R, d, c);
  else if constexpr (KM == KM_ALL)
    return true;
  else
    return jr_key_rel(R.fast.rrel, R.fast.f64 != 0, c.bv[c.li], c.pv0);
}

// Candidates [lo, hi) of a run of n build keys k (ascending) that "build
// rrel x" keeps, x the row's key: a prefix, a suffix or the equal range;
// in the float64 domain NaN keys sort last and are never kept.
__device__ __forceinline__ void key_range(const int64_t* k, int n, int rel, bool f64, int64_t x,
                                          int* lo, int* hi) {
  auto bound = [&](int64_t v, bool upper) {
    int a = 0, b = n;
    while (a < b) {
      const int mid = (a + b) >> 1;
      if (upper ? k[mid] <= v : k[mid] < v)
        a = mid + 1;
      else
        b = mid;
    }
    return a;
  };
  if (f64 && x == kNanKey) {
    *lo = *hi = 0;
    return;
  }
  const int end = f64 ? bound(kNanKey, false) : n;
  switch (rel) {
    case PF_REL_LT: *lo = 0; *hi = bound(x, false); break;
    case PF_REL_LE: *lo = 0; *hi = bound(x, true); break;
    case PF_REL_GT: *lo = bound(x, true); *hi = end; break;
    case PF_REL_GE: *lo = bound(x, false); *hi = end; break;
    default: *lo = bound(x, false); *hi = bound(x, true); break;  // EQ
  }
  if (*hi < *lo) *hi = *lo;
}

// One (round slot)'s candidates of a wave, a window of 64 at a time: lane j
// takes the wave's candidate j (its source lane found by six halving steps
// over the lanes' candidate offsets staged in LDS, which never decrease) and
// evaluates the program on it, so a window's 64 lanes are busy however the
// runs are skewed and the evaluator's uniform work is paid once a window.
// WRITE stores the kept candidates, compacted by ballot, at base + the kept
// candidates before them.  Returns the wave's kept count.
template <int KM, bool WRITE>
__device__ __forceinline__ int res_sweep(const HjParams& p, const JoinResidual& R,
                                         const JrLds* d, PairCtx c, int T, int wb,
                                         const int32_t* s_ex, const int32_t* s_hi,
                                         const int32_t* s_pr, const int64_t* s_p0,
                                         const int64_t* s_p1, const int32_t* rows,
                                         int64_t base) {
  const int lane = lane_id();
  int kept = 0, lo0 = 0;
  for (int j0 = 0; j0 < T; j0 += kWave) {
    const int j = j0 + lane;
    int lo = lo0;  // the last lane whose run starts at or before j
#pragma unroll
    for (int step = kWave / 2; step > 0; step >>= 1) {
      const int nx = lo + step;
      const int ex = s_ex[wb + (nx < kWave ? nx : kWave - 1)];  // branch-free, clamped
      lo = (nx < kWave && ex <= j) ? nx : lo;
    }
    bool keep = false;
    if (j < T) {
      c.li = s_hi[wb + lo] + (j - s_ex[wb + lo]);
      c.prow = s_pr[wb + lo];
      c.pv0 = s_p0[wb + lo];
      c.pv1 = s_p1[wb + lo];
      keep = res_keep<KM>(R, d, c);
    }
    const uint64_t bal = ballot(keep);
    if (WRITE && keep) {
      const int64_t o = base + kept + mbcnt(bal);
      p.out_probe[o] = c.prow;
      p.out_build[o] = (int64_t)rows[c.li];
      for (int q = 0; q < p.npay; ++q) p.pay_out[q][o] = p.pay_src[q][c.cb + c.li];
    }
    kept += __popcll(bal);
    lo0 = __builtin_amdgcn_readlane(lo, kWave - 1);
  }
  return kept;
}

// a[k] of a small register array for a wave-uniform k that the compiler
// cannot fold: a chain of selects, so a loop over k need not be unrolled.
template <int N, typename T>
__device__ __forceinline__ T pick(const T (&a)[N], int k) {
  T v = a[0];
#pragma unroll
  for (int q = 1; q < N; ++q) v = k == q ? a[q] : v;
  return v;
}
template <int N, typename T>
__device__ __forceinline__ void put(T (&a)[N], int k, T v) {
#pragma unroll
  for (int q = 0; q < N; ++q) a[q] = k == q ? v : a[q];
}

// W1 = false: count the kept pairs of each item (and every candidate, into
// x.cand); true: write them at the scanned item positions.  FAST: the
// program's fast form (host-decoded, jr_eval_fast); else the interpreter
// over the program decoded into LDS.  Unique builds evaluate a row's one
// candidate in its lane; multi-match builds sweep each (slot, wave)'s
// candidates in windows (res_sweep): the count pass once, the write pass
// twice (kept totals for the offsets, then the stores).  The loops over a
// round's slots are unrolled, which keeps the per-slot values in registers;
// the evaluators are compact enough for the copies to fit the instruction
// cache.
template <bool MULTI, bool W1, int EV>
__global__ __launch_bounds__(HJ_NT) void hj_probe_res(HjParams p, JoinResidual R, HjRes x) {
  constexpr bool FAST = EV >= 1;
  // the window sweeps' evaluation: the interpreter, the fast form, or in
  // the range form the key comparison (buckets past the keyed sort)
  constexpr int KM = EV == 0 ? KM_INTERP : EV == 1 ? KM_FAST : KM_KEY;
  constexpr int CAP = Lds<MULTI>::CAP, TS = Lds<MULTI>::TS;
  constexpr int IT = hr_it(MULTI), ROUND = HJ_NT * IT;
  constexpr int MT = MULTI ? HJ_NT : 1;
  __shared__ uint64_t keys[CAP];
  __shared__ uint32_t slotw[TS / 2];
  const uint16_t* slot = reinterpret_cast<const uint16_t*>(slotw);
  __shared__ int32_t rows[CAP];
  __shared__ int64_t bv[CAP];
  __shared__ uint16_t glen[MULTI ? CAP : 1];
  __shared__ int64_t cnt[IT * HJ_W];
  __shared__ int32_t s_ex[MT], s_hi[MT], s_pr[MT];
  __shared__ int64_t s_p0[MT], s_p1[MT];
  __shared__ int64_t last_tot;
  __shared__ JrLds dec[FAST ? 1 : 1];
  const int tid = threadIdx.x, lane = lane_id(), w = tid / kWave, wb = w * kWave;
  const int64_t item = blockIdx.x;
  if (item >= p.n_items) return;
  if (!FAST) jr_decode(R, dec);  // read after the chunk's first barrier
  const Item it = item_of(p, item);
  // range form: a bucket the keyed sort ordered by (hash, key) narrows each
  // run to its kept sub-run; larger buckets compare keys per candidate
  const bool nar = EV == 2 && it.be - it.bs <= KEYED_CAP;
  // kept totals kept between the passes: multi-match items of one build
  // chunk that sweep their candidates
  const bool kt_use = MULTI && !nar && x.kt != nullptr && it.be - it.bs <= CAP;
  int32_t* const kt_item = kt_use ? x.kt + item * KT_ITEM : nullptr;
  int64_t pos = W1 ? p.item_pos[item] : 0;
  int64_t kept_n = 0, cand_n = 0;
  const bool has_p0 = x.pv[0] != nullptr, has_p1 = x.pv[1] != nullptr;
  for (int64_t cb = it.bs; cb < it.be; cb += CAP) {
    const int nloc = (int)(it.be - cb < CAP ? it.be - cb : CAP);
    __syncthreads();  // the previous chunk's lookups are done
    for (int s = tid; s < TS / 2; s += HJ_NT) slotw[s] = 0;
    for (int i = tid; i < nloc; i += HJ_NT) {
      keys[i] = p.bh[cb + i];
      rows[i] = p.brow[cb + i];
      if (x.bv0) bv[i] = x.bv0[cb + i];
    }
    __syncthreads();
    for (int i = tid; i < nloc; i += HJ_NT) {
      const uint64_t h = keys[i];
      if (MULTI && i > 0 && keys[i - 1] == h) continue;  // not a run head
      if (MULTI) {
        int len = 1;
        while (i + len < nloc && keys[i + len] == h) ++len;
        glen[i] = (uint16_t)(len > 65535 ? 65535 : len);
      }
      lds_insert<TS>(keys, slotw, h, i);
    }
    __syncthreads();

    const uint64_t* __restrict__ ph = p.ph + it.ps;
    const int32_t* __restrict__ prw = p.prow + it.ps;
    const int64_t* __restrict__ pv0 = has_p0 ? x.pv[0] + it.ps : nullptr;
    const int64_t* __restrict__ pv1 = has_p1 ? x.pv[1] + it.ps : nullptr;
    const int np_item = (int)(it.pe - it.ps);
    PairCtx c;
    c.bv = bv;
    c.rows = rows;
    c.cb = cb;
    for (int i0 = 0; i0 < np_item; i0 += ROUND) {
      uint64_t hv[IT];
      int32_t pr[IT];
      int64_t q0[IT], q1[IT];
      int hidx[IT], m[IT];
#pragma unroll
      for (int k = 0; k < IT; ++k) {
        const int i = i0 + k * HJ_NT + tid;
        const int ic = i < np_item ? i : np_item - 1;
        hv[k] = ph[ic];
        pr[k] = prw[ic];
        q0[k] = has_p0 ? pv0[ic] : 0;
        q1[k] = has_p1 ? pv1[ic] : 0;
        if constexpr (FAST) {  // from here: the row's resolved operands
          const int64_t ra = jr_fast_row(R.fast.asel, R.fast.aconv, R.fast.a, q0[k], q1[k]);
          const int64_t rb = jr_fast_row(R.fast.bsel, R.fast.bconv, R.fast.b, q0[k], q1[k]);
          q0[k] = ra;
          q1[k] = rb;
          if constexpr (EV == 2)  // the range form: the row operand's key
            q0[k] = jr_key(R.fast.xa ? ra : rb, R.fast.f64 != 0, false);
        }
      }
#pragma unroll
      for (int k = 0; k < IT; ++k) {
        const int i = i0 + k * HJ_NT + tid;
        const int f = lds_find<TS>(keys, slot, hv[k]);
        hidx[k] = i < np_item ? f : -1;
        int len = hidx[k] >= 0 ? 1 : 0;
        if (MULTI && hidx[k] >= 0) {
          len = glen[hidx[k]];
          if (len == 65535)  // a run longer than glen's range, counted to its end
            while (hidx[k] + len < nloc && keys[hidx[k] + len] == hv[k]) ++len;
        }
        m[k] = len;
        cand_n += len;
      }
      // kept candidates: unique, this lane's (0 or 1); multi, the wave's in
      // each slot
      int kl[IT], inw[IT];
#pragma unroll
      for (int k = 0; k < IT; ++k) kl[k] = 0;
      if constexpr (MULTI) {
        if (nar) {
#pragma unroll
          for (int k = 0; k < IT; ++k) {
            if (m[k] > 0) {
              int lo, hi;
              key_range(bv + hidx[k], m[k], R.fast.rrel, R.fast.f64 != 0, q0[k], &lo, &hi);
              hidx[k] += lo;
              m[k] = hi - lo;
            }
          }
        }
#pragma unroll
        for (int k = 0; k < IT; ++k) {
          const int incl = wave_incl_scan(m[k]);
          inw[k] = incl - m[k];
          m[k] = __shfl(incl, kWave - 1, kWave);  // from here: the wave's candidates
        }
      }
      // stages slot k's lanes for its window sweeps
      auto stage = [&](int k) {
        s_ex[tid] = pick(inw, k);
        s_hi[tid] = pick(hidx, k);
        s_pr[tid] = pick(pr, k);
        s_p0[tid] = pick(q0, k);
        s_p1[tid] = pick(q1, k);
        wave_lds_sync();  // a wave reads only its own lanes' entries
      };
#pragma unroll
      for (int k = 0; k < IT; ++k) {
        int kk = 0;
        if constexpr (MULTI) {
          const int T = __builtin_amdgcn_readfirstlane(pick(m, k));
          if (T > 0 && nar) {
            kk = T;  // a narrowed run keeps every candidate
          } else if (W1 && kt_use) {
            kk = kt_item[(i0 / ROUND) * (IT * HJ_W) + k * HJ_W + w];  // the count pass's
          } else if (T > 0) {
            stage(k);
            kk = res_sweep<KM, false>(p, R, dec, c, T, wb, s_ex, s_hi, s_pr, s_p0, s_p1, rows,
                                      0);
            wave_lds_sync();  // s_* are rewritten for the next slot
          }
          if (!W1 && lane == 0) {
            kept_n += kk;
            if (kt_use) kt_item[(i0 / ROUND) * (IT * HJ_W) + k * HJ_W + w] = kk;
          }
        } else {
          if (pick(m, k) > 0) {
            c.li = pick(hidx, k);
            c.prow = pick(pr, k);
            c.pv0 = pick(q0, k);
            c.pv1 = pick(q1, k);
            kk = res_keep<KM>(R, dec, c) ? 1 : 0;
          }
          if (!W1) kept_n += kk;
        }
        put(kl, k, kk);
      }
      if (!W1) continue;
      // offsets: within a wave by lane (unique); across (slot, wave) by one
      // wave's scan of the kept totals
#pragma unroll
      for (int k = 0; k < IT; ++k) {
        if constexpr (MULTI) {
          if (lane == 0) cnt[k * HJ_W + w] = kl[k];
        } else {
          const uint64_t bal = ballot(kl[k] != 0);
          inw[k] = mbcnt(bal);
          if (lane == 0) cnt[k * HJ_W + w] = __popcll(bal);
        }
      }
      __syncthreads();
      if (w == 0) {  // PER consecutive (slot, wave) totals a lane
        constexpr int PER = IT * HJ_W / kWave;
        int64_t v[PER], sum = 0;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
          v[q] = cnt[lane * PER + q];
          sum += v[q];
        }
        int64_t ex = wave_incl_scan(sum) - sum;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
          cnt[lane * PER + q] = ex;
          ex += v[q];
        }
        if (lane == kWave - 1) last_tot = ex;
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < (MULTI ? IT : 0); ++k) {
        const int T = __builtin_amdgcn_readfirstlane(pick(m, k));
        if (T == 0) continue;
        stage(k);
        if (EV == 2 && nar)
          res_sweep<KM_ALL, true>(p, R, dec, c, T, wb, s_ex, s_hi, s_pr, s_p0, s_p1, rows,
                                  pos + cnt[k * HJ_W + w]);
        else
          res_sweep<KM, true>(p, R, dec, c, T, wb, s_ex, s_hi, s_pr, s_p0, s_p1, rows,
                              pos + cnt[k * HJ_W + w]);
        wave_lds_sync();  // s_* are rewritten for the next slot
      }
#pragma unroll
      for (int k = 0; k < IT; ++k) {
        if constexpr (!MULTI) {
          if (!kl[k]) continue;
          const int64_t o = pos + cnt[k * HJ_W + w] + inw[k];
          p.out_probe[o] = (int64_t)pr[k];
          p.out_build[o] = (int64_t)rows[hidx[k]];
          for (int q = 0; q < p.npay; ++q) p.pay_out[q][o] = p.pay_src[q][cb + hidx[k]];
        }
      }
      pos += last_tot;
      __syncthreads();  // last_tot and cnt are rewritten by the next round
    }
  }  // build chunks
  if (!W1) {
    const int64_t t = wave_sum(kept_n), a = wave_sum(cand_n);
    if (lane == 0) {
      cnt[w] = t;
      cnt[HJ_W + w] = a;
    }
    __syncthreads();
    if (tid == 0) {
      int64_t s = 0, sa = 0;
      for (int q = 0; q < HJ_W; ++q) {
        s += cnt[q];
        sa += cnt[HJ_W + q];
      }
      p.item_cnt[item] = s;
      if (sa) atomicAdd((unsigned long long*)x.cand, (unsigned long long)sa);
    }
  }
}

// dst[i] = the widened value of a build column at idx[i]: a residual
// operand in partition order.
__global__ void hj_gather_res(const void* __restrict__ src, int type,
                              const int32_t* __restrict__ idx, int64_t n,
                              int64_t* __restrict__ dst) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride)
    dst[i] = rf_widen(src, type, idx[i]);
}

// dst[i] = the comparison key (jr_key) of a build column at idx[i]: the
// range form's build operand in partition order.
__global__ void hj_gather_key(const void* __restrict__ src, int type,
                              const int32_t* __restrict__ idx, int64_t n, int f64, int conv,
                              int64_t* __restrict__ dst) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride)
    dst[i] = jr_key(rf_widen(src, type, idx[i]), f64 != 0, conv != 0);
}

// dst[i] = src[idx[i]]: a build payload column in partition order.
__global__ void hj_gather_pay(const int64_t* __restrict__ src, const int32_t* __restrict__ idx,
                              int64_t n, int64_t* __restrict__ dst) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride)
    dst[i] = src[idx[i]];
}

__global__ void hj_readback(const int64_t* __restrict__ a, const int64_t* __restrict__ b,
                            int64_t* host) {
  if (threadIdx.x == 0) store_host(&host[0], a[0]);
  if (threadIdx.x == 1) store_host(&host[1], a[1]);
  if (threadIdx.x == 2) store_host(&host[2], b ? b[0] : 0);
}

// Bucket b's first row in h sorted by its top `bits` bits (b = P: n).
__global__ void hj_bucket_bounds(const uint64_t* __restrict__ h, int64_t n, int bits,
                                 int64_t* __restrict__ boff) {
  const int P = 1 << bits;
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b > P) return;
  if (b == P) {
    boff[b] = n;
    return;
  }
  int64_t lo = 0, hi = n;  // first index with top bits >= b
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if ((int64_t)(h[mid] >> (64 - bits)) < b)
      lo = mid + 1;
    else
      hi = mid;
  }
  boff[b] = lo;
}

struct HjWs {
  uint64_t* bh;
  int32_t* brow;
  int64_t* boff;
  uint64_t* ph;
  int32_t* prow;
  int64_t* poff;
  int64_t* item_off;
  int32_t* item_bucket;
  int64_t* info;       // [0] items, [1] largest build bucket
  int64_t* total;
  uint64_t* th;        // the first pass's output of a two-pass partition
  int32_t* trow;
  int64_t* item_cnt;   // [max items] (counts, then scanned in place)
  int64_t* sort_off;   // multi: a sort pass's 257 digit offsets
  int64_t* pay_part[MAX_PAY];  // build payloads in partition order
  void* part_ws;
  void* scan;
  size_t bytes;
};

// The partition workspace of both sides.
size_t part_ws_bytes(int64_t nb, int64_t np, int bits) {
  const int64_t n = nb > np ? nb : np;
  if (bits <= kOnePassBits) return radix_partition_ws(n, bits);
  const size_t a = radix_level_ws(nb, bits), b = radix_level_ws(np, bits);
  return a > b ? a : b;
}

// Upper bound of work items: every bucket's probe chunks.
int64_t max_items(int64_t np, int bits, bool multi) {
  return np / pch_of(multi) + (1LL << bits) + 1;
}

HjWs carve_hj(void* ws, int64_t nb, int64_t np, int bits, bool multi) {
  Carve cv(ws);
  HjWs w;
  const int64_t P = 1LL << bits;
  w.bh = cv.take<uint64_t>(nb > 0 ? nb : 1);
  w.brow = cv.take<int32_t>(nb > 0 ? nb : 1);
  w.boff = cv.take<int64_t>(P + 1);
  w.ph = cv.take<uint64_t>(np > 0 ? np : 1);
  w.prow = cv.take<int32_t>(np > 0 ? np : 1);
  w.poff = cv.take<int64_t>(P + 1);
  w.item_off = cv.take<int64_t>(P + 1);
  w.item_bucket = cv.take<int32_t>(max_items(np, bits, multi));
  // the two-pass partition's first output; multi: also the ping-pong
  // buffer of a build bucket sorted in HBM
  const int64_t nt = bits > kOnePassBits ? (nb > np ? nb : np) : multi ? nb : 1;
  w.th = cv.take<uint64_t>(nt > 0 ? nt : 1);
  w.trow = cv.take<int32_t>(nt > 0 ? nt : 1);
  w.info = cv.take<int64_t>(4);
  w.total = cv.take<int64_t>(2);
  const int64_t mi = max_items(np, bits, multi);
  w.item_cnt = cv.take<int64_t>(mi);
  w.sort_off = cv.take<int64_t>(multi ? 257 : 1);
  for (int c = 0; c < MAX_PAY; ++c) w.pay_part[c] = cv.take<int64_t>(nb > 0 ? nb : 1);
  size_t pws = part_ws_bytes(nb, np, bits);
  if (multi && radix_partition_ws(nb, kSortDigit) > pws) pws = radix_partition_ws(nb, kSortDigit);
  w.part_ws = cv.take<char>(pws);
  w.scan = cv.take<char>(scan_i64_scratch(mi));
  w.bytes = cv.bytes();
  return w;
}

// One side as (mix64(key), row id) ordered by the top `bits` bits of the
// hash, stable.  Up to kOnePassBits in one pass; above, the two-level
// partition: one histogram of the whole digit, then two LSD passes of
// about half the bits each (the lower digit first, then the upper one over
// the hashes), so that every pass writes runs long enough to coalesce; the
// bucket bounds come from its offsets.
int partition_side(const int64_t* keys, int64_t n, int bits, uint64_t* h, int32_t* row,
                   int64_t* boff, uint64_t* th, int32_t* trow, void* pws, size_t pws_bytes,
                   hipStream_t st, int who) {
  if (bits <= kOnePassBits)
    return radix_partition(keys, 8, nullptr, n, bits, 0, 1, h, row, boff, pws, pws_bytes, st,
                           who, -1);
  return radix_partition_level(keys, n, bits, th, trow, h, row, boff, pws, pws_bytes, st, who);
}

// Multi-match build buckets of more than SORT_CAP rows, sorted by (hash,
// row) in HBM: a bucket's rows are in row order (the partition is stable),
// so stable LSD passes over the hash bits below the top `bits` order them
// by (hash, row), as the LDS sort orders the smaller buckets.  One wait for
// the bucket bounds.
int sort_big_buckets(const HjWs& w, int64_t nb, int bits, hipStream_t st, bool* all_sorted) {
  const int P = 1 << bits;
  std::vector<int64_t> off(P + 1);
  if (hip_check(hipMemcpyAsync(off.data(), w.boff, (P + 1) * sizeof(int64_t),
                               hipMemcpyDeviceToHost, st), "hash_join bucket bounds"))
    return 1;
  if (wait_stream(st, "hash_join")) return 1;
  size_t pws = radix_partition_ws(nb > 0 ? nb : 1, kSortDigit);
  int big = 0;
  for (int b = 0; b < P; ++b) big += off[b + 1] - off[b] > SORT_CAP;
  if (big > kBigBucketsGlobal) {
    // most buckets large: one stable LSD sort of the whole side by all 64
    // hash bits (8 passes, ending in place) leaves every bucket in its range
    // in (hash, row) order, in far fewer launches than a sort a bucket
    uint64_t* k[2] = {w.bh, w.th};
    int32_t* r[2] = {w.brow, w.trow};
    for (int pass = 0; pass < 64 / kSortDigit; ++pass) {
      const int c = pass & 1;
      if (radix_partition(k[c], 8, r[c], nb, kSortDigit, pass * kSortDigit, 2, k[c ^ 1],
                          r[c ^ 1], w.sort_off, w.part_ws, pws, st, 4, -1))
        return 1;
    }
    *all_sorted = true;
    return 0;
  }
  for (int b = 0; b < P; ++b) {
    const int64_t n = off[b + 1] - off[b];
    if (n <= SORT_CAP) continue;
    uint64_t* k[2] = {w.bh + off[b], w.th};
    int32_t* r[2] = {w.brow + off[b], w.trow};
    int cur = 0;
    for (int shift = 0; shift < 64 - bits; shift += kSortDigit) {
      const int width = 64 - bits - shift < kSortDigit ? 64 - bits - shift : kSortDigit;
      if (radix_partition(k[cur], 8, r[cur], n, width, shift, 2, k[cur ^ 1], r[cur ^ 1],
                          w.sort_off, w.part_ws, pws, st, 4, -1))
        return 1;
      cur ^= 1;
    }
    if (cur == 1) {
      if (hip_check(hipMemcpyAsync(k[0], k[1], n * sizeof(uint64_t), hipMemcpyDeviceToDevice,
                                   st), "hash_join sort copy") ||
          hip_check(hipMemcpyAsync(r[0], r[1], n * sizeof(int32_t), hipMemcpyDeviceToDevice,
                                   st), "hash_join sort copy"))
        return 1;
    }
  }
  return 0;
}

// Value columns the probe side's partition carries (the residual's probe
// operands).
struct ProbeCarry {
  int nv = 0;
  const void* src[2] = {nullptr, nullptr};
  int32_t type[2] = {0, 0};
  int64_t* dst[2] = {nullptr, nullptr};
  int64_t* tx[2] = {nullptr, nullptr};
};

// Multi-match build buckets of more than min_n rows sorted by (hash, row):
// in HBM past the LDS sort's SORT_CAP rows, in LDS below.  *all_sorted: one
// global sort ordered every bucket.
int sort_buckets(const HjWs& w, int64_t nb, int bits, int64_t largest, int min_n,
                 hipStream_t st, bool* all_sorted) {
  const int P = 1 << bits;
  *all_sorted = false;
  if (largest > SORT_CAP && sort_big_buckets(w, nb, bits, st, all_sorted)) return 1;
  if (nb > 0 && !*all_sorted && largest > min_n) {
    {
      KTimer _kt("hj_sort_buckets", st);
      int cap = 2;  // the largest bucket sorted in LDS, to a power of two
      while (cap < largest && cap < SORT_CAP) cap <<= 1;
      const size_t lds = (size_t)cap * (sizeof(uint64_t) + sizeof(int32_t));
      if (hip_check(hipFuncSetAttribute((const void*)hj_sort_buckets,
                                        hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)(SORT_CAP * (sizeof(uint64_t) + sizeof(int32_t)))),
                    "hash_join sort LDS"))
        return 1;
      hj_sort_buckets<<<(unsigned)P, SORT_NT, lds, st>>>(w.bh, w.brow, w.boff, cap, min_n);
    }
    if (launch_check("hash_join sort")) return 1;
  }
  return 0;
}

// Both sides partitioned as (mix64(key), row) by the top `bits` bits of the
// hash (the probe side with its carried columns when pc has any), the work
// items built and read back (one wait), and with `sort`, for multi-match
// builds, every bucket sorted by (hash, row).  *n_items and *largest from
// the read-back.
int prepare_sides(const int64_t* bkeys, int64_t nb, const int64_t* pkeys, int64_t np, int bits,
                  bool multi, const HjWs& w, const ProbeCarry& pc, int64_t* host_out,
                  void* hdev, hipStream_t st, int64_t* n_items, int64_t* largest,
                  bool sort = true) {
  const int P = 1 << bits;
  const size_t pws = part_ws_bytes(nb, np, bits);
  if (partition_side(bkeys, nb, bits, w.bh, w.brow, w.boff, w.th, w.trow, w.part_ws, pws, st, 1))
    return 1;
  if (pc.nv > 0) {
    if (radix_partition_carry(pkeys, np, bits, pc.nv, pc.src, pc.type, pc.dst, w.th, w.trow,
                              pc.tx, w.ph, w.prow, w.poff, w.part_ws, pws, st, 2))
      return 1;
  } else if (partition_side(pkeys, np, bits, w.ph, w.prow, w.poff, w.th, w.trow, w.part_ws, pws,
                            st, 2)) {
    return 1;
  }
  hj_items<<<1, 1024, 0, st>>>(w.boff, w.poff, P, pch_of(multi), max_items(np, bits, multi),
                               w.item_off, w.item_bucket, w.info);
  hj_readback<<<1, 64, 0, st>>>(w.info, nullptr, (int64_t*)hdev);
  if (launch_check("hash_join items")) return 1;
  if (wait_stream(st, "hash_join")) return 1;
  *n_items = host_out[0];
  *largest = host_out[1];
  if (*n_items > max_items(np, bits, multi)) return fail("hash_join: work items overflow");
  bool all_sorted;
  if (multi && sort && sort_buckets(w, nb, bits, *largest, 0, st, &all_sorted)) return 1;
  return 0;
}

// ------------------------------------------ the residual join's host side --

// The residual's extra workspace after the join's own (HjWs): the carried
// probe columns (and the two-level partition's first-pass copies), the
// build columns in partition order, and the candidate counter.
struct HrWs {
  int64_t* px[2] = {nullptr, nullptr};
  int64_t* tx[2] = {nullptr, nullptr};
  int64_t* bx[2] = {nullptr, nullptr};
  int32_t* kt = nullptr;
  int64_t* dev = nullptr;
  size_t bytes = 0;
};

HrWs carve_hr(void* ws, size_t base, int64_t nb, int64_t np, int bits, bool multi, int nv,
              int nbv) {
  Carve cv(ws ? (char*)ws + base : nullptr);
  HrWs r;
  for (int c = 0; c < nv; ++c) r.px[c] = cv.take<int64_t>(np > 0 ? np : 1);
  if (bits > kOnePassBits)
    for (int c = 0; c < nv; ++c) r.tx[c] = cv.take<int64_t>(np > 0 ? np : 1);
  for (int c = 0; c < nbv; ++c) r.bx[c] = cv.take<int64_t>(nb > 0 ? nb : 1);
  if (multi) r.kt = cv.take<int32_t>(max_items(np, bits, true) * KT_ITEM);
  r.dev = cv.take<int64_t>(2);
  r.bytes = base + cv.bytes();
  return r;
}

// Where each column slot's values come from: the first two probe columns
// the program names are carried through the partition, the first build
// column is staged in LDS and the second read in partition order, and any
// other is read at the pair's row.  Every column must cover its side's
// rows.
int residual_plan(const RfProgram& P, int n_cols, int64_t nb, int64_t np, bool multi,
                  JoinResidual* R, int* nv, int* nbv, int pslot[2], int bslot[2]) {
  R->P = P;
  R->bpart1 = nullptr;
  for (int s = 0; s < PF_RES_MAX_COLS; ++s) R->src[s] = RS_ROW;
  *nv = *nbv = 0;
  bool seen[PF_RES_MAX_COLS] = {};
  for (int k = 0; k < P.n_ins; ++k) {
    const int64_t w0 = P.ins[k][0];
    if ((w0 & 0xff) != PF_RES_CMP) continue;
    for (int o = 0; o < 2; ++o) {
      if (!(w0 & (o ? PF_RES_B_COL : PF_RES_A_COL))) continue;
      const int sl = (int)P.ins[k][1 + o];
      if (seen[sl]) continue;
      seen[sl] = true;
      const bool probe = P.side[sl] != 0;
      if (P.rows[sl] < (probe ? np : nb))
        return fail("hash_join residual: column slot " + std::to_string(sl) + " holds " +
                    std::to_string(P.rows[sl]) + " rows, fewer than its side's " +
                    std::to_string(probe ? np : nb));
      if (probe && *nv < 2) {
        R->src[sl] = *nv == 0 ? RS_P0 : RS_P1;
        pslot[(*nv)++] = sl;
      } else if (!probe && *nbv < 2) {
        R->src[sl] = *nbv == 0 ? RS_BLDS : RS_BPART;
        bslot[(*nbv)++] = sl;
      }
    }
  }
  (void)n_cols;
  R->fast = jr_fast_of(*R);
  if (!multi) R->fast.range = 0;  // a unique build's runs are one row
  return 0;
}

template <bool MULTI, bool W1>
void launch_res(int64_t n_items, const HjParams& p, const JoinResidual& R, const HjRes& x,
                hipStream_t st) {
  if (MULTI && R.fast.range)
    hj_probe_res<MULTI, W1, MULTI ? 2 : 1><<<(unsigned)n_items, HJ_NT, 0, st>>>(p, R, x);
  else if (R.fast.on)
    hj_probe_res<MULTI, W1, 1><<<(unsigned)n_items, HJ_NT, 0, st>>>(p, R, x);
  else
    hj_probe_res<MULTI, W1, 0><<<(unsigned)n_items, HJ_NT, 0, st>>>(p, R, x);
}

HjParams params_of(const HjWs& w, int bits, bool multi, int64_t n_items) {
  HjParams p{};
  p.pch = pch_of(multi);
  p.bh = w.bh;
  p.brow = w.brow;
  p.boff = w.boff;
  p.ph = w.ph;
  p.prow = w.prow;
  p.poff = w.poff;
  p.item_off = w.item_off;
  p.item_bucket = w.item_bucket;
  p.P = 1 << bits;
  p.n_items = n_items;
  p.item_cnt = w.item_cnt;
  p.item_pos = w.item_cnt;
  return p;
}

}  // namespace
}  // namespace pf

extern "C" {

size_t pf_hash_join_workspace_bytes(int64_t nb, int64_t np, int32_t bits, int32_t multi) {
  if (bits < 0) bits = 0;
  if (bits > 12) bits = 12;
  return pf::carve_hj(nullptr, nb, np, bits, multi == 1).bytes;
}

int pf_hash_join_prepare(const int64_t* bkeys, int64_t nb, const int64_t* pkeys, int64_t np,
                         int32_t bits, int32_t multi, int64_t* host_out, void* workspace,
                         size_t workspace_bytes, void* stream) {
  using namespace pf;
  hipStream_t st = (hipStream_t)stream;
  if (bits < 0 || bits > 12) return fail("hash_join: bits must be in [0, 12]");
  if (nb < 0 || np < 0 || nb >= (1LL << 31) || np >= (1LL << 31))
    return fail("hash_join: a side takes fewer than 2^31 rows");
  if (multi < 0 || multi > 2) return fail("hash_join: multi is 0, 1 or 2");
  const bool any_order = multi == 2;
  multi = multi == 1;
  HjWs w = carve_hj(workspace, nb, np, bits, multi != 0);
  if (workspace_bytes < w.bytes) return fail("hash_join: workspace too small");
  void* hdev = nullptr;
  if (host_alias(host_out, &hdev)) return 1;
  int64_t n_items = 0, largest = 0;
  if (prepare_sides(bkeys, nb, pkeys, np, bits, multi != 0, w, ProbeCarry{}, host_out, hdev, st,
                    &n_items, &largest))
    return 1;
  if (any_order) {
    // no count pass: the write pass finds the pairs (at most one a probe row)
    host_out[0] = -1;
    host_out[1] = largest;
    host_out[2] = n_items;
    return ok();
  }
  HjParams p = params_of(w, bits, multi != 0, n_items);
  if (n_items > 0) {
    {
    KTimer _kt("hj_count", st);
    if (multi)
      hj_probe<true, 0><<<(unsigned)n_items, HJ_NT, 0, st>>>(p);
    else
      hj_probe<false, 0><<<(unsigned)n_items, HJ_NT, 0, st>>>(p);
  }
    if (launch_check("hash_join count")) return 1;
  }
  if (scan_i64_excl(w.item_cnt, n_items, w.total, w.scan, st)) return 1;
  hj_readback<<<1, 64, 0, st>>>(w.total, w.info, (int64_t*)hdev);
  if (launch_check("hash_join count")) return 1;
  if (wait_stream(st, "hash_join")) return 1;
  // host_out: [0] pairs, [1] (unused), [2] work items
  const int64_t pairs = host_out[0];
  host_out[1] = largest;
  host_out[2] = n_items;
  (void)pairs;
  return ok();
}

int pf_hash_join_write(int64_t nb, int64_t np, int32_t bits, int32_t multi, int64_t n_items,
                       const int64_t* const* bpay, int32_t npay, int64_t* out_probe,
                       int64_t* out_build, int64_t* const* pay_out, int64_t* host_out,
                       void* workspace, size_t workspace_bytes, void* stream) {
  using namespace pf;
  hipStream_t st = (hipStream_t)stream;
  if (npay < 0 || npay > MAX_PAY) return fail("hash_join: at most 2 payload columns");
  if (multi < 0 || multi > 2) return fail("hash_join: multi is 0, 1 or 2");
  const bool any_order = multi == 2;
  multi = multi == 1;
  HjWs w = carve_hj(workspace, nb, np, bits, multi != 0);
  if (workspace_bytes < w.bytes) return fail("hash_join: workspace too small");
  void* hdev = nullptr;
  if (any_order) {
    if (!host_out) return fail("hash_join: the any-order write reads its pair count back");
    if (host_alias(host_out, &hdev)) return 1;
    if (n_items <= 0) {
      host_out[0] = 0;
      return ok();
    }
    if (fill_i64(w.total, 1, 0, st)) return 1;
  }
  if (n_items <= 0) return ok();
  HjParams p = params_of(w, bits, multi != 0, n_items);
  p.next_out = w.total;
  p.out_probe = out_probe;
  p.out_build = out_build;
  p.npay = npay;
  for (int c = 0; c < npay; ++c) {
    {
      KTimer _kt("hj_gather_pay", st);
      hj_gather_pay<<<grid_for(nb, 256) < 16384 ? grid_for(nb, 256) : 16384, 256, 0, st>>>(
          bpay[c], w.brow, nb, w.pay_part[c]);
    }
    if (launch_check("hash_join payload")) return 1;
    p.pay_src[c] = w.pay_part[c];
    p.pay_out[c] = pay_out[c];
  }
  {
    KTimer _kt("hj_write", st);
    if (multi)
      hj_probe<true, 1><<<(unsigned)n_items, HJ_NT, 0, st>>>(p);
    else if (any_order)
      hj_probe<false, 2><<<(unsigned)n_items, HJ_NT, 0, st>>>(p);
    else
      hj_probe<false, 1><<<(unsigned)n_items, HJ_NT, 0, st>>>(p);
  }
  if (launch_check("hash_join write")) return 1;
  if (any_order) {
    hj_readback<<<1, 64, 0, st>>>(w.total, nullptr, (int64_t*)hdev);
    if (launch_check("hash_join write")) return 1;
    if (wait_stream(st, "hash_join write")) return 1;
  }
  return ok();
}

size_t pf_hash_join_residual_workspace_bytes(int64_t nb, int64_t np, int32_t bits, int32_t multi,
                                             int32_t probe_cols, int32_t build_cols) {
  if (bits < 0) bits = 0;
  if (bits > 12) bits = 12;
  const int nv = probe_cols < 0 ? 0 : probe_cols > 2 ? 2 : probe_cols;
  const int nbv = build_cols < 0 ? 0 : build_cols > 2 ? 2 : build_cols;
  const size_t base = pf::carve_hj(nullptr, nb, np, bits, multi == 1).bytes;
  return pf::carve_hr(nullptr, base, nb, np, bits, multi == 1, nv, nbv).bytes;
}

int pf_hash_join_residual_prepare(const int64_t* bkeys, int64_t nb, const int64_t* pkeys,
                                  int64_t np, int32_t bits, int32_t multi,
                                  const void* const* cols, const int32_t* col_types,
                                  const int32_t* col_sides, const int64_t* col_rows,
                                  int32_t n_cols, const int64_t* ins, int32_t n_ins,
                                  int64_t* host_out, void* workspace, size_t workspace_bytes,
                                  void* stream) {
  using namespace pf;
  hipStream_t st = (hipStream_t)stream;
  if (bits < 0 || bits > 12) return fail("hash_join: bits must be in [0, 12]");
  if (nb < 0 || np < 0 || nb >= (1LL << 31) || np >= (1LL << 31))
    return fail("hash_join: a side takes fewer than 2^31 rows");
  if (multi != 0 && multi != 1) return fail("hash_join residual: multi is 0 or 1");
  RfProgram P;
  if (rf_program_of(cols, col_types, col_sides, col_rows, n_cols, ins, n_ins, &P)) return 1;
  JoinResidual R;
  int nv, nbv, pslot[2], bslot[2];
  if (residual_plan(P, n_cols, nb, np, multi != 0, &R, &nv, &nbv, pslot, bslot)) return 1;
  HjWs w = carve_hj(workspace, nb, np, bits, multi != 0);
  HrWs r = carve_hr(workspace, w.bytes, nb, np, bits, multi != 0, nv, nbv);
  if (workspace_bytes < r.bytes) return fail("hash_join: workspace too small");
  void* hdev = nullptr;
  if (host_alias(host_out, &hdev)) return 1;
  ProbeCarry pc;
  pc.nv = nv;
  for (int c = 0; c < nv; ++c) {
    pc.src[c] = P.col[pslot[c]];
    pc.type[c] = P.type[pslot[c]];
    pc.dst[c] = r.px[c];
    pc.tx[c] = r.tx[c];
  }
  int64_t n_items = 0, largest = 0;
  const bool range = R.fast.range != 0;
  if (prepare_sides(bkeys, nb, pkeys, np, bits, multi != 0, w, pc, host_out, hdev, st, &n_items,
                    &largest, !range))
    return 1;
  const unsigned gg = grid_for(nb, 256) < 16384 ? grid_for(nb, 256) : 16384;
  if (range) {
    // the range form: buckets past the keyed sort by (hash, row), the build
    // operand's keys gathered, then the smaller buckets by (hash, key, row)
    bool all_sorted;
    if (sort_buckets(w, nb, bits, largest, KEYED_CAP, st, &all_sorted)) return 1;
    hj_gather_key<<<gg, 256, 0, st>>>(P.col[bslot[0]], P.type[bslot[0]], w.brow, nb,
                                      R.fast.f64, R.fast.bv_conv, r.bx[0]);
    if (launch_check("hash_join residual keys")) return 1;
    if (nb > 1) {
      KTimer _kt("hj_sort_keyed", st);
      int cap = 2;
      while (cap < largest && cap < KEYED_CAP) cap <<= 1;
      const size_t per = sizeof(uint64_t) + sizeof(int64_t) + sizeof(int32_t);
      if (hip_check(hipFuncSetAttribute((const void*)hj_sort_keyed,
                                        hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)(KEYED_CAP * per)),
                    "hash_join keyed sort LDS"))
        return 1;
      hj_sort_keyed<<<(unsigned)(1 << bits), SORT_NT, cap * per, st>>>(w.bh, r.bx[0], w.brow,
                                                                      w.boff, cap);
    }
    if (launch_check("hash_join keyed sort")) return 1;
  } else {
    // the build operands in the (sorted) partition order
    for (int c = 0; c < nbv; ++c) {
      hj_gather_res<<<gg, 256, 0, st>>>(P.col[bslot[c]], P.type[bslot[c]], w.brow, nb, r.bx[c]);
      if (launch_check("hash_join residual operands")) return 1;
    }
  }
  if (fill_i64(r.dev, 2, 0, st)) return 1;
  HjParams p = params_of(w, bits, multi != 0, n_items);
  R.bpart1 = nbv > 1 ? r.bx[1] : nullptr;
  HjRes x{{nv > 0 ? r.px[0] : nullptr, nv > 1 ? r.px[1] : nullptr},
          nbv > 0 ? r.bx[0] : nullptr, r.dev, r.kt};
  if (n_items > 0) {
    {
      KTimer _kt("hj_count_res", st);
      if (multi)
        launch_res<true, false>(n_items, p, R, x, st);
      else
        launch_res<false, false>(n_items, p, R, x, st);
    }
    if (launch_check("hash_join residual count")) return 1;
  }
  if (scan_i64_excl(w.item_cnt, n_items, w.total, w.scan, st)) return 1;
  hj_readback<<<1, 64, 0, st>>>(w.total, r.dev, (int64_t*)hdev);
  if (launch_check("hash_join residual count")) return 1;
  if (wait_stream(st, "hash_join")) return 1;
  // host_out: [0] kept pairs, [1] the largest build bucket, [2] work items,
  // [3] candidate pairs before the residual
  host_out[3] = host_out[2];
  host_out[1] = largest;
  host_out[2] = n_items;
  return ok();
}

int pf_hash_join_residual_write(int64_t nb, int64_t np, int32_t bits, int32_t multi,
                                int64_t n_items, const int64_t* const* bpay, int32_t npay,
                                int64_t* out_probe, int64_t* out_build, int64_t* const* pay_out,
                                const void* const* cols, const int32_t* col_types,
                                const int32_t* col_sides, const int64_t* col_rows,
                                int32_t n_cols, const int64_t* ins, int32_t n_ins,
                                void* workspace, size_t workspace_bytes, void* stream) {
  using namespace pf;
  hipStream_t st = (hipStream_t)stream;
  if (bits < 0 || bits > 12) return fail("hash_join: bits must be in [0, 12]");
  if (npay < 0 || npay > MAX_PAY) return fail("hash_join: at most 2 payload columns");
  if (multi != 0 && multi != 1) return fail("hash_join residual: multi is 0 or 1");
  RfProgram P;
  if (rf_program_of(cols, col_types, col_sides, col_rows, n_cols, ins, n_ins, &P)) return 1;
  JoinResidual R;
  int nv, nbv, pslot[2], bslot[2];
  if (residual_plan(P, n_cols, nb, np, multi != 0, &R, &nv, &nbv, pslot, bslot)) return 1;
  HjWs w = carve_hj(workspace, nb, np, bits, multi != 0);
  HrWs r = carve_hr(workspace, w.bytes, nb, np, bits, multi != 0, nv, nbv);
  if (workspace_bytes < r.bytes) return fail("hash_join: workspace too small");
  if (n_items <= 0) return ok();
  if (n_items > max_items(np, bits, multi != 0)) return fail("hash_join: work items overflow");
  HjParams p = params_of(w, bits, multi != 0, n_items);
  p.out_probe = out_probe;
  p.out_build = out_build;
  p.npay = npay;
  for (int c = 0; c < npay; ++c) {
    {
      KTimer _kt("hj_gather_pay", st);
      hj_gather_pay<<<grid_for(nb, 256) < 16384 ? grid_for(nb, 256) : 16384, 256, 0, st>>>(
          bpay[c], w.brow, nb, w.pay_part[c]);
    }
    if (launch_check("hash_join payload")) return 1;
    p.pay_src[c] = w.pay_part[c];
    p.pay_out[c] = pay_out[c];
  }
  R.bpart1 = nbv > 1 ? r.bx[1] : nullptr;
  HjRes x{{nv > 0 ? r.px[0] : nullptr, nv > 1 ? r.px[1] : nullptr},
          nbv > 0 ? r.bx[0] : nullptr, r.dev, r.kt};
  {
    KTimer _kt("hj_write_res", st);
    if (multi)
      launch_res<true, true>(n_items, p, R, x, st);
    else
      launch_res<false, true>(n_items, p, R, x, st);
  }
  if (launch_check("hash_join residual write")) return 1;
  return ok();
}

}  // extern "C"
