// Stable radix partition of (key, row id) pairs into up to 2^12 buckets: the
// partition of the batched and large-join tiers (one pass, no padding), and
// the per-digit pass of an LSD sort.
//
// A bucket is the top `bits` bits of mix64(key) below its top `shift` bits
// (hashed modes; the join's partition function generalised past the legacy
// 512 buckets), or bits [shift, shift + bits) of the unsigned key (raw mode).
// Rows keep their input order inside a bucket.
//
// gfx950 design.  Each workgroup owns a chunk of rows (at least 64 rows a
// bucket on average) and walks it in tiles: per tile, wave64 ballots rank
// rows inside their bucket (one ballot per digit bit plus validity), per-wave
// counts in LDS give each wave its start, the tile is staged in LDS in bucket
// order and each bucket's run is written at the chunk's running cursor for
// that bucket.  A chunk's cursors come from a device-wide scan of the chunk
// histograms laid out bucket-major, so the whole output is one stable
// partition and a bucket's runs from consecutive tiles of a chunk are
// contiguous in HBM.
//
//   rp_hist     per-chunk bucket counts (LDS atomics), hist[b][chunk]
//   scan        exclusive scan of hist in bucket-major order -> cursors
//   rp_scatter  per tile: rank, stage, write runs; advance the cursors
//   rp_offsets  each bucket's first row (and n) into bucket_offsets
#include <hip/hip_runtime.h>

#include <string>

#include "pf_common.hpp"
#include "pf_residual.hpp"
#include "../../include/pageflow_join.h"

namespace pf {

int radix_partition(const void* keys, int key_bytes, const int32_t* vals, int64_t n, int bits,
                    int shift, int mode, void* keys_out, int32_t* vals_out,
                    int64_t* bucket_offsets, void* ws, size_t ws_bytes, hipStream_t st,
                    int who, int shape);

namespace {

// RP_HASH32: the digit of fmix32 of an int32 key (the batch tier's slots),
// taken as the low 32 bits of a 64-bit value, and fmix32 of the key written
// in its place (a bijection: equal outputs are equal keys)
enum RpMode : int { RP_HASH_KEY = 0, RP_HASH_OUT = 1, RP_RAW = 2, RP_HASH32 = 3 };

// Up to two value columns that travel with the rows of a partition (the
// join's probe-side residual operands): src[c] is read at the row's input
// index as a PF_COL_* column and widened to 64 bits (rf_widen), dst[c]
// gets it at the row's output position.
struct Carry {
  const void* src[2];
  int32_t type[2];
  int64_t* dst[2];
};

template <typename K>
struct Unsigned;
template <>
struct Unsigned<int64_t> { typedef uint64_t T; };
template <>
struct Unsigned<int32_t> { typedef uint32_t T; };

// digit of a key and the value written in its place (the key, or its hash)
template <typename K, typename KO, int MODE>
__device__ __forceinline__ uint32_t rp_digit(K key, int bits, int shift, KO& outv) {
  if (MODE == RP_RAW) {
    const uint64_t u = (uint64_t)(typename Unsigned<K>::T)key;
    outv = (KO)key;
    return (uint32_t)((u >> shift) & ((1ULL << bits) - 1));
  }
  const uint64_t h = MODE == RP_HASH32 ? (uint64_t)fmix32((uint32_t)key)
                                       : mix64((uint64_t)(int64_t)key);
  outv = (MODE == RP_HASH_OUT || MODE == RP_HASH32) ? (KO)h : (KO)key;
  return bits ? (uint32_t)((h << shift) >> (64 - bits)) : 0u;
}

// The digit again from the value written in the key's place (the hash
// itself, or the raw key); RP_HASH_KEY would need the hash again.
template <typename KO, int MODE>
__device__ __forceinline__ uint32_t rp_digit_of_out(KO outv, int bits, int shift) {
  static_assert(MODE != RP_HASH_KEY, "the key would be hashed again");
  const uint64_t u = MODE == RP_HASH32 ? (uint64_t)(uint32_t)outv
                                       : (uint64_t)(typename Unsigned<KO>::T)outv;
  if (MODE == RP_RAW) return (uint32_t)((u >> shift) & ((1ULL << bits) - 1));
  return bits ? (uint32_t)((u << shift) >> (64 - bits)) : 0u;
}

// Rows of a chunk: at least 64 a bucket on average, a multiple of the tile.
int64_t chunk_rows(int bits, int tile) {
  int64_t c = (int64_t)64 << bits;
  if (c < 32768) c = 32768;
  return ((c + tile - 1) / tile) * tile;
}

template <int LOGB>
struct RpShape;
template <>
struct RpShape<8> {
  static constexpr int NT = 256, IT = 8;
};
template <>
struct RpShape<9> {
  static constexpr int NT = 256, IT = 8;
};
template <>
struct RpShape<12> {
  static constexpr int NT = 512, IT = 8;
};

// A chunk's histogram of the digit into hist: bucket-major
// (hist[b * nchunks + chunk], the layout the offsets scan reads) or, with
// CHUNK_MAJOR, the chunk's counts contiguous (hist[chunk << bits | b]).
template <typename K, typename KO, int MODE, int LOGB, bool CHUNK_MAJOR = false>
__device__ __forceinline__ void rp_hist_body(const K* __restrict__ keys, int64_t n,
                                             int64_t crows, int bits, int shift, int64_t nchunks,
                                             int32_t* __restrict__ hist, int64_t chunk) {
  constexpr int NT = RpShape<LOGB>::NT;
  __shared__ int32_t h[1 << LOGB];
  const int B = 1 << bits;
  for (int b = threadIdx.x; b < B; b += NT) h[b] = 0;
  __syncthreads();
  const int64_t c0 = chunk * crows;
  const int64_t c1 = c0 + crows < n ? c0 + crows : n;
  // U keys a thread in flight before their atomics
  constexpr int U = 8;
  for (int64_t b0 = c0 + threadIdx.x; b0 < c1; b0 += (int64_t)NT * U) {
    K k[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t i = b0 + (int64_t)u * NT;
      k[u] = i < c1 ? keys[i] : (K)0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      KO o;
      if (b0 + (int64_t)u * NT < c1) atomicAdd(&h[rp_digit<K, KO, MODE>(k[u], bits, shift, o)], 1);
    }
  }
  __syncthreads();
  for (int b = threadIdx.x; b < B; b += NT)
    hist[CHUNK_MAJOR ? (chunk << bits) + b : (int64_t)b * nchunks + chunk] = h[b];
}

template <typename K, typename KO, int MODE, int LOGB, bool CHUNK_MAJOR = false>
__global__ __launch_bounds__(RpShape<LOGB>::NT) void rp_hist(const K* __restrict__ keys,
                                                            int64_t n, int64_t crows, int bits,
                                                            int shift, int64_t nchunks,
                                                            int32_t* __restrict__ hist) {
  rp_hist_body<K, KO, MODE, LOGB, CHUNK_MAJOR>(keys, n, crows, bits, shift, nchunks, hist,
                                               blockIdx.x);
}

// The segment block `blk` works for: the last whose first block is <= it.
__device__ __forceinline__ int seg_of_block(const RpSeg* __restrict__ segs, int nseg,
                                            int64_t blk) {
  int a = 0, b = nseg;
  while (b - a > 1) {
    const int mid = (a + b) >> 1;
    if (segs[mid].blk_start <= blk)
      a = mid;
    else
      b = mid;
  }
  return a;
}

__global__ __launch_bounds__(RpShape<kSegMaxBits>::NT) void rp_hist_seg(const int32_t* __restrict__ keys,
                                                             const RpSeg* __restrict__ segs,
                                                             int nseg,
                                                             int32_t* __restrict__ hist) {
  const RpSeg& g = segs[seg_of_block(segs, nseg, blockIdx.x)];
  rp_hist_body<int32_t, int32_t, RP_HASH32, kSegMaxBits>(keys + g.in_off, g.n, kSegChunkRows, g.bits,
                                               g.shift, g.nchunks, hist + g.hist_off,
                                               (int64_t)blockIdx.x - g.blk_start);
}

// Lanes of the wave holding the same digit (d == ~0u: not a row).
template <int MAXB>
__device__ __forceinline__ uint64_t match_digit(uint32_t d, int bits) {
  // For each bit (and the valid flag), ballot(bit) XOR (all ones when this
  // lane's bit is set) has the lanes that differ from this lane in it; the
  // peers are the lanes that differ in none.  A bit costs a bit extract, a
  // ballot and an XOR-OR per half of the mask.
  const uint32_t inv = d == ~0u ? ~0u : 0u;
  const uint64_t bi = ballot(inv != 0);
  uint32_t lo = (uint32_t)bi ^ inv, hi = (uint32_t)(bi >> 32) ^ inv;
#pragma unroll
  for (int bit = 0; bit < MAXB; ++bit) {
    if (bit < bits) {  // uniform
      const uint32_t on = (uint32_t)((int32_t)(d << (31 - bit)) >> 31);  // bit -> 0 or ~0
      const uint64_t b = ballot(on != 0);
      lo |= (uint32_t)b ^ on;
      hi |= (uint32_t)(b >> 32) ^ on;
    }
  }
  return ~(((uint64_t)hi << 32) | lo);
}

template <typename K, typename KO, int MODE, int LOGB>
__global__ __launch_bounds__(RpShape<LOGB>::NT) void rp_scatter(
    const K* __restrict__ keys, const int32_t* __restrict__ vals, int64_t n, int64_t crows,
    int bits, int shift, int64_t nchunks, const int64_t* __restrict__ offs,
    KO* __restrict__ kout, int32_t* __restrict__ vout) {
  constexpr int NT = RpShape<LOGB>::NT, IT = RpShape<LOGB>::IT;
  constexpr int W = NT / kWave, TILE = NT * IT, BMAX = 1 << LOGB;
  constexpr int WROWS = kWave * IT;
  __shared__ uint16_t wh[W * BMAX];
  __shared__ uint16_t tstart[BMAX];
  __shared__ uint32_t cursor[BMAX];
  __shared__ KO stage_k[TILE];
  __shared__ int32_t stage_v[TILE];
  __shared__ uint16_t sb[TILE];
  __shared__ int32_t tmp[W + 1];

  const int tid = threadIdx.x, lane = lane_id(), w = tid / kWave;
  const int B = 1 << bits;
  const int64_t c0 = (int64_t)blockIdx.x * crows;
  const int64_t c1 = c0 + crows < n ? c0 + crows : n;
  for (int b = tid; b < B; b += NT) {
    cursor[b] = (uint32_t)offs[(int64_t)b * nchunks + blockIdx.x];
#pragma unroll
    for (int q = 0; q < W; ++q) wh[q * B + b] = 0;
  }
  __syncthreads();
  for (int64_t t0 = c0; t0 < c1; t0 += TILE) {
    uint32_t d[IT];
    KO kv[IT];
    int32_t vv[IT];
    const int64_t seg = t0 + (int64_t)w * WROWS;
#pragma unroll
    for (int k = 0; k < IT; ++k) {
      const int64_t i = seg + k * kWave + lane;
      if (i < c1) {
        d[k] = rp_digit<K, KO, MODE>(keys[i], bits, shift, kv[k]);
        vv[k] = vals ? vals[i] : (int32_t)i;
      } else {
        d[k] = ~0u;
      }
    }
    // stable rank inside (wave, bucket)
    int rk[IT];
#pragma unroll
    for (int k = 0; k < IT; ++k) {
      const uint64_t peers = match_digit<LOGB>(d[k], bits);
      const int below = mbcnt(peers);
      const bool valid = d[k] != ~0u;
      const int base = valid ? (int)wh[w * B + d[k]] : 0;
      rk[k] = base + below;
      if (valid && below == 0) wh[w * B + d[k]] = (uint16_t)(base + __popcll(peers));
    }
    __syncthreads();
    // per bucket: each wave's start in the tile's run; the run's length
    for (int b = tid; b < B; b += NT) {
      int run = 0;
#pragma unroll
      for (int q = 0; q < W; ++q) {
        const int t = wh[q * B + b];
        wh[q * B + b] = (uint16_t)run;
        run += t;
      }
      tstart[b] = (uint16_t)run;
    }
    __syncthreads();
    // the runs' starts in bucket order: B / NT consecutive buckets a thread
    {
      constexpr int PER = BMAX / NT > 0 ? BMAX / NT : 1;
      const int per = B >= NT ? B / NT : 1;
      const int b0 = tid * per;
      int loc[PER];
      int s = 0;
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        loc[q] = (q < per && b0 + q < B) ? tstart[b0 + q] : 0;
        s += loc[q];
      }
      int tot;
      int ex = block_excl_scan<NT>(s, tmp, &tot);
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        if (q < per && b0 + q < B) tstart[b0 + q] = (uint16_t)ex;
        ex += loc[q];
      }
    }
    __syncthreads();
    int tvalid = 0;
#pragma unroll
    for (int k = 0; k < IT; ++k) {
      if (d[k] != ~0u) {
        const int pos = tstart[d[k]] + wh[w * B + d[k]] + rk[k];
        stage_k[pos] = kv[k];
        stage_v[pos] = vv[k];
        sb[pos] = (uint16_t)d[k];
      }
    }
    {
      const int64_t left = c1 - t0;
      tvalid = left < TILE ? (int)left : TILE;
    }
    __syncthreads();
    for (int s = tid; s < tvalid; s += NT) {
      const int b = sb[s];
      const uint32_t o = cursor[b] + (uint32_t)(s - tstart[b]);
      kout[o] = stage_k[s];
      vout[o] = stage_v[s];
    }
    __syncthreads();
    for (int b = tid; b < B; b += NT) {
      const int next = b + 1 < B ? tstart[b + 1] : tvalid;
      cursor[b] += (uint32_t)(next - tstart[b]);
#pragma unroll
      for (int q = 0; q < W; ++q) wh[q * B + b] = 0;
    }
    __syncthreads();
  }
}

// Up to 256 buckets (512 for 4-byte keys): the same stable partition with
// the next tile's keys loaded while the current one is ranked (registers
// double-buffered), each bucket owned by one thread across the per-tile
// bookkeeping (wave prefix, tile scan, global base, cursor).  Rows [c0, c1)
// of `keys`; bucket b's first output row is offs_col[b * ostride] + delta.
// NV value columns of `cy` ride along.
//   LDS_MATCH  a wave's lanes of one bucket by LDS lane masks, else ballots
//   BMAX 512   9-bit digits: ballot ranking, and no staged bucket bytes (the
//              write-out takes each row's digit again from its staged hash)
//   ROW_IDS    vals is null: a row's id is its index, taken at staging
template <typename K, typename KO, int MODE, int NT, int IT, int NV = 0, bool LDS_MATCH = true,
          int BMAX = 256, bool ROW_IDS = false>
__device__ __forceinline__ void rp_scatter8_body(
    const K* __restrict__ keys, const int32_t* __restrict__ vals, int64_t c0, int64_t c1,
    int bits, int shift, const int64_t* __restrict__ offs_col, int64_t ostride,
    KO* __restrict__ kout, int32_t* __restrict__ vout, int64_t delta, const Carry& cy) {
  constexpr int W = NT / kWave, TILE = NT * IT, WROWS = kWave * IT;
  constexpr int XV = NV > 0 ? NV : 1, XT = NV > 0 ? TILE : 1;
  static_assert(BMAX <= NT, "a thread owns each bucket");
  static_assert(!ROW_IDS || sizeof(KO) == 4, "row ids from indexes: 4-byte keys (staged in pairs)");
  static_assert(BMAX == 256 || (BMAX == 512 && !LDS_MATCH && MODE != RP_HASH_KEY &&
                                 sizeof(KO) == 4 && NV == 0),
                "digits of 8 bits, or of 9 with ballots, 4-byte keys and a re-derivable digit");
  constexpr bool KEEP_B = BMAX == 256;  // the staged rows' buckets as bytes
  // wh: per (wave, bucket) counts while ranking, then the bucket's rows in
  // the waves before; tstart: the bucket's run in the tile.  A 4-byte key
  // and its row id are staged as one 8-byte word.
  constexpr bool PACK = sizeof(KO) == 4;
  __shared__ uint16_t wh[W * BMAX];
  __shared__ uint32_t gb[BMAX];
  __shared__ uint16_t tstart[BMAX];
  __shared__ uint64_t stage_p[PACK ? TILE : 1];
  __shared__ KO stage_k[PACK ? 1 : TILE];
  __shared__ int32_t stage_v[PACK ? 1 : TILE];
  __shared__ uint8_t sb[KEEP_B ? TILE : 1];
  __shared__ int64_t stage_x[XV][XT];
  __shared__ int32_t tmp[W + 1];
  // LDS_MATCH: each lane ORs its lane bit into the (wave, bucket) mask
  // (order-free, so the ranks stay stable) and reads it back; the bucket's
  // first lane clears it.  Three LDS operations a row, where a ballot a
  // digit bit (match_digit) costs four vector instructions a bit.  Which is
  // faster depends on the kernel (measured: LDS masks for 8-byte keys,
  // ballots for the 4-byte segmented partition).
  __shared__ uint64_t pm[LDS_MATCH ? W * BMAX : 1];

  const int tid = threadIdx.x, lane = lane_id(), w = tid / kWave;
  const int B = 1 << bits;
  uint32_t cursor = tid < B ? (uint32_t)(offs_col[(int64_t)tid * ostride] + delta) : 0u;
  for (int x = tid; x < W * B; x += NT) {
    wh[x] = 0;
    if (LDS_MATCH) pm[x] = 0;
  }

  K kn[IT];
  int32_t vn[IT];
  int64_t xn[XV][IT];
  // tiles start on a 64-row boundary (a wave's loads cover whole lines);
  // rows before c0 are loaded by no lane and ranked as no row
  const int64_t a0 = c0 & ~(int64_t)(kWave - 1);
  auto load_tile = [&](int64_t t0) {
    const int64_t seg = t0 + (int64_t)w * WROWS;
#pragma unroll
    for (int k = 0; k < IT; ++k) {
      const int64_t i = seg + k * kWave + lane;
      if (i >= c0 && i < c1) {
        kn[k] = keys[i];
        if (!ROW_IDS) vn[k] = vals ? vals[i] : (int32_t)i;
#pragma unroll
        for (int c = 0; c < NV; ++c) xn[c][k] = rf_widen(cy.src[c], cy.type[c], i);
      }
    }
  };
  // digits taken again from ov where that is cheap, not held in registers
  constexpr bool REDERIVE = MODE != RP_HASH_KEY;
  // the next tile's rows are loaded while this one is ranked (registers
  // double-buffered; the 4-byte 16-row tiles fit with their row ids taken
  // from indexes and their digits re-derived)
  load_tile(a0);
  __syncthreads();
  for (int64_t t0 = a0; t0 < c1; t0 += TILE) {
    K kc[IT];
    int32_t vc[IT];
    int64_t xc[XV][IT];
#pragma unroll
    for (int k = 0; k < IT; ++k) {
      kc[k] = kn[k];
      vc[k] = ROW_IDS ? 0 : vn[k];
#pragma unroll
      for (int c = 0; c < NV; ++c) xc[c][k] = xn[c][k];
    }
    if (t0 + TILE < c1) load_tile(t0 + TILE);
    const int64_t seg = t0 + (int64_t)w * WROWS;
    uint32_t dh[REDERIVE ? 1 : IT];
    KO ov[IT];
#pragma unroll
    for (int k = 0; k < IT; ++k) {
      const int64_t i = seg + k * kWave + lane;
      const uint32_t dk =
          (i >= c0 && i < c1) ? rp_digit<K, KO, MODE>(kc[k], bits, shift, ov[k]) : ~0u;
      if (!REDERIVE) dh[k] = dk;
    }
    // row k's digit, or ~0u past the rows
    auto dig = [&](int k) -> uint32_t {
      if (!REDERIVE) return dh[k];
      const int64_t i = seg + k * kWave + lane;
      return (i >= c0 && i < c1)
                 ? rp_digit_of_out<KO, (REDERIVE ? MODE : RP_RAW)>(ov[k], bits, shift)
                 : ~0u;
    };
    int rk[IT];
#pragma unroll
    for (int k = 0; k < IT; ++k) {
      const uint32_t dk = dig(k);
      const bool valid = dk != ~0u;
      const int at = w * B + (valid ? (int)dk : 0);
      uint64_t peers;
      if (LDS_MATCH) {
        if (valid) atomicOr((unsigned long long*)&pm[at], 1ULL << lane);
        peers = valid ? pm[at] : 0ULL;
      } else {
        peers = match_digit<(BMAX > 256 ? 9 : 8)>(dk, bits);
      }
      const int below = mbcnt(peers);
      const int base = valid ? (int)wh[at] : 0;
      rk[k] = base + below;
      if (valid && below == 0) {
        wh[at] = (uint16_t)(base + __popcll(peers));
        if (LDS_MATCH) pm[at] = 0;
      }
    }
    __syncthreads();
    // thread b: bucket b's wave starts, its tile run's start and global base
    int tot = 0;
    if (tid < B) {
#pragma unroll
      for (int q = 0; q < W; ++q) {
        const int t = wh[q * B + tid];
        wh[q * B + tid] = (uint16_t)tot;
        tot += t;
      }
    }
    {
      const int incl = wave_incl_scan(tot);
      if (lane == kWave - 1) tmp[w] = incl;
      __syncthreads();
      if (w == 0) {
        const int x = lane < W ? tmp[lane] : 0;
        const int xs = wave_incl_scan(x);
        if (lane < W) tmp[lane] = xs - x;
      }
      __syncthreads();
      if (tid < B) {
        const int ex = tmp[w] + incl - tot;
        gb[tid] = cursor - (uint32_t)ex;
        cursor += (uint32_t)tot;
        tstart[tid] = (uint16_t)ex;
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < IT; ++k) {
      const uint32_t dk = dig(k);
      if (dk != ~0u) {
        const int pos = tstart[dk] + wh[w * B + dk] + rk[k];
        if (PACK) {
          // ROW_IDS: the row id is the row's index (vals is null)
          const int32_t v = ROW_IDS ? (int32_t)(seg + k * kWave + lane) : vc[k];
          stage_p[pos] = ((uint64_t)(uint32_t)v << 32) | (uint32_t)ov[k];
        } else {
          stage_k[pos] = ov[k];
          stage_v[pos] = vc[k];
        }
        if (KEEP_B) sb[pos] = (uint8_t)dk;
#pragma unroll
        for (int c = 0; c < NV; ++c) stage_x[c][pos] = xc[c][k];
      }
    }
    // the tile's rows: those of [t0, t0 + TILE) inside [c0, c1)
    const int64_t lo_row = t0 > c0 ? t0 : c0;
    const int64_t hi_row = t0 + TILE < c1 ? t0 + TILE : c1;
    const int tvalid = (int)(hi_row - lo_row);
    __syncthreads();
    for (int s = tid; s < tvalid; s += NT) {
      const uint64_t kv = PACK ? stage_p[s] : 0;
      const uint32_t b = KEEP_B ? (uint32_t)sb[s]
                                : rp_digit_of_out<KO, (KEEP_B ? RP_RAW : MODE)>(
                                      (KO)(uint32_t)kv, bits, shift);
      const uint32_t o = gb[b] + (uint32_t)s;
      if (PACK) {
        kout[o] = (KO)(uint32_t)kv;
        vout[o] = (int32_t)(kv >> 32);
      } else {
        kout[o] = stage_k[s];
        vout[o] = stage_v[s];
      }
#pragma unroll
      for (int c = 0; c < NV; ++c) cy.dst[c][o] = stage_x[c][s];
    }
    for (int x = tid; x < W * B; x += NT) wh[x] = 0;
    __syncthreads();
  }
}

template <typename K, typename KO, int MODE, int NT, int IT, int NV = 0>
__global__ __launch_bounds__(NT) void rp_scatter8(
    const K* __restrict__ keys, const int32_t* __restrict__ vals, int64_t n, int64_t crows,
    int bits, int shift, int64_t nchunks, const int64_t* __restrict__ offs,
    KO* __restrict__ kout, int32_t* __restrict__ vout, Carry cy) {
  const int64_t c0 = (int64_t)blockIdx.x * crows;
  rp_scatter8_body<K, KO, MODE, NT, IT, NV>(keys, vals, c0, c0 + crows < n ? c0 + crows : n,
                                            bits, shift, offs + blockIdx.x, nchunks, kout, vout,
                                            0, cy);
}

// Digits of up to 9 bits (BMAX 512) in 16K-row tiles: a bucket's run a
// tile is 32 rows or more; row ids are the rows' indexes.
template <int NT, int IT, int BMAX>
__global__ __launch_bounds__(NT) void rp_scatter8_seg(const int32_t* __restrict__ keys,
                                                      const RpSeg* __restrict__ segs, int nseg,
                                                      const int64_t* __restrict__ offs,
                                                      int32_t* __restrict__ kout,
                                                      int32_t* __restrict__ vout) {
  const RpSeg& g = segs[seg_of_block(segs, nseg, blockIdx.x)];
  const int64_t chunk = (int64_t)blockIdx.x - g.blk_start;
  const int64_t c0 = chunk * kSegChunkRows;
  rp_scatter8_body<int32_t, int32_t, RP_HASH32, NT, IT, 0, false, BMAX, true>(
      keys + g.in_off, nullptr, c0, c0 + kSegChunkRows < g.n ? c0 + kSegChunkRows : g.n, g.bits,
      g.shift, offs + g.hist_off + chunk, g.nchunks, kout, vout, g.delta, Carry{});
}

__global__ void rp_offsets(const int64_t* __restrict__ offs, int64_t nchunks, int B, int64_t n,
                           int64_t* __restrict__ bucket_offsets) {
  for (int b = blockIdx.x * blockDim.x + threadIdx.x; b <= B; b += gridDim.x * blockDim.x)
    bucket_offsets[b] = b < B ? offs[(int64_t)b * nchunks] : n;
}

struct RpWs {
  int32_t* hist;
  int64_t* offs;
  void* scan;
  size_t bytes;
};

// Shapes of the <= 8-bit scatter (threads, rows a thread per tile): 0..3
// and 5; the default is the measured best on MI355X, a caller (the
// autotuner) may pick another.  8192-row tiles give each bucket twice the
// rows a tile of 512 x 8 (longer runs per write) but take 96-114 KB of LDS,
// one workgroup a CU: faster from 64M rows up, slower on a few hundred
// chunks (10M rows: 0.07 -> 0.10 ms), where 512 x 8 tiles keep several
// workgroups a CU.  Of the 8192-row tiles, 1024 x 8 (16 waves to hide the
// loads) beats 512 x 16: 6-bit passes over 640M rows 4.18 (512 x 8) ->
// 4.00 (512 x 16) -> 3.71 ms (1024 x 8); 400M: 2.57 -> 2.43 -> 2.25 ms.
// shape: 0 = 256 x 8, 1 = 256 x 16, 2 = 512 x 8, 3 = 512 x 16, 5 = 1024 x 8
constexpr int kShape1k = 5;                     // 1024 threads x 8 rows
constexpr int64_t kLargeTileRows = 64LL << 20;  // rows from which 8192-row tiles win
// up to 6-bit digits 1024 x 8; 7-8 bits 512 x 16, whose waves see twice the
// rows a bucket (8 bits over 640M rows: 5.05 ms against 5.15 for 1024 x 8)
int default_shape8(int64_t n, int bits) {
  return n < kLargeTileRows ? 2 : bits <= 6 ? kShape1k : 3;
}
bool shape8_valid(int shape) { return (shape >= 0 && shape <= 3) || shape == kShape1k; }

// The largest tile of any shape fixes the chunk rows, so one workspace size
// serves every shape.
// From 64M rows a chunk is at least 128K rows (16 or more 8K-row tiles a
// workgroup, its start spread over them; still thousands of chunks).
int64_t chunk_rows_of(int64_t n, int bits) {
  const int64_t c = chunk_rows(bits, 8192);
  return n >= (64LL << 20) && c < 131072 ? 131072 : c;
}

RpWs carve_rp(void* ws, int64_t n, int bits) {
  Carve cv(ws);
  RpWs r;
  const int64_t crows = chunk_rows_of(n, bits);
  const int64_t nchunks = n > 0 ? (n + crows - 1) / crows : 1;
  const int64_t entries = nchunks << bits;
  r.hist = cv.take<int32_t>(entries);
  r.offs = cv.take<int64_t>(entries);
  r.scan = cv.take<char>(scan_i32_scratch(entries));
  r.bytes = cv.bytes();
  return r;
}

// kernel names for the profiler, by caller: the C entry point, the join's
// build side, the join's probe side, the zero-copy gather's index sort
const char* const kHistName[6] = {"rp_hist", "hj_build_hist", "hj_probe_hist", "zc_sort_hist",
                                  "hj_bigsort_hist", "gk_sort_hist"};
const char* const kScatterName[6] = {"rp_scatter", "hj_build_scatter", "hj_probe_scatter",
                                     "zc_sort_scatter", "hj_bigsort_scatter", "gk_sort_scatter"};
// the second pass of a two-level partition
const char* const kScatter2Name[6] = {"rp_scatter_2", "hj_build_scatter_2", "hj_probe_scatter_2",
                                      "zc_sort_scatter_2", "hj_bigsort_scatter_2",
                                      "gk_sort_scatter_2"};

template <typename K, typename KO, int MODE, int NT, int IT>
void launch8(const K* keys, const int32_t* vals, int64_t n, int64_t crows, int bits, int shift,
             int64_t nchunks, const int64_t* offs, KO* kout, int32_t* vout, hipStream_t st) {
  rp_scatter8<K, KO, MODE, NT, IT><<<(unsigned)nchunks, NT, 0, st>>>(
      keys, vals, n, crows, bits, shift, nchunks, offs, kout, vout, Carry{});
}

template <typename K, typename KO, int MODE, int LOGB>
int rp_launch(const K* keys, const int32_t* vals, int64_t n, int bits, int shift, KO* kout,
              int32_t* vout, int64_t* bucket_offsets, const RpWs& r, hipStream_t st,
              int who, int shape) {
  const int64_t crows = chunk_rows_of(n, bits);
  const int64_t nchunks = n > 0 ? (n + crows - 1) / crows : 1;
  const int64_t entries = nchunks << bits;
  {
    KTimer _kt(kHistName[who], st);
    rp_hist<K, KO, MODE, LOGB><<<(unsigned)nchunks, RpShape<LOGB>::NT, 0, st>>>(
        keys, n, crows, bits, shift, nchunks, r.hist);
  }
  if (launch_check("radix_partition hist")) return 1;
  if (scan_i32_to_i64_excl(r.hist, r.offs, entries, nullptr, r.scan, st)) return 1;
  if (n > 0) {
    KTimer _kt(kScatterName[who], st);
    if (LOGB > 8) {
      rp_scatter<K, KO, MODE, 12><<<(unsigned)nchunks, RpShape<12>::NT, 0, st>>>(
          keys, vals, n, crows, bits, shift, nchunks, r.offs, kout, vout);
    } else {
      switch (shape8_valid(shape) ? shape : default_shape8(n, bits)) {
        case 0: launch8<K, KO, MODE, 256, 8>(keys, vals, n, crows, bits, shift, nchunks, r.offs,
                                             kout, vout, st); break;
        case 1: launch8<K, KO, MODE, 256, 16>(keys, vals, n, crows, bits, shift, nchunks,
                                              r.offs, kout, vout, st); break;
        case 2: launch8<K, KO, MODE, 512, 8>(keys, vals, n, crows, bits, shift, nchunks, r.offs,
                                             kout, vout, st); break;
        case 3: launch8<K, KO, MODE, 512, 16>(keys, vals, n, crows, bits, shift, nchunks,
                                              r.offs, kout, vout, st); break;
        default: launch8<K, KO, MODE, 1024, 8>(keys, vals, n, crows, bits, shift, nchunks,
                                               r.offs, kout, vout, st); break;
      }
    }
  }
  if (launch_check("radix_partition scatter")) return 1;
  rp_offsets<<<grid_for((1 << bits) + 1, 256), 256, 0, st>>>(r.offs, nchunks, 1 << bits, n,
                                                             bucket_offsets);
  return launch_check("radix_partition offsets");
}

template <typename K>
int rp_dispatch(const K* keys, const int32_t* vals, int64_t n, int bits, int shift, int mode,
                void* kout, int32_t* vout, int64_t* bucket_offsets, const RpWs& r,
                hipStream_t st, int who, int shape) {
  if (mode == RP_HASH_OUT) {
    if (bits <= 8)
      return rp_launch<K, int64_t, RP_HASH_OUT, 8>(keys, vals, n, bits, shift, (int64_t*)kout,
                                                   vout, bucket_offsets, r, st, who, shape);
    return rp_launch<K, int64_t, RP_HASH_OUT, 12>(keys, vals, n, bits, shift, (int64_t*)kout,
                                                  vout, bucket_offsets, r, st, who, shape);
  }
  if (mode == RP_HASH_KEY) {
    if (bits <= 8)
      return rp_launch<K, K, RP_HASH_KEY, 8>(keys, vals, n, bits, shift, (K*)kout, vout,
                                             bucket_offsets, r, st, who, shape);
    return rp_launch<K, K, RP_HASH_KEY, 12>(keys, vals, n, bits, shift, (K*)kout, vout,
                                            bucket_offsets, r, st, who, shape);
  }
  if (bits <= 8)
    return rp_launch<K, K, RP_RAW, 8>(keys, vals, n, bits, shift, (K*)kout, vout,
                                      bucket_offsets, r, st, who, shape);
  return rp_launch<K, K, RP_RAW, 12>(keys, vals, n, bits, shift, (K*)kout, vout,
                                     bucket_offsets, r, st, who, shape);
}

// Bucket bounds of a partitioned output by binary search on the full
// digit: bucket b's first row (b = 2^bits: n).  KO is the output key type;
// HASH_OUT outputs hold mix64 already.
template <typename KO, int MODE>
__global__ void rp_bounds(const KO* __restrict__ kout, int64_t n, int bits, int shift,
                          int64_t* __restrict__ offs) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b > (1 << bits)) return;
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    uint32_t d;
    if (MODE == RP_HASH_OUT) {
      const uint64_t h = (uint64_t)kout[mid];
      d = bits ? (uint32_t)((h << shift) >> (64 - bits)) : 0u;
    } else {
      KO unused;
      d = rp_digit<KO, KO, MODE>(kout[mid], bits, shift, unused);
    }
    if ((int)d < b)
      lo = mid + 1;
    else
      hi = mid;
  }
  offs[b] = lo;
}

constexpr int kOnePassMax = 8;   // wider digits: two stable passes by default
constexpr int kShapeOnePass = 4; // shape: one pass whatever the width

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

size_t two_pass_ws(int64_t n, int bits) {
  const int hi = bits / 2, lo = bits - hi;
  size_t a = carve_rp(nullptr, n, lo).bytes, b = carve_rp(nullptr, n, hi).bytes;
  return align256(a > b ? a : b) + align256((size_t)n * 8) + align256((size_t)n * 4);
}

template <typename KO, int MODE>
int launch_bounds_of(const void* kout, int64_t n, int bits, int shift, int64_t* offs,
                     hipStream_t st) {
  rp_bounds<KO, MODE><<<grid_for((1 << bits) + 1, 256), 256, 0, st>>>((const KO*)kout, n, bits,
                                                                        shift, offs);
  return launch_check("radix_partition bounds");
}

// Two stable LSD passes, the low half of the digit first: the output is
// ordered by the whole digit with each bucket's rows in input order, as
// one pass would leave it, at two well-coalesced scatters instead of one
// scatter over 2^bits runs.
int two_pass(const void* keys, int key_bytes, const int32_t* vals, int64_t n, int bits,
             int shift, int mode, void* keys_out, int32_t* vals_out, int64_t* offs, void* ws,
             hipStream_t st, int who, int shape) {
  const int hi = bits / 2, lo = bits - hi;
  const size_t wsb = two_pass_ws(n, bits);
  char* base = (char*)ws;
  const size_t hb = wsb - align256((size_t)n * 8) - align256((size_t)n * 4);
  void* tk = base + hb;
  int32_t* tv = (int32_t*)(base + hb + align256((size_t)n * 8));
  const int sh = shape8_valid(shape) ? shape : -1;
  if (mode == RP_RAW) {
    if (radix_partition(keys, key_bytes, vals, n, lo, shift, RP_RAW, tk, tv, offs, ws, hb, st,
                        who, sh) ||
        radix_partition(tk, key_bytes, tv, n, hi, shift + lo, RP_RAW, keys_out, vals_out, offs,
                        ws, hb, st, who, sh))
      return 1;
    return key_bytes == 8
               ? launch_bounds_of<int64_t, RP_RAW>(keys_out, n, bits, shift, offs, st)
               : launch_bounds_of<int32_t, RP_RAW>(keys_out, n, bits, shift, offs, st);
  }
  if (mode == RP_HASH_KEY) {
    if (radix_partition(keys, key_bytes, vals, n, lo, shift + hi, RP_HASH_KEY, tk, tv, offs, ws,
                        hb, st, who, sh) ||
        radix_partition(tk, key_bytes, tv, n, hi, shift, RP_HASH_KEY, keys_out, vals_out, offs,
                        ws, hb, st, who, sh))
      return 1;
    return key_bytes == 8
               ? launch_bounds_of<int64_t, RP_HASH_KEY>(keys_out, n, bits, shift, offs, st)
               : launch_bounds_of<int32_t, RP_HASH_KEY>(keys_out, n, bits, shift, offs, st);
  }
  // HASH_OUT: the first pass writes the hashes, the second reads their bits
  if (radix_partition(keys, key_bytes, vals, n, lo, shift + hi, RP_HASH_OUT, tk, tv, offs, ws,
                      hb, st, who, sh) ||
      radix_partition(tk, 8, tv, n, hi, 64 - shift - hi, RP_RAW, keys_out, vals_out, offs, ws,
                      hb, st, who, sh))
    return 1;
  return launch_bounds_of<int64_t, RP_HASH_OUT>(keys_out, n, bits, shift, offs, st);
}

// ------------------------------------------------- two-level partition --
//
// A digit of 9-12 bits in two stable passes (the low part L of the digit,
// then the high part H), both fed by ONE histogram read of the keys: each
// chunk's histogram of the whole digit.  The first pass's offsets are its
// sums over H; the second pass works on units of first-pass output (the
// rows of one L from a group of consecutive chunks), so a unit's count of
// each H, and with them its output offsets, are sums of the same
// histogram.  The second pass reads no histogram of its own, and the
// bucket bounds come from the offsets (no search).

// Chunk rows of the histogram and the first pass: about 2048 chunks,
// a whole number of 8192-row tiles, 8192 to 131072 rows.
#ifndef PF_LEVEL_CHUNK_MAX
#define PF_LEVEL_CHUNK_MAX 131072
#endif
int64_t level_chunk_rows(int64_t n) {
  int64_t c = (n / 2048 + 8191) / 8192 * 8192;
  return c < 8192 ? 8192 : (c > PF_LEVEL_CHUNK_MAX ? PF_LEVEL_CHUNK_MAX : c);
}

struct LvPlan {
  int lo, hi;          // digit bits of the two passes
  int64_t crows, nc;   // chunk rows, chunks
  int group;           // chunks a second-pass unit spans
  int64_t ng;          // chunk groups
};

LvPlan level_plan(int64_t n, int bits) {
  LvPlan p;
  p.hi = bits / 2;
  p.lo = bits - p.hi;
  p.crows = level_chunk_rows(n);
  p.nc = n > 0 ? (n + p.crows - 1) / p.crows : 1;
  // about 32768 rows a unit: as long a run a bucket as a one-pass chunk has
  int64_t grp = (32768LL << p.lo) / p.crows;
  p.group = (int)(grp < 1 ? 1 : grp);
  p.ng = (p.nc + p.group - 1) / p.group;
  return p;
}

struct LvWs {
  int32_t* J;      // [chunk][digit]
  int32_t* cnt1;   // [L][chunk]
  int64_t* offs1;
  int32_t* cnt2;   // [digit][group]
  int64_t* offs2;
  void* scan;
  size_t bytes;
};

LvWs carve_lv(void* ws, int64_t n, int bits) {
  const LvPlan p = level_plan(n, bits);
  Carve cv(ws);
  LvWs w;
  const int64_t e1 = (1LL << p.lo) * p.nc, e2 = (1LL << bits) * p.ng;
  w.J = cv.take<int32_t>((1LL << bits) * p.nc);
  w.cnt1 = cv.take<int32_t>(e1);
  w.offs1 = cv.take<int64_t>(e1);
  w.cnt2 = cv.take<int32_t>(e2);
  w.offs2 = cv.take<int64_t>(e2);
  w.scan = cv.take<char>(scan_i32_scratch(e1 > e2 ? e1 : e2));
  w.bytes = cv.bytes();
  return w;
}

// J[c << bits | d]: chunk c's rows of digit d (d = H << lo | L).
// cnt1[L * nc + c]: chunk c's rows of low part L; cnt2[d * ng + g]: group
// g's rows of digit d.
__global__ void rp_level_counts(const int32_t* __restrict__ J, int64_t nc, int lo, int hi,
                                int64_t ng, int group, int32_t* __restrict__ cnt1,
                                int32_t* __restrict__ cnt2) {
  const int64_t n1 = (1LL << lo) * nc, n2 = (1LL << (lo + hi)) * ng;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < n1 + n2; x += stride) {
    // consecutive threads read consecutive counters of J
    int32_t s = 0;
    if (x < n1) {
      const int64_t c = x >> lo, L = x & ((1LL << lo) - 1);
      const int32_t* row = J + (c << (lo + hi));
      for (int64_t H = 0; H < (1LL << hi); ++H) s += row[(H << lo) | L];
      cnt1[L * nc + c] = s;
    } else {
      const int64_t y = x - n1, g = y >> (lo + hi), d = y & ((1LL << (lo + hi)) - 1);
      const int64_t a = g * group, b = a + group < nc ? a + group : nc;
      for (int64_t c = a; c < b; ++c) s += J[(c << (lo + hi)) + d];
      cnt2[d * ng + g] = s;
    }
  }
}

// Second pass: unit u = (L, g) is the first-pass output rows of low part
// L from chunks [g * group, (g + 1) * group); its bucket H starts at
// offs2[(H << lo | L) * ng + g].
template <typename K, typename KO, int MODE, int NT, int IT, int NV = 0>
__global__ __launch_bounds__(NT) void rp_scatter8_units(
    const K* __restrict__ keys, const int32_t* __restrict__ vals, int64_t n, int lo, int64_t nc,
    int64_t ng, int group, const int64_t* __restrict__ offs1, int bits, int shift,
    const int64_t* __restrict__ offs2, KO* __restrict__ kout, int32_t* __restrict__ vout,
    Carry cy) {
  const int64_t u = blockIdx.x;
  const int64_t L = u / ng, g = u - L * ng;
  const int64_t c0 = offs1[L * nc + g * group];
  const int64_t c1 = (g + 1) * group < nc ? offs1[L * nc + (g + 1) * group]
                                          : (L + 1 < (1LL << lo) ? offs1[(L + 1) * nc] : n);
  rp_scatter8_body<K, KO, MODE, NT, IT, NV>(keys, vals, c0, c1, bits, shift, offs2 + u,
                                            (1LL << lo) * ng, kout, vout, 0, cy);
}

__global__ void rp_level_offsets(const int64_t* __restrict__ offs2, int64_t ng, int P,
                                 int64_t n, int64_t* __restrict__ boff) {
  for (int d = blockIdx.x * blockDim.x + threadIdx.x; d <= P; d += gridDim.x * blockDim.x)
    boff[d] = d < P ? offs2[(int64_t)d * ng] : n;
}

template <typename K, typename KO, int MODE, int NT, int IT>
void launch_units(const K* keys, const int32_t* vals, int64_t n, const LvPlan& p,
                  const int64_t* offs1, int bits, int shift, const int64_t* offs2, KO* kout,
                  int32_t* vout, hipStream_t st) {
  rp_scatter8_units<K, KO, MODE, NT, IT><<<(unsigned)((1LL << p.lo) * p.ng), NT, 0, st>>>(
      keys, vals, n, p.lo, p.nc, p.ng, p.group, offs1, bits, shift, offs2, kout, vout,
      Carry{});
}

template <typename K, typename KO, int MODE>
void units_by_shape(int shape, const K* keys, const int32_t* vals, int64_t n, const LvPlan& p,
                    const int64_t* offs1, int bits, int shift, const int64_t* offs2, KO* kout,
                    int32_t* vout, hipStream_t st) {
  switch (shape) {
    case 0: launch_units<K, KO, MODE, 256, 8>(keys, vals, n, p, offs1, bits, shift, offs2, kout, vout, st); break;
    case 1: launch_units<K, KO, MODE, 256, 16>(keys, vals, n, p, offs1, bits, shift, offs2, kout, vout, st); break;
    case 2: launch_units<K, KO, MODE, 512, 8>(keys, vals, n, p, offs1, bits, shift, offs2, kout, vout, st); break;
    case 3: launch_units<K, KO, MODE, 512, 16>(keys, vals, n, p, offs1, bits, shift, offs2, kout, vout, st); break;
    default: launch_units<K, KO, MODE, 1024, 8>(keys, vals, n, p, offs1, bits, shift, offs2, kout, vout, st); break;
  }
}

template <typename K, typename KO, int MODE>
void pass1_by_shape(int shape, const K* keys, const int32_t* vals, int64_t n, const LvPlan& p,
                    int shift, const int64_t* offs1, KO* kout, int32_t* vout, hipStream_t st) {
  switch (shape) {
    case 0: launch8<K, KO, MODE, 256, 8>(keys, vals, n, p.crows, p.lo, shift, p.nc, offs1, kout, vout, st); break;
    case 1: launch8<K, KO, MODE, 256, 16>(keys, vals, n, p.crows, p.lo, shift, p.nc, offs1, kout, vout, st); break;
    case 2: launch8<K, KO, MODE, 512, 8>(keys, vals, n, p.crows, p.lo, shift, p.nc, offs1, kout, vout, st); break;
    case 3: launch8<K, KO, MODE, 512, 16>(keys, vals, n, p.crows, p.lo, shift, p.nc, offs1, kout, vout, st); break;
    default: launch8<K, KO, MODE, 1024, 8>(keys, vals, n, p.crows, p.lo, shift, p.nc, offs1, kout, vout, st); break;
  }
}

// The whole two-level partition of n keys by `bits` (9-12) bits of the
// digit `mode` selects below `shift` (the hashed modes) or from `shift`
// up (raw); pass 1 writes (tk, tv), pass 2 (keys_out, vals_out).
template <typename K>
int level_run(const K* keys, const int32_t* vals, int64_t n, int bits, int shift, int mode,
              void* tk, int32_t* tv, void* keys_out, int32_t* vals_out, int64_t* boff,
              const LvWs& w, hipStream_t st, int who, int shape) {
  const LvPlan p = level_plan(n, bits);
  const int P = 1 << bits;
  {
    KTimer _kt(kHistName[who], st);
    if (mode == RP_RAW)
      rp_hist<K, K, RP_RAW, 12, true><<<(unsigned)p.nc, RpShape<12>::NT, 0, st>>>(
          keys, n, p.crows, bits, shift, p.nc, w.J);
    else
      rp_hist<K, K, RP_HASH_KEY, 12, true><<<(unsigned)p.nc, RpShape<12>::NT, 0, st>>>(
          keys, n, p.crows, bits, shift, p.nc, w.J);
  }
  if (launch_check("radix_partition level hist")) return 1;
  const int64_t e1 = (1LL << p.lo) * p.nc, e2 = (int64_t)P * p.ng;
  rp_level_counts<<<grid_for(e1 + e2, 256) < 8192 ? grid_for(e1 + e2, 256) : 8192, 256, 0, st>>>(
      w.J, p.nc, p.lo, p.hi, p.ng, p.group, w.cnt1, w.cnt2);
  if (launch_check("radix_partition level counts")) return 1;
  if (scan_i32_to_i64_excl(w.cnt1, w.offs1, e1, nullptr, w.scan, st) ||
      scan_i32_to_i64_excl(w.cnt2, w.offs2, e2, nullptr, w.scan, st))
    return 1;
  if (n > 0) {
    const int sh1 = shape8_valid(shape) ? shape : default_shape8(n, p.lo);
    const int sh2 = shape8_valid(shape) ? shape : default_shape8(n, p.hi);
    {
      KTimer _kt(kScatterName[who], st);
      if (mode == RP_RAW)
        pass1_by_shape<K, K, RP_RAW>(sh1, keys, vals, n, p, shift, w.offs1, (K*)tk, tv, st);
      else if (mode == RP_HASH_KEY)
        pass1_by_shape<K, K, RP_HASH_KEY>(sh1, keys, vals, n, p, shift + p.hi, w.offs1, (K*)tk,
                                          tv, st);
      else  // HASH_OUT: the first pass writes the hashes, the second reads their bits
        pass1_by_shape<K, int64_t, RP_HASH_OUT>(sh1, keys, vals, n, p, shift + p.hi, w.offs1,
                                                (int64_t*)tk, tv, st);
    }
    KTimer _kt(kScatter2Name[who], st);
    if (mode == RP_RAW)
      units_by_shape<K, K, RP_RAW>(sh2, (const K*)tk, tv, n, p, w.offs1, p.hi, shift + p.lo,
                                   w.offs2, (K*)keys_out, vals_out, st);
    else if (mode == RP_HASH_KEY)
      units_by_shape<K, K, RP_HASH_KEY>(sh2, (const K*)tk, tv, n, p, w.offs1, p.hi, shift,
                                        w.offs2, (K*)keys_out, vals_out, st);
    else
      units_by_shape<int64_t, int64_t, RP_RAW>(sh2, (const int64_t*)tk, tv, n, p, w.offs1, p.hi,
                                               64 - shift - p.hi, w.offs2,
                                               (int64_t*)keys_out, vals_out, st);
  }
  if (launch_check("radix_partition level scatter")) return 1;
  rp_level_offsets<<<grid_for(P + 1, 256), 256, 0, st>>>(w.offs2, p.ng, P, n, boff);
  return launch_check("radix_partition level offsets");
}

size_t level_ws(int64_t n, int bits) {
  return align256(carve_lv(nullptr, n, bits).bytes) + align256((size_t)(n > 0 ? n : 1) * 8) +
         align256((size_t)(n > 0 ? n : 1) * 4);
}

// ------------------------------------------- partition with carried values --
//
// The join's probe side when residual operands ride along: the hashed-out
// partition (mix64(key), row id) with NV value columns.  A tile of 1024 x 4
// rows keeps the staged values (8 B a row each) within the LDS next to the
// hashes and row ids.
constexpr int kCarryNT = 1024, kCarryIT = 4;

template <int NV>
int carry_run(const int64_t* keys, int64_t n, int bits, const Carry& first, const Carry& second,
              uint64_t* tk, int32_t* tv, uint64_t* h_out, int32_t* row_out, int64_t* boff,
              void* ws, hipStream_t st, int who) {
  if (bits <= kOnePassMax) {
    RpWs r = carve_rp(ws, n, bits);
    const int64_t crows = chunk_rows_of(n, bits);
    const int64_t nchunks = n > 0 ? (n + crows - 1) / crows : 1;
    {
      KTimer _kt(kHistName[who], st);
      rp_hist<int64_t, int64_t, RP_HASH_KEY, 8><<<(unsigned)nchunks, RpShape<8>::NT, 0, st>>>(
          keys, n, crows, bits, 0, nchunks, r.hist);
    }
    if (launch_check("radix_partition hist")) return 1;
    if (scan_i32_to_i64_excl(r.hist, r.offs, nchunks << bits, nullptr, r.scan, st)) return 1;
    if (n > 0) {
      KTimer _kt(kScatterName[who], st);
      rp_scatter8<int64_t, int64_t, RP_HASH_OUT, kCarryNT, kCarryIT, NV>
          <<<(unsigned)nchunks, kCarryNT, 0, st>>>(keys, nullptr, n, crows, bits, 0, nchunks,
                                                   r.offs, (int64_t*)h_out, row_out, first);
    }
    if (launch_check("radix_partition scatter")) return 1;
    rp_offsets<<<grid_for((1 << bits) + 1, 256), 256, 0, st>>>(r.offs, nchunks, 1 << bits, n,
                                                               boff);
    return launch_check("radix_partition offsets");
  }
  const LvPlan p = level_plan(n, bits);
  const int P = 1 << bits;
  LvWs w = carve_lv(ws, n, bits);
  {
    KTimer _kt(kHistName[who], st);
    rp_hist<int64_t, int64_t, RP_HASH_KEY, 12, true><<<(unsigned)p.nc, RpShape<12>::NT, 0, st>>>(
        keys, n, p.crows, bits, 0, p.nc, w.J);
  }
  if (launch_check("radix_partition level hist")) return 1;
  const int64_t e1 = (1LL << p.lo) * p.nc, e2 = (int64_t)P * p.ng;
  rp_level_counts<<<grid_for(e1 + e2, 256) < 8192 ? grid_for(e1 + e2, 256) : 8192, 256, 0, st>>>(
      w.J, p.nc, p.lo, p.hi, p.ng, p.group, w.cnt1, w.cnt2);
  if (launch_check("radix_partition level counts")) return 1;
  if (scan_i32_to_i64_excl(w.cnt1, w.offs1, e1, nullptr, w.scan, st) ||
      scan_i32_to_i64_excl(w.cnt2, w.offs2, e2, nullptr, w.scan, st))
    return 1;
  if (n > 0) {
    {
      KTimer _kt(kScatterName[who], st);
      rp_scatter8<int64_t, int64_t, RP_HASH_OUT, kCarryNT, kCarryIT, NV>
          <<<(unsigned)p.nc, kCarryNT, 0, st>>>(keys, nullptr, n, p.crows, p.lo, p.hi, p.nc,
                                                w.offs1, (int64_t*)tk, tv, first);
    }
    KTimer _kt(kScatter2Name[who], st);
    rp_scatter8_units<int64_t, int64_t, RP_RAW, kCarryNT, kCarryIT, NV>
        <<<(unsigned)((1LL << p.lo) * p.ng), kCarryNT, 0, st>>>(
            (const int64_t*)tk, tv, n, p.lo, p.nc, p.ng, p.group, w.offs1, p.hi, 64 - p.hi,
            w.offs2, (int64_t*)h_out, row_out, second);
  }
  if (launch_check("radix_partition level scatter")) return 1;
  rp_level_offsets<<<grid_for(P + 1, 256), 256, 0, st>>>(w.offs2, p.ng, P, n, boff);
  return launch_check("radix_partition level offsets");
}

}  // namespace

// The join's probe side with nv (1 or 2) value columns carried: rows as
// (mix64(key), row id) by the top `bits` bits of the hash, stable, each
// row's values src[c] (PF_COL_* type[c], widened to 64 bits) into dst[c]
// at its output row.  Two-level partitions (9-12 bits) write their first
// pass into (tk, tv, tx).  ws: radix_partition_ws(n, bits) for bits <= 8,
// else radix_level_ws(n, bits).
int radix_partition_carry(const int64_t* keys, int64_t n, int bits, int nv,
                          const void* const* src, const int32_t* type, int64_t* const* dst,
                          uint64_t* tk, int32_t* tv, int64_t* const* tx, uint64_t* h_out,
                          int32_t* row_out, int64_t* boff, void* ws, size_t ws_bytes,
                          hipStream_t st, int who) {
  if (bits < 0 || bits > 12) return fail("radix_partition_carry: bits must be in [0, 12]");
  if (n < 0 || n >= (1LL << 31)) return fail("radix_partition_carry: n must be below 2^31");
  if (nv < 1 || nv > 2) return fail("radix_partition_carry: one or two carried columns");
  const size_t need = bits <= kOnePassMax ? carve_rp(nullptr, n, bits).bytes
                                          : carve_lv(nullptr, n, bits).bytes;
  if (ws_bytes < need) return fail("radix_partition_carry: workspace too small");
  Carry first{}, second{};
  for (int c = 0; c < nv; ++c) {
    first.src[c] = src[c];
    first.type[c] = type[c];
    // one pass: straight to the output; two: through the first pass's
    // int64 copies
    first.dst[c] = bits <= kOnePassMax ? dst[c] : tx[c];
    second.src[c] = tx ? tx[c] : nullptr;
    second.type[c] = PF_COL_I64;
    second.dst[c] = dst[c];
  }
  if (nv == 1)
    return carry_run<1>(keys, n, bits, first, second, tk, tv, h_out, row_out, boff, ws, st, who);
  return carry_run<2>(keys, n, bits, first, second, tk, tv, h_out, row_out, boff, ws, st, who);
}

// Host-side use by the other tiers of the library.
int radix_partition_segments(const int32_t* keys, const RpSeg* d_segs, int nseg,
                             int64_t nblocks, int64_t hist_entries, int32_t* hist, int64_t* offs,
                             void* scan_ws, int32_t* kout, int32_t* vout, hipStream_t st,
                             const char* hist_name, const char* scatter_name) {
  if (nseg <= 0 || nblocks <= 0) return 0;
  // 1024 x 16 tiles with digits of up to 9 bits: 0.17 ms less than 1024 x 8
  // tiles on the batch's 8-bit segments, and one launch for 8- and 9-bit
  // segments costs no more than one each
  static_assert(kSegChunkRows % (1024 * 16) == 0, "a segment chunk is whole scatter tiles");
  {
    KTimer _kt(hist_name, st);
    rp_hist_seg<<<(unsigned)nblocks, RpShape<kSegMaxBits>::NT, 0, st>>>(keys, d_segs, nseg, hist);
  }
  if (launch_check("segmented partition hist")) return 1;
  if (scan_i32_to_i64_excl(hist, offs, hist_entries, nullptr, scan_ws, st)) return 1;
  {
    KTimer _kt(scatter_name, st);
    rp_scatter8_seg<1024, 16, 1 << kSegMaxBits><<<(unsigned)nblocks, 1024, 0, st>>>(
        keys, d_segs, nseg, offs, kout, vout);
  }
  return launch_check("segmented partition scatter");
}

size_t radix_partition_ws(int64_t n, int bits) {
  const size_t one = carve_rp(nullptr, n, bits).bytes;
  if (bits <= kOnePassMax) return one;
  const size_t two = level_ws(n, bits);
  return one > two ? one : two;
}

size_t radix_level_ws(int64_t n, int bits) { return carve_lv(nullptr, n, bits).bytes; }

int radix_partition_level(const int64_t* keys, int64_t n, int bits, uint64_t* tk, int32_t* tv,
                          uint64_t* h_out, int32_t* row_out, int64_t* boff, void* ws,
                          size_t ws_bytes, hipStream_t st, int who) {
  if (bits <= kOnePassMax || bits > 12) return fail("radix_partition_level: 9 to 12 bits");
  if (n < 0 || n >= (1LL << 31)) return fail("radix_partition_level: n must be below 2^31");
  LvWs w = carve_lv(ws, n, bits);
  if (ws_bytes < w.bytes) return fail("radix_partition_level: workspace too small");
  return level_run<int64_t>(keys, nullptr, n, bits, 0, RP_HASH_OUT, tk, tv, h_out, row_out, boff,
                            w, st, who, -1);
}

int radix_partition(const void* keys, int key_bytes, const int32_t* vals, int64_t n, int bits,
                    int shift, int mode, void* keys_out, int32_t* vals_out,
                    int64_t* bucket_offsets, void* ws, size_t ws_bytes, hipStream_t st,
                    int who, int shape) {
  if (key_bytes != 8 && key_bytes != 4) return fail("radix_partition: keys are int64 or int32");
  if (bits < 0 || bits > 12) return fail("radix_partition: bits must be in [0, 12]");
  // row ids are int32: the iota ids of a call without vals, and the
  // partitioned ids of every caller, stay below 2^31
  if (n < 0 || n >= (1LL << 31))
    return fail("radix_partition: n must be below 2^31 (row ids are int32)");
  if (mode < RP_HASH_KEY || mode > RP_RAW) return fail("radix_partition: unknown mode");
  const int kbits = mode == RP_RAW ? key_bytes * 8 : 64;
  if (shift < 0 || shift + bits > kbits) return fail("radix_partition: shift + bits too large");
  if (mode == RP_HASH_OUT && key_bytes != 8 && keys_out == keys)
    return fail("radix_partition: a hashed output cannot overwrite int32 keys");
  if (bits > kOnePassMax && shape != kShapeOnePass) {
    if (ws_bytes < level_ws(n, bits)) return fail("radix_partition: workspace too small");
    const size_t lb = align256(carve_lv(nullptr, n, bits).bytes);
    char* base = (char*)ws;
    void* tk = base + lb;
    int32_t* tv = (int32_t*)(base + lb + align256((size_t)(n > 0 ? n : 1) * 8));
    LvWs w = carve_lv(ws, n, bits);
    if (key_bytes == 8)
      return level_run<int64_t>((const int64_t*)keys, vals, n, bits, shift, mode, tk, tv,
                                keys_out, vals_out, bucket_offsets, w, st, who, shape);
    return level_run<int32_t>((const int32_t*)keys, vals, n, bits, shift, mode, tk, tv,
                              keys_out, vals_out, bucket_offsets, w, st, who, shape);
  }
  RpWs r = carve_rp(ws, n, bits);
  if (ws_bytes < r.bytes) return fail("radix_partition: workspace too small");
  if (key_bytes == 8)
    return rp_dispatch<int64_t>((const int64_t*)keys, vals, n, bits, shift, mode, keys_out,
                                vals_out, bucket_offsets, r, st, who, shape);
  return rp_dispatch<int32_t>((const int32_t*)keys, vals, n, bits, shift, mode, keys_out,
                              vals_out, bucket_offsets, r, st, who, shape);
}

}  // namespace pf

extern "C" {

size_t pf_radix_partition_workspace_bytes(int64_t n, int32_t bits) {
  if (bits < 0) bits = 0;
  if (bits > 12) bits = 12;
  return pf::radix_partition_ws(n, bits);
}

int pf_radix_partition_shape(const void* keys, int32_t key_bytes, const int32_t* vals,
                             int64_t n, int32_t bits, int32_t shift, int32_t mode,
                             void* keys_out, int32_t* vals_out, int64_t* bucket_offsets,
                             void* workspace, size_t workspace_bytes, int32_t shape,
                             void* stream) {
  if (pf::radix_partition(keys, key_bytes, vals, n, bits, shift, mode, keys_out, vals_out,
                          bucket_offsets, workspace, workspace_bytes, (hipStream_t)stream, 0,
                          shape))
    return 1;
  return pf::ok();
}

int pf_radix_partition(const void* keys, int32_t key_bytes, const int32_t* vals, int64_t n,
                       int32_t bits, int32_t shift, int32_t mode, void* keys_out,
                       int32_t* vals_out, int64_t* bucket_offsets, void* workspace,
                       size_t workspace_bytes, void* stream) {
  return pf_radix_partition_shape(keys, key_bytes, vals, n, bits, shift, mode, keys_out,
                                  vals_out, bucket_offsets, workspace, workspace_bytes, -1,
                                  stream);
}

}  // extern "C"
