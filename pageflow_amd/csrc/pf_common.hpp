// Shared device and host pieces of the PageFlow join kernels for gfx950
// (CDNA4, wave64).  Every kernel in this library includes this header.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

namespace pf {

constexpr int kWave = 64;
constexpr int64_t kEmptyKey = INT64_MIN;  // empty slot of an int64-key table
constexpr int64_t kNull = -64;            // the engine's null sentinel (constants.null)

// ------------------------------------------------------------ hashing --

// Murmur3's 64-bit finalizer (fmix64).  A key's radix bucket is the top
// log2(B) bits of mix64(key); table slots come from its low 32 bits, which
// are independent of the bucket bits.
__device__ __forceinline__ uint64_t mix64(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdULL;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ULL;
  k ^= k >> 33;
  return k;
}

// Murmur3's 32-bit finalizer (fmix32), a bijection on 32-bit values: the
// batch tier's table slots of int32 keys (a third of mix64's instructions on
// 64-bit values).
__device__ __forceinline__ uint32_t fmix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}

// The hash of a 2-key tuple: the first key's mix64 combined with the
// second's (boost-style combine), finalised once more.
__device__ __forceinline__ uint64_t tuple_hash(int64_t k0, int64_t k1, bool two) {
  uint64_t h = mix64((uint64_t)k0);
  if (two) {
    const uint64_t h1 = mix64((uint64_t)k1);
    h ^= h1 + 0x9E3779B97F4A7C15ULL + (h << 6) + (h >> 2);
    h = mix64(h);
  }
  return h;
}

// First slot of a hash in a table of `cap` slots (cap < 2^32): the low 32
// bits of the hash scaled onto [0, cap) (multiply-shift, no modulo).
__device__ __forceinline__ uint32_t slot_of(uint64_t h, uint32_t cap) {
  return (uint32_t)(((uint64_t)(uint32_t)h * (uint64_t)cap) >> 32);
}
__device__ __forceinline__ uint32_t next_slot(uint32_t s, uint32_t cap) {
  return s + 1 == cap ? 0u : s + 1;
}

// ------------------------------------------------------------- wave64 --

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// Lanes below this one that are set in m.
__device__ __forceinline__ int mbcnt(uint64_t m) {
  return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                        __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ uint64_t ballot(bool p) { return (uint64_t)__ballot(p); }

// Orders a wave's LDS writes before its own later LDS reads (and the reads
// before later writes) when lanes exchange data through LDS within the wave.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}


// Inclusive prefix sum across the 64 lanes of a wave.
template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v) {
  const int lane = lane_id();
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const T n = __shfl_up(v, o, kWave);
    if (lane >= o) v += n;
  }
  return v;
}

// The same for 32-bit values with DPP: Hillis-Steele row shifts by 1, 2, 4
// and 8 lanes inside each row of 16 (a lane with no source adds 0), then
// lane 15 of rows 0 and 2 broadcast into rows 1 and 3, and lane 31 into
// rows 2 and 3.  Six VALU adds, no LDS permutes.
__device__ __forceinline__ int wave_incl_scan(int v) {
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);  // row_shr:1
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);  // row_shr:2
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);  // row_shr:4
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);  // row_shr:8
  v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15
  v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// Exclusive prefix sum over the threads of a block of NT threads; *total
// gets the block's sum.  `tmp` holds NT / 64 + 1 values of shared memory.
// Contains __syncthreads (every thread must call it).
template <int NT, typename T>
__device__ __forceinline__ T block_excl_scan(T v, T* tmp, T* total) {
  constexpr int W = NT / kWave;
  const int lane = lane_id(), w = threadIdx.x / kWave;
  const T incl = wave_incl_scan(v);
  if (lane == kWave - 1) tmp[w] = incl;
  __syncthreads();
  if (w == 0) {
    T x = lane < W ? tmp[lane] : T(0);
    T xs = wave_incl_scan(x);
    if (lane < W) tmp[lane] = xs - x;
    if (lane == W - 1) tmp[W] = xs;
  }
  __syncthreads();
  const T r = tmp[w] + incl - v;
  *total = tmp[W];
  __syncthreads();
  return r;
}

// ------------------------------------------------------- device-scope --

// Plain stores into pinned host memory for the few values a call reads back.
__device__ __forceinline__ void store_host(int64_t* p, int64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// ---------------------------------------------------------------- host --

// The calling thread's last error message (pf_last_error).
void set_error(const std::string& msg);
// Clears the calling thread's error; returns 0.
int ok();
// Records msg and returns a nonzero status.
int fail(const std::string& msg);
// Checks a HIP call; on error records "<what>: <hip error>" and returns
// nonzero.
int hip_check(hipError_t e, const char* what);
// Checks the last launch.
int launch_check(const char* what);
// One reusable event per (device, stream) for the calling thread.
hipEvent_t stream_event(hipStream_t st);
// Records the stream's event and waits for it on the host.
int wait_stream(hipStream_t st, const char* what);
// The device alias of a pinned host pointer (kernels store read-backs into
// it).
int host_alias(void* host, void** dev);

// Per-kernel timing (pf_profile): while enabled, a KTimer scope records a
// HIP event pair around the launches it encloses on `st`, accumulated by
// kernel name when read back.  Disabled, it costs one flag test.
struct KTimer {
  KTimer(const char* name, hipStream_t st);
  ~KTimer();
  const char* name;
  hipStream_t st;
  hipEvent_t a = nullptr, b = nullptr;
};

// Bump allocator over a caller-provided workspace (every piece 256-B
// aligned).  Passing base == nullptr only measures.
struct Carve {
  char* base;
  size_t off = 0;
  explicit Carve(void* b) : base((char*)b) {}
  template <typename T>
  T* take(size_t n) {
    off = (off + 255) & ~(size_t)255;
    T* p = base ? (T*)(base + off) : nullptr;
    off += n * sizeof(T);
    return p;
  }
  size_t bytes() const { return (off + 255) & ~(size_t)255; }
};

inline unsigned grid_for(int64_t n, int per_block) {
  int64_t g = (n + per_block - 1) / per_block;
  return (unsigned)(g < 1 ? 1 : g);
}

// Device-wide exclusive prefix sum of n int64 values in place, their total
// into *total (device), through scratch from `carve` (tiles of 8192 values,
// then one block over the tile sums, recursively).  scan_i64_scratch gives
// the scratch a call takes.
size_t scan_i64_scratch(int64_t n);
int scan_i64_excl(int64_t* data, int64_t n, int64_t* total, void* scratch, hipStream_t st);
int scan_i64_excl_from(const int64_t* in, int64_t* out, int64_t n, int64_t* total,
                       void* scratch, hipStream_t st);
// The same over int32 counts, producing int64 offsets (out may not alias in).
size_t scan_i32_scratch(int64_t n);
int scan_i32_to_i64_excl(const int32_t* in, int64_t* out, int64_t n, int64_t* total,
                         void* scratch, hipStream_t st);

// One segment of a segmented radix partition (many arrays partitioned in
// one launch each of hist, scan and scatter): its blocks start at
// blk_start; its keys at in_off of the key buffer (n of them, int32); its
// histogram (bucket-major, nchunks a bucket) at hist_off; its outputs at
// the scanned position + delta; its digit is `bits` bits of fmix32(key),
// held as a 64-bit value, below its top `shift` bits.  The key output is
// fmix32(key) (a bijection of the key); row ids are segment-local.
struct RpSeg {
  int64_t blk_start, in_off, n, nchunks, hist_off, delta;
  int32_t bits, shift;
};
// rows a chunk, every segment: four 16K-row scatter tiles, so a workgroup's
// start (bucket bases, first tile's loads) is spread over 64K rows
constexpr int64_t kSegChunkRows = 65536;
constexpr int kSegMaxBits = 9;            // digit bits of a segment, at most

// Stable segmented partition: one hist, one scan and one scatter launch
// over the nseg segments in d_segs (nblocks chunks in all, hist_entries
// histogram entries); hist/offs/scan_ws sized for hist_entries.  kout gets
// each row's fmix32(key), vout its segment-local row id.
int radix_partition_segments(const int32_t* keys, const RpSeg* d_segs, int nseg,
                             int64_t nblocks, int64_t hist_entries, int32_t* hist, int64_t* offs,
                             void* scan_ws, int32_t* kout, int32_t* vout, hipStream_t st,
                             const char* hist_name, const char* scatter_name);

// Fills n values of a device array (a kernel, never hipMemsetAsync, so that
// graph capture records no memset node).
int fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t st);
int fill_i64(int64_t* p, int64_t n, int64_t v, hipStream_t st);

}  // namespace pf
