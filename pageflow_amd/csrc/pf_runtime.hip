// Host runtime of the library (error state, events, pinned read-back
// aliases), the device-wide scans and fills every entry point shares, and the
// hashing entry points (pf_mix64, pf_bucket_ids).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "pf_common.hpp"
#include "../../include/pageflow_join.h"

namespace pf {

namespace {
thread_local std::string g_error;
}

void set_error(const std::string& msg) { g_error = msg; }
int ok() {
  g_error.clear();
  return 0;
}
int fail(const std::string& msg) {
  g_error = msg;
  return 1;
}

int hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return 0;
  return fail(std::string(what) + ": " + hipGetErrorString(e));
}

int launch_check(const char* what) { return hip_check(hipGetLastError(), what); }

hipEvent_t stream_event(hipStream_t st) {
  thread_local std::map<std::pair<int, hipStream_t>, hipEvent_t> events;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  auto key = std::make_pair(dev, st);
  auto it = events.find(key);
  if (it != events.end()) return it->second;
  hipEvent_t ev = nullptr;
  if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) return nullptr;
  events.emplace(key, ev);
  return ev;
}

int wait_stream(hipStream_t st, const char* what) {
  hipEvent_t ev = stream_event(st);
  if (!ev) return fail(std::string(what) + ": cannot create an event");
  if (hip_check(hipEventRecord(ev, st), what)) return 1;
  return hip_check(hipEventSynchronize(ev), what);
}

int host_alias(void* host, void** dev) {
  *dev = nullptr;
  hipError_t e = hipHostGetDevicePointer(dev, host, 0);
  if (e != hipSuccess || *dev == nullptr) {
    (void)hipGetLastError();
    return fail("host_out is not pinned, device-mapped host memory");
  }
  return 0;
}

// ----------------------------------------------------------- profile --

namespace {
struct Pending {
  std::string name;
  hipEvent_t a, b;
};
std::mutex g_prof_mu;
bool g_prof_on = false;
std::vector<Pending> g_pending;
std::map<std::string, std::pair<double, int64_t>> g_prof;  // ms, launches

void drain_locked() {
  for (auto& p : g_pending) {
    float ms = 0.f;
    if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      auto& e = g_prof[p.name];
      e.first += ms;
      e.second += 1;
    }
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  g_pending.clear();
}
}  // namespace

KTimer::KTimer(const char* n, hipStream_t s) : name(n), st(s) {
  std::lock_guard<std::mutex> lock(g_prof_mu);
  if (!g_prof_on) return;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
    a = b = nullptr;
    return;
  }
  (void)hipEventRecord(a, st);
}

KTimer::~KTimer() {
  if (!a) return;
  (void)hipEventRecord(b, st);
  std::lock_guard<std::mutex> lock(g_prof_mu);
  g_pending.push_back({name, a, b});
}

// ------------------------------------------------------------- fills --

namespace {

template <typename T>
__global__ void k_fill(T* __restrict__ p, int64_t n, T v) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

template <typename T>
int fill_t(T* p, int64_t n, T v, hipStream_t st) {
  if (n <= 0) return 0;
  int64_t g = (n + 255) / 256;
  if (g > 8192) g = 8192;
  k_fill<T><<<(unsigned)g, 256, 0, st>>>(p, n, v);
  return launch_check("fill");
}

// ------------------------------------------------------------- scans --

constexpr int SCAN_NT = 256;
constexpr int SCAN_IT = 16;
constexpr int SCAN_TILE = SCAN_NT * SCAN_IT;  // 4096 values a tile

// Sum of each tile of `in`.
template <typename TI>
__global__ __launch_bounds__(SCAN_NT) void k_tile_sums(const TI* __restrict__ in, int64_t n,
                                                       int64_t* __restrict__ sums) {
  __shared__ int64_t red[SCAN_NT / kWave];
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  int64_t s = 0;
#pragma unroll
  for (int j = 0; j < SCAN_IT; ++j) {
    const int64_t i = base + j * SCAN_NT + threadIdx.x;
    if (i < n) s += (int64_t)in[i];
  }
  s = wave_sum(s);
  if (lane_id() == 0) red[threadIdx.x / kWave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t t = 0;
#pragma unroll
    for (int w = 0; w < SCAN_NT / kWave; ++w) t += red[w];
    sums[blockIdx.x] = t;
  }
}

// Exclusive scan of each tile of `in` into `out`, plus the tile's offset
// (tile_off, or 0 when null); the last tile's thread stores the grand total.
template <typename TI>
__global__ __launch_bounds__(SCAN_NT) void k_tile_scan(const TI* in, int64_t* out, int64_t n,
                                                       const int64_t* __restrict__ tile_off,
                                                       int64_t* total) {
  // one pad value every SCAN_IT values: a thread's run spans distinct banks
  __shared__ int64_t s[SCAN_TILE + SCAN_NT];
  __shared__ int64_t tmp[SCAN_NT / kWave + 1];
  const int tid = threadIdx.x;
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
#pragma unroll
  for (int j = 0; j < SCAN_IT; ++j) {
    const int idx = j * SCAN_NT + tid;
    const int64_t i = base + idx;
    s[idx + idx / SCAN_IT] = i < n ? (int64_t)in[i] : 0;
  }
  __syncthreads();
  int64_t local = 0;
#pragma unroll
  for (int j = 0; j < SCAN_IT; ++j) local += s[tid * (SCAN_IT + 1) + j];
  int64_t tot;
  int64_t run = block_excl_scan<SCAN_NT>(local, tmp, &tot);
  run += tile_off ? tile_off[blockIdx.x] : 0;
#pragma unroll
  for (int j = 0; j < SCAN_IT; ++j) {
    const int64_t v = s[tid * (SCAN_IT + 1) + j];
    s[tid * (SCAN_IT + 1) + j] = run;
    run += v;
  }
  if (total && blockIdx.x == gridDim.x - 1 && tid == SCAN_NT - 1) *total = run;
  __syncthreads();
#pragma unroll
  for (int j = 0; j < SCAN_IT; ++j) {
    const int idx = j * SCAN_NT + tid;
    const int64_t i = base + idx;
    if (i < n) out[i] = s[idx + idx / SCAN_IT];
  }
}

__global__ void k_store_zero(int64_t* p) { *p = 0; }

size_t scan_scratch_values(int64_t n) {
  // tile sums at every level (each level's sums are scanned in place)
  size_t v = 0;
  int64_t m = n;
  while (m > SCAN_TILE) {
    m = (m + SCAN_TILE - 1) / SCAN_TILE;
    v += (size_t)m + 64;
  }
  return v;
}

template <typename TI>
int scan_impl(const TI* in, int64_t* out, int64_t n, int64_t* total, int64_t* scratch,
              hipStream_t st) {
  if (n <= 0) {
    if (total) {
      k_store_zero<<<1, 1, 0, st>>>(total);
      return launch_check("scan");
    }
    return 0;
  }
  const int64_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (tiles == 1) {
    k_tile_scan<TI><<<1, SCAN_NT, 0, st>>>(in, out, n, nullptr, total);
    return launch_check("scan");
  }
  int64_t* sums = scratch;
  int64_t* rest = scratch + tiles + 64;
  k_tile_sums<TI><<<(unsigned)tiles, SCAN_NT, 0, st>>>(in, n, sums);
  if (launch_check("scan")) return 1;
  if (scan_impl<int64_t>(sums, sums, tiles, nullptr, rest, st)) return 1;
  k_tile_scan<TI><<<(unsigned)tiles, SCAN_NT, 0, st>>>(in, out, n, sums, total);
  return launch_check("scan");
}

}  // namespace

int fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t st) { return fill_t(p, n, v, st); }
int fill_i64(int64_t* p, int64_t n, int64_t v, hipStream_t st) { return fill_t(p, n, v, st); }

size_t scan_i64_scratch(int64_t n) { return scan_scratch_values(n) * sizeof(int64_t); }
size_t scan_i32_scratch(int64_t n) { return scan_scratch_values(n) * sizeof(int64_t); }

int scan_i64_excl(int64_t* data, int64_t n, int64_t* total, void* scratch, hipStream_t st) {
  return scan_impl<int64_t>(data, data, n, total, (int64_t*)scratch, st);
}

int scan_i64_excl_from(const int64_t* in, int64_t* out, int64_t n, int64_t* total,
                       void* scratch, hipStream_t st) {
  return scan_impl<int64_t>(in, out, n, total, (int64_t*)scratch, st);
}

int scan_i32_to_i64_excl(const int32_t* in, int64_t* out, int64_t n, int64_t* total,
                         void* scratch, hipStream_t st) {
  return scan_impl<int32_t>(in, out, n, total, (int64_t*)scratch, st);
}

// ----------------------------------------------------------- hashing --

namespace {

__global__ void k_mix64(const int64_t* __restrict__ in, int64_t n, int64_t* __restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride)
    out[i] = (int64_t)mix64((uint64_t)in[i]);
}

__global__ void k_bucket_ids(const int64_t* __restrict__ in, int64_t n, int logB,
                             int32_t* __restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride)
    out[i] = logB ? (int32_t)(mix64((uint64_t)in[i]) >> (64 - logB)) : 0;
}

// Source types of pf_join_key_bits (pageflow_join.h PF_KEY_*).
enum KeyType : int { KT_I64 = 0, KT_I32 = 1, KT_I16 = 2, KT_I8 = 3, KT_U8 = 4, KT_F64 = 5,
                     KT_F32 = 6, KT_BOOL = 7 };

template <typename T>
__device__ __forceinline__ double key_as_f64(T v) { return (double)v; }

// The int64 image of a join key under which two keys are equal exactly when
// the reference's comparison finds them equal.  Integers: the value itself.
// Floats (and integers joined against floats, compared as float64): the bit
// pattern of the value with -0.0 folded into +0.0; a NaN, equal to nothing,
// becomes one NaN pattern per side, so the two sides' NaNs never meet.
template <typename T>
__global__ void k_key_bits(const T* __restrict__ in, int64_t n, bool as_float,
                           int64_t nan_bits, int64_t* __restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) {
    const T v = in[i];
    if (!as_float) {
      out[i] = (int64_t)v;
      continue;
    }
    double d = key_as_f64(v);
    int64_t b;
    if (d != d) {
      b = nan_bits;
    } else {
      if (d == 0.0) d = 0.0;  // -0.0 -> +0.0
      b = __double_as_longlong(d);
    }
    out[i] = b;
  }
}

unsigned stream_grid(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g > 16384) g = 16384;
  return (unsigned)(g < 1 ? 1 : g);
}

}  // namespace

}  // namespace pf

extern "C" {

int pf_version(void) { return PF_ABI_VERSION; }

int pf_profile(int32_t enable) {
  std::lock_guard<std::mutex> lock(pf::g_prof_mu);
  pf::drain_locked();
  pf::g_prof.clear();
  pf::g_prof_on = enable != 0;
  return pf::ok();
}

int32_t pf_profile_read(char* names, int64_t names_bytes, double* total_ms, int64_t* launches,
                        int32_t max_kernels) {
  std::lock_guard<std::mutex> lock(pf::g_prof_mu);
  pf::drain_locked();
  int32_t k = 0;
  int64_t at = 0;
  for (auto& e : pf::g_prof) {
    if (k >= max_kernels) break;
    const int64_t len = (int64_t)e.first.size() + 1;
    if (at + len > names_bytes) break;
    memcpy(names + at, e.first.c_str(), (size_t)len);
    at += len;
    total_ms[k] = e.second.first;
    launches[k] = e.second.second;
    ++k;
  }
  return k;
}

const char* pf_last_error(void) { return pf::g_error.c_str(); }

int pf_mix64(const int64_t* keys, int64_t n, int64_t* out, void* stream) {
  if (n < 0) return pf::fail("pf_mix64: negative row count");
  if (n == 0) return pf::ok();
  pf::k_mix64<<<pf::stream_grid(n), 256, 0, (hipStream_t)stream>>>(keys, n, out);
  if (pf::launch_check("pf_mix64")) return 1;
  return pf::ok();
}

int pf_join_key_bits(const void* keys, int32_t key_type, int64_t n, int32_t as_float,
                     int32_t side, int64_t* out, void* stream) {
  using namespace pf;
  if (n < 0) return fail("join_key_bits: negative row count");
  if (side != 0 && side != 1) return fail("join_key_bits: side is 0 (build) or 1 (probe)");
  const bool fl = as_float != 0 || key_type == KT_F64 || key_type == KT_F32;
  // quiet NaN patterns no canonical value takes: payload 1 (build), 2 (probe)
  const int64_t nan_bits = (int64_t)0x7ff8000000000001ULL + side;
  if (n == 0) return ok();
  hipStream_t st = (hipStream_t)stream;
  const unsigned g = stream_grid(n);
  switch (key_type) {
    case KT_I64: k_key_bits<int64_t><<<g, 256, 0, st>>>((const int64_t*)keys, n, fl, nan_bits, out); break;
    case KT_I32: k_key_bits<int32_t><<<g, 256, 0, st>>>((const int32_t*)keys, n, fl, nan_bits, out); break;
    case KT_I16: k_key_bits<int16_t><<<g, 256, 0, st>>>((const int16_t*)keys, n, fl, nan_bits, out); break;
    case KT_I8: k_key_bits<int8_t><<<g, 256, 0, st>>>((const int8_t*)keys, n, fl, nan_bits, out); break;
    case KT_U8:
    case KT_BOOL: k_key_bits<uint8_t><<<g, 256, 0, st>>>((const uint8_t*)keys, n, fl, nan_bits, out); break;
    case KT_F64: k_key_bits<double><<<g, 256, 0, st>>>((const double*)keys, n, true, nan_bits, out); break;
    case KT_F32: k_key_bits<float><<<g, 256, 0, st>>>((const float*)keys, n, true, nan_bits, out); break;
    default: return fail("join_key_bits: unknown key type " + std::to_string(key_type));
  }
  if (launch_check("pf_join_key_bits")) return 1;
  return ok();
}

int pf_bucket_ids(const int64_t* keys, int64_t n, int32_t B, int32_t* out, void* stream) {
  if (n < 0) return pf::fail("pf_bucket_ids: negative row count");
  if (B < 1 || (B & (B - 1))) return pf::fail("pf_bucket_ids: B must be a power of two");
  int logB = 0;
  while ((1 << logB) < B) ++logB;
  if (logB > 31) return pf::fail("pf_bucket_ids: B too large");
  if (n == 0) return pf::ok();
  pf::k_bucket_ids<<<pf::stream_grid(n), 256, 0, (hipStream_t)stream>>>(keys, n, logB, out);
  if (pf::launch_check("pf_bucket_ids")) return 1;
  return pf::ok();
}

}  // extern "C"
