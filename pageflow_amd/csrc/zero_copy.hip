// Zero-copy reads of a column held in pinned host memory: the out-of-core
// payload gather of a join's late materialisation, and the masked read of a
// scan (the reference's IO/zero_copy_reader.cu:54-120, driven by
// IO/pipeline/rearrange_pipeline.py:78-93).  The GPU reads the host rows
// itself over its host link; nothing is staged through a copy engine.
//
// A column is a list of segments: runs of whole rows back to back in pinned
// host memory.  seg_ptrs[s] is the device address of segment s's first row,
// seg_rows[s] its first row in the column's row space (ascending), and
// seg_rows[num_segs] the end of the last.  Row i of a read is row
// row_start + i of the row space.
//
//   gather   out[r] = row row_start + idx[r], or -64 in every element of
//            the row when idx[r] is -64 (the null row of an outer join).
//            The indices are first radix-sorted on the GPU (stable LSD, 8
//            bits a pass over the bits their range spans), so the host rows
//            are read in ascending order: lanes of a wave read neighbouring
//            rows and their reads merge into whole host-link requests.  The
//            rows are then written to their output positions.
//   mask     out[k] = row row_start + i for the k-th i with mask[i] != 0.
//
// A row outside every segment is not read: its output is -64 and it is
// counted into a device counter the caller checks (the reference traps).
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

constexpr int64_t kNullRow = -64;
constexpr int ZC_NT = 256;
constexpr int ZC_SORT_MIN = 1 << 16;  // fewer indices are read in output order
constexpr int kRaw = 2;               // radix_partition's raw-digit mode
constexpr int kWhoSort = 3;           // its profiler names for this caller
constexpr int MT = 4096;              // mask rows a tile

// The segment holding `row`: the last s with seg_rows[s] <= row, or
// num_segs when the row is past the end.
__device__ __forceinline__ int zc_segment(const int64_t* __restrict__ seg_rows, int num_segs,
                                          int64_t row) {
  int lo = 0, hi = num_segs;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (seg_rows[mid + 1] <= row)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// Element `off` of source row `src` (row-space row row_start + src), or -64.
template <typename T>
__device__ __forceinline__ T zc_read(const int64_t* __restrict__ seg_ptrs,
                                     const int64_t* __restrict__ seg_rows, int num_segs,
                                     int64_t row_start, int64_t src, int64_t width, int64_t off,
                                     unsigned long long* bad) {
  if (src == kNullRow) return (T)kNullRow;
  const int64_t row = row_start + src;
  const int s = zc_segment(seg_rows, num_segs, row);
  if (s >= num_segs || row < seg_rows[0]) {
    if (off == 0) atomicAdd(bad, 1ULL);
    return (T)kNullRow;
  }
  const T* seg = (const T*)seg_ptrs[s];
  return seg[(row - seg_rows[s]) * width + off];
}

// out[pos[r]] = source row idx[r] (pos null: r), one element a lane.
template <typename T>
__global__ __launch_bounds__(ZC_NT) void zc_gather(const int64_t* __restrict__ seg_ptrs,
                                                   const int64_t* __restrict__ seg_rows,
                                                   int num_segs, int64_t row_start,
                                                   const int64_t* __restrict__ idx,
                                                   const int32_t* __restrict__ pos, int64_t n,
                                                   int64_t width, T* __restrict__ out,
                                                   unsigned long long* bad) {
  const int64_t total = n * width;
  const int64_t stride = (int64_t)gridDim.x * ZC_NT;
  for (int64_t e = blockIdx.x * (int64_t)ZC_NT + threadIdx.x; e < total; e += stride) {
    const int64_t r = width == 1 ? e : e / width;
    const int64_t off = e - r * width;
    const int64_t p = pos ? (int64_t)pos[r] : r;
    out[p * width + off] =
        zc_read<T>(seg_ptrs, seg_rows, num_segs, row_start, idx[r], width, off, bad);
  }
}

// Mask tiles: the kept rows of each tile of MT mask bytes.
__global__ __launch_bounds__(ZC_NT) void zc_mask_count(const uint8_t* __restrict__ mask,
                                                       int64_t n, int32_t* __restrict__ cnt) {
  __shared__ int32_t red[ZC_NT / kWave];
  const int64_t t0 = blockIdx.x * (int64_t)MT;
  int c = 0;
  for (int i = threadIdx.x; i < MT && t0 + i < n; i += ZC_NT) c += mask[t0 + i] != 0;
  c = wave_sum(c);
  if (lane_id() == 0) red[threadIdx.x / kWave] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0;
    for (int w = 0; w < ZC_NT / kWave; ++w) s += red[w];
    cnt[blockIdx.x] = s;
  }
}

// Block b: tile b's kept rows, in order, from its first output offs[b];
// outputs past n_out are not written (the caller's count disagrees).
template <typename T>
__global__ __launch_bounds__(ZC_NT) void zc_mask_gather(
    const int64_t* __restrict__ seg_ptrs, const int64_t* __restrict__ seg_rows, int num_segs,
    int64_t row_start, const uint8_t* __restrict__ mask, int64_t n,
    const int64_t* __restrict__ offs, int64_t width, T* __restrict__ out, int64_t n_out,
    unsigned long long* bad) {
  __shared__ int32_t wcnt[ZC_NT / kWave];
  __shared__ int64_t s_base;
  const int64_t t0 = blockIdx.x * (int64_t)MT;
  const int w = threadIdx.x / kWave;
  if (threadIdx.x == 0) s_base = offs[blockIdx.x];
  for (int i0 = 0; i0 < MT && t0 + i0 < n; i0 += ZC_NT) {
    const int64_t i = t0 + i0 + threadIdx.x;
    const bool keep = i < n && i < t0 + MT && mask[i] != 0;
    const uint64_t bal = ballot(keep);
    if (lane_id() == 0) wcnt[w] = (int)__popcll(bal);
    __syncthreads();
    int before = 0, tot = 0;
#pragma unroll
    for (int q = 0; q < ZC_NT / kWave; ++q) {
      before += q < w ? wcnt[q] : 0;
      tot += wcnt[q];
    }
    const int64_t k = s_base + before + mbcnt(bal);
    if (keep && k < n_out)
      for (int64_t off = 0; off < width; ++off)
        out[k * width + off] =
            zc_read<T>(seg_ptrs, seg_rows, num_segs, row_start, i, width, off, bad);
    __syncthreads();
    if (threadIdx.x == 0) s_base += tot;
  }
}

// the element type of an elem_bytes-wide column (floats move as their bits,
// except the null: -64 is written as a value of the column's type)
template <template <typename> class F, typename... A>
int by_type(int elem_bytes, int is_float, A... a) {
  if (is_float) {
    if (elem_bytes == 4) return F<float>::run(a...);
    if (elem_bytes == 8) return F<double>::run(a...);
    return fail("zero_copy: float columns are 4 or 8 bytes");
  }
  switch (elem_bytes) {
    case 1: return F<int8_t>::run(a...);
    case 2: return F<int16_t>::run(a...);
    case 4: return F<int32_t>::run(a...);
    case 8: return F<int64_t>::run(a...);
    default: return fail("zero_copy: elements are 1, 2, 4 or 8 bytes");
  }
}

template <typename T>
struct GatherLaunch {
  static int run(const int64_t* seg_ptrs, const int64_t* seg_rows, int num_segs,
                 int64_t row_start, const int64_t* idx, const int32_t* pos, int64_t n,
                 int64_t width, void* out, unsigned long long* bad, hipStream_t st) {
    if (n > 0) {
      KTimer _kt("zc_gather", st);
      const unsigned grid = (unsigned)std::min<int64_t>(grid_for(n * width, ZC_NT), 8192);
      zc_gather<T><<<grid, ZC_NT, 0, st>>>(seg_ptrs, seg_rows, num_segs, row_start, idx, pos,
                                            n, width, (T*)out, bad);
    }
    return launch_check("zero_copy gather");
  }
};

template <typename T>
struct MaskLaunch {
  static int run(const int64_t* seg_ptrs, const int64_t* seg_rows, int num_segs,
                 int64_t row_start, const uint8_t* mask, int64_t n, const int64_t* offs,
                 int64_t width, void* out, int64_t n_out, unsigned long long* bad,
                 hipStream_t st) {
    const int64_t tiles = (n + MT - 1) / MT;
    if (tiles > 0) {
      KTimer _kt("zc_mask_gather", st);
      zc_mask_gather<T><<<(unsigned)tiles, ZC_NT, 0, st>>>(
          seg_ptrs, seg_rows, num_segs, row_start, mask, n, offs, width, (T*)out, n_out, bad);
    }
    return launch_check("zero_copy mask");
  }
};

int check_common(const int64_t* seg_ptrs, const int64_t* seg_rows, int32_t num_segs,
                 int64_t row_start, int64_t row_stop, int64_t width, int64_t* bad) {
  if (!seg_ptrs || !seg_rows || num_segs < 1) return fail("zero_copy: no segment to read");
  if (row_start < 0 || row_start > row_stop)
    return fail("zero_copy: rows [" + std::to_string(row_start) + ", " +
                std::to_string(row_stop) + ") are not a range");
  if (width < 1) return fail("zero_copy: width must be positive");
  if (!bad) return fail("zero_copy: no bad-row counter");
  return 0;
}

// bits an index of a read of `span` rows needs
int span_bits(int64_t span) {
  int b = 0;
  while (b < 62 && (int64_t{1} << b) < span) ++b;
  return b;
}

struct ZcWs {
  int64_t *k0, *k1, *offs;
  int32_t *v0, *v1;
  void* rp;
  size_t rp_bytes;
  size_t bytes;
};

ZcWs carve_zc(void* ws, int64_t n) {
  Carve cv(ws);
  ZcWs z;
  const int64_t m = n > 0 ? n : 1;
  z.k0 = cv.take<int64_t>(m);
  z.k1 = cv.take<int64_t>(m);
  z.v0 = cv.take<int32_t>(m);
  z.v1 = cv.take<int32_t>(m);
  z.offs = cv.take<int64_t>(257);
  z.rp_bytes = radix_partition_ws(m, 8);
  z.rp = cv.take<char>(z.rp_bytes);
  z.bytes = cv.bytes();
  return z;
}

}  // namespace
}  // namespace pf

extern "C" {

int pf_device_alias(const void* host, void** dev) {
  using namespace pf;
  if (!host || !dev) return fail("device_alias: null pointer");
  (void)hipGetLastError();
  void* alias = nullptr;
  const hipError_t e = hipHostGetDevicePointer(&alias, const_cast<void*>(host), 0);
  if (e != hipSuccess)
    return fail(std::string("device_alias: not pinned host memory (") + hipGetErrorString(e) +
                ")");
  *dev = alias;
  return ok();
}

size_t pf_zero_copy_rearrange_workspace_bytes(int64_t n) {
  return pf::carve_zc(nullptr, n).bytes;
}

int pf_zero_copy_rearrange(const int64_t* seg_ptrs, const int64_t* seg_rows, int32_t num_segs,
                           int64_t row_start, int64_t row_stop, const int64_t* rearrange,
                           int64_t n, int32_t elem_bytes, int32_t is_float, int64_t width,
                           void* out, int64_t* bad, void* workspace, size_t workspace_bytes,
                           void* stream) {
  using namespace pf;
  hipStream_t st = (hipStream_t)stream;
  if (check_common(seg_ptrs, seg_rows, num_segs, row_start, row_stop, width, bad)) return 1;
  if (n < 0 || n >= (int64_t{1} << 31))
    return fail("zero_copy_rearrange: fewer than 2^31 output rows a call");
  if (n > 0 && !rearrange) return fail("zero_copy_rearrange: no indices");
  if (workspace_bytes < pf_zero_copy_rearrange_workspace_bytes(n))
    return fail("zero_copy_rearrange: workspace too small");
  ZcWs z = carve_zc(workspace, n);
  const int64_t* idx = rearrange;
  const int32_t* pos = nullptr;
  const int bits = span_bits(row_stop - row_start);
  // rows sharing a 128-B host line need no order among themselves: the sort
  // skips the index bits below the line (4 for 8-B rows, one pass of four
  // for a 2^28-row span)
  const int64_t row_bytes = (int64_t)elem_bytes * width;
  int low = 0;
  while (low < bits && (row_bytes << (low + 1)) <= 128) ++low;
  if (n >= ZC_SORT_MIN && bits > low) {
    // stable LSD passes of 8 bits: the indices ascending by line, each with
    // its output position
    const int64_t* kin = rearrange;
    const int32_t* vin = nullptr;
    int64_t* kout = z.k0;
    int32_t* vout = z.v0;
    for (int shift = low; shift < bits; shift += 8) {
      const int b = bits - shift < 8 ? bits - shift : 8;
      if (radix_partition(kin, 8, vin, n, b, shift, kRaw, kout, vout, z.offs, z.rp, z.rp_bytes,
                          st, kWhoSort, -1))
        return 1;
      kin = kout;
      vin = vout;
      kout = kout == z.k0 ? z.k1 : z.k0;
      vout = vout == z.v0 ? z.v1 : z.v0;
    }
    idx = kin;
    pos = vin;
  }
  return by_type<GatherLaunch>(elem_bytes, is_float, seg_ptrs, seg_rows, (int)num_segs,
                               row_start, idx, pos, n, width, out,
                               (unsigned long long*)bad, st) ||
         ok();
}

size_t pf_zero_copy_mask_workspace_bytes(int64_t n) {
  pf::Carve cv(nullptr);
  const int64_t tiles = (n + pf::MT - 1) / pf::MT;
  cv.take<int32_t>(tiles > 0 ? tiles : 1);
  cv.take<int64_t>(tiles > 0 ? tiles : 1);
  cv.take<int64_t>(2);
  cv.take<char>(pf::scan_i32_scratch(tiles));
  return cv.bytes();
}

int pf_zero_copy_mask(const int64_t* seg_ptrs, const int64_t* seg_rows, int32_t num_segs,
                      int64_t row_start, int64_t row_stop, const uint8_t* mask, int32_t elem_bytes,
                      int32_t is_float, int64_t width, void* out, int64_t n_out,
                      int64_t* kept, int64_t* bad, void* workspace, size_t workspace_bytes,
                      void* stream) {
  using namespace pf;
  hipStream_t st = (hipStream_t)stream;
  if (check_common(seg_ptrs, seg_rows, num_segs, row_start, row_stop, width, bad)) return 1;
  const int64_t n = row_stop - row_start;
  if (n > 0 && !mask) return fail("zero_copy_mask: no mask");
  if (n_out < 0 || !kept) return fail("zero_copy_mask: bad output");
  if (workspace_bytes < pf_zero_copy_mask_workspace_bytes(n))
    return fail("zero_copy_mask: workspace too small");
  Carve cv(workspace);
  const int64_t tiles = (n + MT - 1) / MT;
  int32_t* cnt = cv.take<int32_t>(tiles > 0 ? tiles : 1);
  int64_t* offs = cv.take<int64_t>(tiles > 0 ? tiles : 1);
  int64_t* total = cv.take<int64_t>(2);
  void* scan = cv.take<char>(scan_i32_scratch(tiles));
  if (tiles > 0) {
    zc_mask_count<<<(unsigned)tiles, ZC_NT, 0, st>>>(mask, n, cnt);
    if (launch_check("zero_copy mask count")) return 1;
  }
  if (scan_i32_to_i64_excl(cnt, offs, tiles, kept, scan, st)) return 1;
  (void)total;
  return by_type<MaskLaunch>(elem_bytes, is_float, seg_ptrs, seg_rows, (int)num_segs,
                             row_start, mask, n, (const int64_t*)offs, width, out, n_out,
                             (unsigned long long*)bad, st) ||
         ok();
}

}  // extern "C"
