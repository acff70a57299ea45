// Payload gather with the null row: the join's output columns at the
// matched rows (operators/hashjoin.py:359-377, _gather_join_result's
// index_select) and the right-outer gather that writes a null row where the
// index is the null sentinel (conversion.py:39-50, index_with_null).
//
// Rows of any width: elements of 1, 2, 4 or 8 bytes, row_elems a row.
// Consecutive threads take consecutive output elements, so writes coalesce;
// reads follow the index.
#include <hip/hip_runtime.h>

#include "pf_common.hpp"
#include "../../include/pageflow_join.h"

namespace pf {
namespace {

// T is an integer of the element's width: rows are copied as bits, and a
// null row's elements take `nul` (the -64 of the column's own type).
template <typename T>
__global__ void k_gather(const T* __restrict__ src, int64_t nsrc, int64_t row_elems,
                         const int64_t* __restrict__ idx, int64_t n, T* __restrict__ out,
                         int64_t* bad, T nul) {
  const int64_t total = n * row_elems;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += stride) {
    const int64_t i = e / row_elems, q = e - i * row_elems;
    const int64_t r = idx[i];
    if (r == kNull || nsrc == 0) {
      out[e] = nul;
    } else if (r < 0 || r >= nsrc) {
      out[e] = nul;
      if (bad) *bad = 1;
    } else {
      out[e] = src[r * row_elems + q];
    }
  }
}

// flags[idx[i]] = value for every i (idempotent on repeats).
__global__ void k_mark(const int64_t* __restrict__ idx, int64_t n, uint8_t* __restrict__ flags,
                       int64_t nflags, uint8_t value) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) {
    const int64_t r = idx[i];
    if (r >= 0 && r < nflags) flags[r] = value;
  }
}

unsigned grid_of(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g > 16384) g = 16384;
  return (unsigned)(g < 1 ? 1 : g);
}

}  // namespace
}  // namespace pf

extern "C" {

int pf_gather_rows(const void* src, int64_t nsrc, int32_t elem_bytes, int64_t row_elems,
                   const int64_t* idx, int64_t n, void* out, int64_t* bad, int64_t null_bits,
                   void* stream) {
  using namespace pf;
  hipStream_t st = (hipStream_t)stream;
  if (n < 0 || nsrc < 0 || row_elems < 1) return fail("gather_rows: bad sizes");
  if (n == 0) return ok();
  {
    KTimer _kt("gather_rows", st);
    const unsigned g = grid_of(n * row_elems);
    switch (elem_bytes) {
      case 1: k_gather<int8_t><<<g, 256, 0, st>>>((const int8_t*)src, nsrc, row_elems, idx, n,
                                                  (int8_t*)out, bad, (int8_t)null_bits); break;
      case 2: k_gather<int16_t><<<g, 256, 0, st>>>((const int16_t*)src, nsrc, row_elems, idx, n,
                                                   (int16_t*)out, bad, (int16_t)null_bits); break;
      case 4: k_gather<int32_t><<<g, 256, 0, st>>>((const int32_t*)src, nsrc, row_elems, idx, n,
                                                   (int32_t*)out, bad, (int32_t)null_bits); break;
      case 8: k_gather<int64_t><<<g, 256, 0, st>>>((const int64_t*)src, nsrc, row_elems, idx, n,
                                                   (int64_t*)out, bad, (int64_t)null_bits); break;
      default: return fail("gather_rows: elements of 1, 2, 4 or 8 bytes");
    }
  }
  if (launch_check("gather_rows")) return 1;
  return ok();
}

int pf_mark_rows(const int64_t* idx, int64_t n, uint8_t* flags, int64_t nflags, int32_t value,
                 void* stream) {
  using namespace pf;
  hipStream_t st = (hipStream_t)stream;
  if (n < 0 || nflags < 0) return fail("mark_rows: bad sizes");
  if (n == 0) return ok();
  k_mark<<<grid_of(n), 256, 0, st>>>(idx, n, flags, nflags, (uint8_t)value);
  if (launch_check("mark_rows")) return 1;
  return ok();
}

}  // extern "C"
