/*
 * pageflow_join.h: the C ABI of the MI355X (gfx950) equi-join hot path.
 *
 * Drop-in boundary for the reference's pybind module TQPlib.tqpkernels
 * (operators/kernels/tqpkernels.cu:22-46 of the reference repository) plus
 * the entry points of the batched and serving tiers.  Plain pointers and
 * sizes only:
 *
 *  - device pointers are HIP device memory (torch-ROCm tensors' data_ptr());
 *  - `stream` is a hipStream_t (torch.cuda.current_stream().cuda_stream);
 *  - `host_out` is pinned, device-mapped host memory (torch pin_memory);
 *  - outputs are allocated by the caller at the documented upper bound and
 *    narrowed by the caller to the count the call reads back, as the
 *    reference's Python wrappers narrow theirs;
 *  - scratch comes in as a caller-allocated workspace of the size the
 *    matching *_workspace_bytes() function returns;
 *  - every entry point returns 0 on success, nonzero on error, with the
 *    message in pf_last_error() (per calling thread).
 *
 * A call that "waits" makes exactly one host wait (hipEventSynchronize) on
 * its stream, as the reference's calls do (kernel_utils.cuh:37-40).
 */
#ifndef PAGEFLOW_JOIN_H
#define PAGEFLOW_JOIN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PF_ABI_VERSION 1

/* Library ABI version (PF_ABI_VERSION). */
int pf_version(void);

/* The calling thread's last error message ("" after a successful call). */
const char* pf_last_error(void);

/* ------------------------------------------------------------ hashing --
 * mix64 (kernel_utils.cuh:18-22) of n int64 keys into out (int64 bits).
 * Replaces the torch-op twin operators/aggregate.py:95-104. */
int pf_mix64(const int64_t* keys, int64_t n, int64_t* out, void* stream);

/* The radix bucket of n int64 hashes among B (a power of two): the top
 * log2(B) bits of mix64, 0 when B is 1, as int32.
 * Replaces operators/aggregate.py:107-116 (_bucket_ids). */
int pf_bucket_ids(const int64_t* keys, int64_t n, int32_t B, int32_t* out, void* stream);

/* ---------------------------------------------------- partition_sort --
 * Replaces tqpkernels.partition_sort (operators/kernels/partition_sort.cu:610-613;
 * its Python caller operators/partition_sort.py:125-127).
 *
 * Stable hash partition of one tick into the padded per-bucket layout.
 * cols_in[c]: column c (ncols <= 6), row_bytes[c] bytes a row, raw (ndom
 * rows: the mask's, else the key's) or compacted by the mask (n_valid rows)
 * per comp[c]; the key (cols_in[key_idx]) is int64.  mask: one byte a
 * pre-mask row, or NULL.  n_valid: the mask's kept rows (ndom without a
 * mask).  cols_out[c] and gidx_out (int64) are sized n_valid + B*(u-1) rows,
 * u = lcm(row bytes, 512) / row bytes of the output.  host_out: pinned
 * int64[B + 5]; receives the B bucket counts, each row-width class's padded
 * total (classes numbered by first appearance among the columns, then the
 * gidx's 8 B) and, in [B + 4], the kept rows a mask prologue found (-1 when
 * none ran).  B: a power of two <= 512.  select_first: collect the kept
 * rows' ids first (only under a mask).  Waits once, for the counts. */
size_t pf_partition_sort_workspace_bytes(int64_t ndom, int64_t n_valid, int32_t B,
                                         int32_t has_mask, int32_t select_first,
                                         int32_t any_comp);
int pf_partition_sort(const void* const* cols_in, void* const* cols_out,
                      const int32_t* row_bytes, const int32_t* comp, int32_t ncols,
                      int32_t key_idx, const uint8_t* mask, int64_t ndom, int64_t n_valid,
                      int64_t goff, int32_t B, int32_t select_first, int64_t* gidx_out,
                      int64_t* host_out, void* workspace, size_t workspace_bytes,
                      void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PAGEFLOW_JOIN_H */
