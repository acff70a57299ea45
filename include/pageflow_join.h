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

/* Per-kernel timing with HIP events (no counterpart in the reference, whose
 * timing is Nsight's): pf_profile(1) clears and starts recording an event
 * pair around every kernel the library launches, pf_profile(0) stops.
 * pf_profile_read waits for the recorded events and fills, per kernel name
 * (NUL-separated in `names`), the total milliseconds and the launches;
 * returns the kernels filled. */
int pf_profile(int32_t enable);
int32_t pf_profile_read(char* names, int64_t names_bytes, double* total_ms, int64_t* launches,
                        int32_t max_kernels);

/* ------------------------------------------------------------ hashing --
 * mix64 (kernel_utils.cuh:18-22) of n int64 keys into out (int64 bits).
 * Replaces the torch-op twin operators/aggregate.py:95-104. */
int pf_mix64(const int64_t* keys, int64_t n, int64_t* out, void* stream);

/* The radix bucket of n int64 hashes among B (a power of two): the top
 * log2(B) bits of mix64, 0 when B is 1, as int32.
 * Replaces operators/aggregate.py:107-116 (_bucket_ids). */
int pf_bucket_ids(const int64_t* keys, int64_t n, int32_t B, int32_t* out, void* stream);

/* A join key column of any scalar type as int64 join keys, for the pair path
 * (replaces the dtype handling the reference's sort/search core inherits
 * from torch.sort/searchsorted, operators/hashjoin.py:408-421).  key_type:
 * PF_KEY_*.  Integer keys become their int64 value.  Float keys, and
 * integer keys when as_float is set (an integer side joined against a float
 * side compares in float64), become the float64 bit pattern with -0.0
 * folded into +0.0, so equal values have equal images; NaN, which equals
 * nothing, becomes a NaN pattern of its own per side (side 0 build, 1
 * probe), so NaNs never match.  No host wait. */
#define PF_KEY_I64 0
#define PF_KEY_I32 1
#define PF_KEY_I16 2
#define PF_KEY_I8 3
#define PF_KEY_U8 4
#define PF_KEY_F64 5
#define PF_KEY_F32 6
#define PF_KEY_BOOL 7
int pf_join_key_bits(const void* keys, int32_t key_type, int64_t n, int32_t as_float,
                     int32_t side, int64_t* out, void* stream);

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

/* The partition of a partitioned aggregation's pass-1 tick, with the
 * buckets given.  Replaces operators/aggregate.py:120-152
 * (_partition_outputs; its buckets come from _bucket_ids, pf_bucket_ids
 * above).  cols_in[c]: n rows of row_bytes[c] bytes (ncols <= 6, at most 4
 * row widths); bucket: int32[n], each row's bucket in [0, B).  cols_out[c]
 * is sized n + B*(u-1) rows as above; rows keep their order inside a bucket
 * and each non-empty bucket starts on a whole lcm(row bytes, 512)-byte unit,
 * padding unwritten.  host_out: pinned int64[B + 6]: the B bucket counts,
 * each row-width class's padded total (classes by first appearance), -1,
 * and the rows whose bucket was out of range (dropped).  Waits once. */
size_t pf_partition_by_bucket_workspace_bytes(int64_t n, int32_t B);
int pf_partition_by_bucket(const void* const* cols_in, void* const* cols_out,
                           const int32_t* row_bytes, int32_t ncols, const int32_t* bucket,
                           int64_t n, int32_t B, int64_t* host_out, void* workspace,
                           size_t workspace_bytes, void* stream);

/* ------------------------------------------------------- unique hash --
 * Replaces tqpkernels.join_unique_hash (operators/kernels/join_core.cu:301-351;
 * Python caller operators/join_core.py:110-111).
 * Inner join on a unique build tuple of nkeys (1 or 2) int64 columns
 * (b1/p1 ignored for 1 key), nb, np >= 1.  build_idx/probe_idx: int64[np];
 * the first host_out[0] hold the matches in probe-row order, a probe row
 * taking the lowest build row with its tuple.  host_out: pinned int64[2].
 * Waits once. */
size_t pf_join_unique_hash_workspace_bytes(int64_t nb, int64_t np, int32_t nkeys);
int pf_join_unique_hash(const int64_t* b0, const int64_t* b1, int64_t nb, const int64_t* p0,
                        const int64_t* p1, int64_t np, int32_t nkeys, int64_t* build_idx,
                        int64_t* probe_idx, int64_t* host_out, void* workspace,
                        size_t workspace_bytes, void* stream);

/* pf_join_unique_hash with no workspace and no allocation of any kind.  It has
 * no counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under stream capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The table, partner and scan scratch live in one
 * fixed device array of PF_NOALLOC_SCRATCH_BYTES that each device holds from
 * the first call on it until the process ends; a call whose
 * pf_join_unique_hash_workspace_bytes(nb, np, nkeys) exceeds it fails with an
 * error and launches nothing.  The caller still passes the inputs, the
 * outputs (int64[np] each) and host_out, as for pf_join_unique_hash.  Every
 * call on a device shares that device's array: calls that overlap on one
 * device (other threads, other streams) race and must be serialised by the
 * caller.  Each call initialises all the scratch it reads.  Waits once. */
#define PF_NOALLOC_SCRATCH_BYTES (64u << 20)
int pf_join_unique_hash_noalloc(const int64_t* b0, const int64_t* b1, int64_t nb,
                                const int64_t* p0, const int64_t* p1, int64_t np,
                                int32_t nkeys, int64_t* build_idx, int64_t* probe_idx,
                                int64_t* host_out, void* stream);

/* ------------------------------------------------------- semi / anti --
 * Replaces tqpkernels.join_semi_anti (operators/kernels/join_core.cu:358-400;
 * Python caller operators/join_core.py:138-139).
 * op: 0 none, 1 '<>', 2 '<', 3 '<=', 4 '>', 5 '>=' (build.v OP probe.v,
 * bv/pv int64 value columns when op != 0).  out: int64[np], the first
 * host_out[0] the kept probe rows ascending; host_out[1] = 1 when a build
 * key equals int64 min (the result is then not valid).  Waits once. */
size_t pf_join_semi_anti_workspace_bytes(int64_t nb, int64_t np, int32_t op);
int pf_join_semi_anti(const int64_t* bk, int64_t nb, const int64_t* pk, int64_t np,
                      const int64_t* bv, const int64_t* pv, int32_t op, int32_t anti,
                      int64_t* out, int64_t* host_out, void* workspace, size_t workspace_bytes,
                      void* stream);

/* The same predicate as a row mask (no counterpart in the reference: the
 * mask form of semi/anti output for the skewed-join workload): mask_out[i] =
 * 1 when probe row i is kept.  bad_out (pinned, optional) receives the
 * int64-min-key flag.  No host wait. */
size_t pf_semi_anti_mask_workspace_bytes(int64_t nb, int64_t np, int32_t op);
int pf_semi_anti_mask(const int64_t* bk, int64_t nb, const int64_t* pk, int64_t np,
                      const int64_t* bv, const int64_t* pv, int32_t op, int32_t anti,
                      uint8_t* mask_out, int64_t* bad_out, void* workspace,
                      size_t workspace_bytes, void* stream);

/* ------------------------------------------------------------ expand --
 * Replaces tqpkernels.join_expand (operators/kernels/join_core.cu:263-295;
 * Python caller operators/join_core.py:60-61), split in two because the
 * caller owns the outputs: the count call scans len (m rows) into the
 * workspace and waits for the pair count (host_out[0]); the caller sizes
 * cs/cr (int64[n_flat]) and the write call, on the same workspace, writes
 * (sorted_ids[start[i] + j], i) for every probe row i and j < len[i], in
 * probe-row then run order. */
size_t pf_join_expand_workspace_bytes(int64_t m);
int pf_join_expand_count(const int64_t* len, int64_t m, int64_t* host_out, void* workspace,
                         size_t workspace_bytes, void* stream);
int pf_join_expand_write(const int64_t* sorted_ids, const int64_t* start, int64_t m,
                         int64_t n_flat, int64_t* cs, int64_t* cr, void* workspace,
                         void* stream);

/* --------------------------------------------------- radix partition --
 * No single counterpart in the reference: the legacy partition
 * (pf_partition_sort, B <= 512, padded, int64 keys) generalised for the
 * batched and large-join tiers to 2^12 buckets, int32 or int64 keys and
 * int32 row ids, unpadded.  Stable: rows keep their input order in a bucket.
 * keys: n < 2^31 keys of key_bytes (8 or 4); vals: int32 row ids (NULL: the
 * row index).  mode 0: digit = the `bits` bits of mix64(key) below its top
 * `shift` bits, keys_out gets the keys; mode 1: the same digit, keys_out
 * gets mix64(key) as int64; mode 2: digit = bits [shift, shift + bits) of
 * the unsigned key.  bucket_offsets: device int64[2^bits + 1], each bucket's
 * first row and n.  No host wait. */
size_t pf_radix_partition_workspace_bytes(int64_t n, int32_t bits);
int pf_radix_partition(const void* keys, int32_t key_bytes, const int32_t* vals, int64_t n,
                       int32_t bits, int32_t shift, int32_t mode, void* keys_out,
                       int32_t* vals_out, int64_t* bucket_offsets, void* workspace,
                       size_t workspace_bytes, void* stream);
/* The same with the scatter's tile shape chosen by the caller (the
 * autotuner): shape 0..3 = (256 threads x 8 rows, 256 x 16, 512 x 8,
 * 512 x 16), 5 = 1024 x 8, a tile for digits of up to 8 bits; -1 the
 * default (from 64M rows 1024 x 8 up to 6-bit digits and 512 x 16 above,
 * 512 x 8 below 64M rows).  Above 8
 * bits the partition runs as two stable passes (the low half of the digit,
 * then the high half, each with that tile shape) and a bucket-bounds
 * kernel, unless shape is 4: one pass of the wide-digit scatter. */
int pf_radix_partition_shape(const void* keys, int32_t key_bytes, const int32_t* vals,
                             int64_t n, int32_t bits, int32_t shift, int32_t mode,
                             void* keys_out, int32_t* vals_out, int64_t* bucket_offsets,
                             void* workspace, size_t workspace_bytes, int32_t shape,
                             void* stream);

/* --------------------------------------------------- partitioned join --
 * The large-join tier (no single counterpart in the reference; it replaces
 * the per-call unique-hash core, join_core.cu:301-351, and the sort/search
 * core, operators/hashjoin.py:380-441 + join_core.cu:263-295, for whole
 * sides): both sides radix-partitioned by the top `bits` bits of mix64(key)
 * (bits <= 12), each bucket joined against an LDS hash table of its build
 * rows.  multi = 0: a unique build (a probe row takes the lowest build row
 * with its key); multi = 1: every matching build row, each build bucket
 * sorted by (hash, row) so a key's rows are one run (in LDS up to 8192
 * rows; a larger bucket by stable 8-bit LSD passes of the radix partition
 * over its hashes in HBM, after one more wait); multi = 2: a unique build
 * whose pairs may come in any order (no count pass, see below).  Two calls
 * on one workspace: prepare partitions, counts and waits, leaving in
 * host_out (pinned int64[3]) [0] the pair count (-1 for multi = 2), [1] the
 * largest build bucket, [2] the work items; write fills out_probe/out_build
 * (int64, at least the pair count rows; np rows for multi = 2) and gathers
 * npay (<= 2) int64 build payload columns by build row into pay_out.  With
 * multi 0 or 1 pairs come in a deterministic order (bucket, then probe row
 * in partition order, then build row ascending) and write does not wait
 * (host_out may be NULL); with multi = 2 each round of a work item takes its
 * output rows with an atomic add, so pairs come grouped by round in no
 * fixed order, and write waits once and leaves the pair count in
 * host_out[0].  Keys are int64; np, nb < 2^31. */
size_t pf_hash_join_workspace_bytes(int64_t nb, int64_t np, int32_t bits, int32_t multi);
int pf_hash_join_prepare(const int64_t* bkeys, int64_t nb, const int64_t* pkeys, int64_t np,
                         int32_t bits, int32_t multi, int64_t* host_out, void* workspace,
                         size_t workspace_bytes, void* stream);
int pf_hash_join_write(int64_t nb, int64_t np, int32_t bits, int32_t multi, int64_t n_items,
                       const int64_t* const* bpay, int32_t npay, int64_t* out_probe,
                       int64_t* out_build, int64_t* const* pay_out, int64_t* host_out,
                       void* workspace, size_t workspace_bytes, void* stream);

/* ------------------------------------------- residual inside the join --
 * Replaces the sort/search core's residual step (operators/hashjoin.py:
 * 421-439: the pair count, expand, the per-condition evaluate_join masks,
 * nonzero and the two index selects) for the partitioned join above: the
 * residual program (the PF_RES_* program of the pair filter below, column
 * side 0 = build, 1 = probe) is evaluated on every candidate pair inside the
 * count and write passes, so only kept pairs are counted and written.
 * The first two probe columns the program names are carried through the
 * probe side's partition (their values ride with the hashes), the first two
 * build columns are gathered into partition order (the first staged in
 * LDS); any other column is read at the pair's row.  Every column must hold
 * at least its side's rows.  multi: 0 unique build (a probe row's lowest
 * build row with its hash), 1 multi-match; pairs come in the deterministic
 * order of pf_hash_join_write.
 * workspace: pf_hash_join_residual_workspace_bytes with the program's
 * count of distinct probe-side and build-side columns.
 * prepare: partitions, counts kept pairs (two waits); host_out (pinned
 * int64[4]) [0] kept pairs, [1] the largest build bucket, [2] work items,
 * [3] candidate pairs before the residual (the pair count the reference's
 * INT_ROWS_MAX check reads).  write: the kept pairs (and npay gathered
 * int64 build payloads) into arrays of host_out[0] rows; no wait.  The
 * program passed to write must be the one passed to prepare. */
size_t pf_hash_join_residual_workspace_bytes(int64_t nb, int64_t np, int32_t bits, int32_t multi,
                                             int32_t probe_cols, int32_t build_cols);
int pf_hash_join_residual_prepare(const int64_t* bkeys, int64_t nb, const int64_t* pkeys,
                                  int64_t np, int32_t bits, int32_t multi,
                                  const void* const* cols, const int32_t* col_types,
                                  const int32_t* col_sides, const int64_t* col_rows,
                                  int32_t n_cols, const int64_t* ins, int32_t n_ins,
                                  int64_t* host_out, void* workspace, size_t workspace_bytes,
                                  void* stream);
int pf_hash_join_residual_write(int64_t nb, int64_t np, int32_t bits, int32_t multi,
                                int64_t n_items, const int64_t* const* bpay, int32_t npay,
                                int64_t* out_probe, int64_t* out_build, int64_t* const* pay_out,
                                const void* const* cols, const int32_t* col_types,
                                const int32_t* col_sides, const int64_t* col_rows,
                                int32_t n_cols, const int64_t* ins, int32_t n_ins,
                                void* workspace, size_t workspace_bytes, void* stream);

/* --------------------------------------------------- keyed group-by --
 * Replaces the grouping of a keyed aggregation's pass 1
 * (operators/aggregate.py:1511-1563: the linear group hash of
 * _hash_key_columns, :551-564, then torch.unique(return_inverse)) and its
 * SUM/COUNT (operators/operator_utils.py:342-399).  Rows group by the exact
 * tuple of nkeys (1 or 2) int64 key columns k0, k1 (k1 ignored for 1 key),
 * n < 2^30 rows.  Each group's hash is the reference's linear hash
 * sum_i key_i * coeffs[i] (int64, wrapping); coeffs is a host array of
 * nkeys values.  Outputs (caller-allocated, n rows each; the first G hold
 * the groups, in ascending signed hash order): hash_out, key0_out (and
 * key1_out for 2 keys) the group's key tuple, sum_out/cnt_out the exact
 * int64 SUM(val) and COUNT(val != -64) when val is not NULL, bucket_out
 * (optional) the hash's bucket among B (the top log2 B bits of its mix64,
 * B a power of two <= 4096).  host_out (pinned int64[2 + B]): [0] G, [1]
 * the groups whose hash equals the previous group's (distinct tuples the
 * reference's hash would merge; 0 when the hash is injective on the
 * input), [2 + b] the groups of bucket b.  Waits twice (for G, then for
 * the counts). */
size_t pf_group_by_keys_workspace_bytes(int64_t n, int32_t nkeys, int32_t B);
int pf_group_by_keys(const int64_t* k0, const int64_t* k1, int32_t nkeys, int64_t n,
                     const int64_t* coeffs, const int64_t* val, int32_t B, int64_t* hash_out,
                     int64_t* key0_out, int64_t* key1_out, int64_t* sum_out, int64_t* cnt_out,
                     int32_t* bucket_out, int64_t* host_out, void* workspace,
                     size_t workspace_bytes, void* stream);

/* ------------------------------------------------------------ gather --
 * Replaces the payload gather of operators/hashjoin.py:359-377
 * (torch.index_select per output column) and conversion.py:39-50
 * (index_with_null): out[i] = src[idx[i]] for rows of row_elems elements of
 * elem_bytes (1, 2, 4, 8), copied as bits, so any element type gathers;
 * where idx[i] is the null sentinel (-64), or src has no row, every element
 * of the row is null_bits (its low elem_bytes bytes): the bits of -64 in the
 * column's type, -64.0 for a float column.  An index outside [0, nsrc) also
 * writes the null row and sets *bad (optional, device).  No wait. */
int pf_gather_rows(const void* src, int64_t nsrc, int32_t elem_bytes, int64_t row_elems,
                   const int64_t* idx, int64_t n, void* out, int64_t* bad, int64_t null_bits,
                   void* stream);

/* The join-type tails (operators/hashjoin.py:293-352, the bool scatter of
 * matched rows and its nonzero): pf_mark_rows sets flags[idx[i]] = value;
 * pf_select_rows writes the rows whose flag equals `want` (0 or nonzero),
 * ascending, into out (int64[n]) and their count into host_out[0] (pinned);
 * one wait. */
int pf_mark_rows(const int64_t* idx, int64_t n, uint8_t* flags, int64_t nflags, int32_t value,
                 void* stream);
size_t pf_select_rows_workspace_bytes(int64_t n);
int pf_select_rows(const uint8_t* flags, int64_t n, int32_t want, int64_t* out,
                   int64_t* host_out, void* workspace, size_t workspace_bytes, void* stream);

/* ---------------------------------------------------- residual filter --
 * Replaces the residual step on a join's pairs: the per-condition
 * evaluate_join masks, their `&=`, the nonzero and the two index selects of
 * operators/hashjoin.py:264-290 (_apply_fast_path_conditions) and :428-439
 * (the sort/search core), with evaluate_join itself at :235-242 and the
 * comparison semantics of operators/operator_utils.py:438-478.
 *
 * The conditions come as one postfix program of n_ins instructions, three
 * int64 words each (ins[3*k .. 3*k+2]):
 *   word 0, bits 0-7: PF_RES_CMP, PF_RES_AND, PF_RES_OR, PF_RES_NOT,
 *     PF_RES_TRUE (AND/OR pop two results and push one, NOT flips the top,
 *     TRUE pushes true, CMP pushes a comparison);
 *   for CMP: bits 8-15 the relation (PF_REL_*), PF_RES_F64 compares in
 *     float64 (else int64), PF_RES_A_COL / PF_RES_B_COL: operand a / b is a
 *     column slot (else a constant: the int64 value, or the float64's bits
 *     under PF_RES_F64); words 1 and 2 are operands a and b.
 * PF_REL_CLOSE is torch.isclose(a, b, rtol=1e-5, atol=1e-5) in float64,
 * the reference's float `=`.  The program leaves exactly one result, at
 * most 32 deep; at most PF_RES_MAX_INS instructions.
 * Column slot s (n_cols <= PF_RES_MAX_COLS): cols[s] (device, col_rows[s]
 * elements of col_types[s], PF_COL_*), read at left_idx[i] when
 * col_sides[s] is 0, else at right_idx[i].  An integer column converts to
 * float64 in a float64 comparison; a float64 column only compares in
 * float64.
 *
 * pf_pair_filter_count evaluates the program on the n pairs (left_idx[i],
 * right_idx[i]), writes mask_out[i] (optional, bool bytes: evaluate_join's
 * result), keeps its per-pair results in the workspace, and reads back
 * host_out[0] = pairs kept, host_out[1] = pair indexes outside their
 * column's rows (read as 0; the caller raises); one wait.
 * pf_pair_filter_write, on the same workspace, writes the kept pairs in
 * pair order into left_out / right_out (int64[host_out[0]]); no wait. */
#define PF_RES_CMP 1
#define PF_RES_AND 2
#define PF_RES_OR 3
#define PF_RES_NOT 4
#define PF_RES_TRUE 5
#define PF_REL_EQ 0
#define PF_REL_NE 1
#define PF_REL_LT 2
#define PF_REL_LE 3
#define PF_REL_GT 4
#define PF_REL_GE 5
#define PF_REL_CLOSE 6
#define PF_RES_F64 (1LL << 16)
#define PF_RES_A_COL (1LL << 17)
#define PF_RES_B_COL (1LL << 18)
#define PF_COL_I64 0
#define PF_COL_I32 1
#define PF_COL_I16 2
#define PF_COL_I8 3
#define PF_COL_U8 4
#define PF_COL_F64 5
#define PF_RES_MAX_COLS 16
#define PF_RES_MAX_INS 64
size_t pf_pair_filter_workspace_bytes(int64_t n);
int pf_pair_filter_count(const int64_t* left_idx, const int64_t* right_idx, int64_t n,
                         const void* const* cols, const int32_t* col_types,
                         const int32_t* col_sides, const int64_t* col_rows, int32_t n_cols,
                         const int64_t* ins, int32_t n_ins, uint8_t* mask_out,
                         int64_t* host_out, void* workspace, size_t workspace_bytes,
                         void* stream);
int pf_pair_filter_write(const int64_t* left_idx, const int64_t* right_idx, int64_t n,
                         int64_t* left_out, int64_t* right_out, void* workspace,
                         void* stream);

/* ---------------------------------------------------------- group-by --
 * Replaces the SUM/COUNT evaluation of operators/operator_utils.py:342-399
 * under KeyedAggregate.pass1 (operators/aggregate.py:1511-1563), in exact
 * int64: sum[g] += val[i], cnt[g] += (val[i] != -64) for gid[i] = g in
 * [0, G); other gids are skipped and counted into *bad (optional, device).
 * accumulate = 0 zeroes sum/cnt first.  No wait. */
int pf_group_sum_count(const int32_t* gid, const int64_t* val, int64_t n, int32_t G,
                       int64_t* sum, int64_t* cnt, int32_t accumulate, int64_t* bad,
                       void* stream);

/* --------------------------------------------------------- serving --
 * No counterpart in the reference (the serving tier): a resident dimension
 * table, int32 key -> int32 group in [0, G), G <= 1024, in `slots` uint64
 * of memory (pf_dim_table_slots: a power of two >= 2n).  info (device
 * int64[4]) receives [0] the groups out of range, [1] the repeated keys,
 * [2] the smallest key and [3] the layout: keys spanning at most 4n values
 * with G <= 255 are stored dense (a byte a key value; [3] = the key range),
 * others in an open-addressing table ([3] = 0).  The table and its info go
 * together to every lookup.  pf_fact_join_groupby joins n fact rows (int32
 * foreign key, int64 amount) to the table and writes SUM(amount) and
 * COUNT(amount) (values != -64) per group into sum/cnt (int64[G], zeroed
 * first); fact rows without a match are dropped.  No wait. */
int64_t pf_dim_table_slots(int64_t n_dim);
int pf_dim_table_build(const int32_t* key, const int32_t* grp, int64_t n, int32_t G,
                       uint64_t* table, int64_t slots, int64_t* info, void* stream);
int pf_fact_join_groupby(const int32_t* fk, const int64_t* amt, int64_t n,
                         const uint64_t* table, int64_t slots, const int64_t* info, int32_t G,
                         int64_t* sum, int64_t* cnt, void* stream);

/* One (request, fact rows) item of a serving step.  No counterpart in the
 * reference (the serving tier).  fk/amt: the request's fact rows to add in
 * this step (n of them); table/slots/info/G: the request's resident
 * dimension table and its info (pf_dim_table_build); sum/cnt: the request's
 * int64[G] accumulators, zeroed by the caller before its first step. */
#define PF_SERVE_MAX_ITEMS 40
typedef struct pf_serve_item {
  const int32_t* fk;
  const int64_t* amt;
  int64_t n;
  const uint64_t* table;
  int64_t slots;
  const int64_t* info;
  int64_t* sum;
  int64_t* cnt;
  int32_t G;
  int32_t reserved;
} pf_serve_item;

/* One launch over up to PF_SERVE_MAX_ITEMS items (host array, passed to the
 * kernel by value: nothing is copied or allocated), rows_per_wg fact rows a
 * workgroup: SUM(amount), COUNT(amount != -64) per dimension group added to
 * each item's accumulators.  Items of one request in different steps add
 * up to the request's result.  No wait. */
int pf_serve_step(const pf_serve_item* items, int32_t n_items, int32_t rows_per_wg,
                  void* stream);

/* ------------------------------------------------- zero-copy reads --
 * Replace the reference's zero-copy reader extension
 * (IO/zero_copy_reader.cu:224-318, pybind module TQPlib.zero_copy_reader):
 * reads of a column held in pinned host memory, by the GPU over its host
 * link.  The column is num_segs segments: seg_ptrs[s] (device int64) the
 * device address of segment s's first row, seg_rows[s] (device int64,
 * num_segs + 1 entries, ascending) its first row in the column's row space,
 * seg_rows[num_segs] the end.  Row i of a read is row row_start + i.  Rows
 * are width elements of elem_bytes (1, 2, 4, 8; is_float: 4 or 8).  A row
 * outside every segment is not read: its output is -64 and it is counted
 * into *bad (device int64, added to); the reference traps instead.  No
 * wait.
 *
 * pf_device_alias replaces device_alias (zero_copy_reader.cu:224-233): the
 * device address of a pinned host pointer. */
int pf_device_alias(const void* host, void** dev);

/* Replaces zero_copy_rearrange (zero_copy_reader.cu:267-292, kernel
 * :87-120): out[r] (n rows) = row row_start + rearrange[r], or -64 in every
 * element when rearrange[r] is -64 (the null row).  n < 2^31.  The indices
 * are radix-sorted on the GPU first (from 2^16 of them), so host rows are
 * read in ascending order. */
size_t pf_zero_copy_rearrange_workspace_bytes(int64_t n);
int pf_zero_copy_rearrange(const int64_t* seg_ptrs, const int64_t* seg_rows, int32_t num_segs,
                           int64_t row_start, int64_t row_stop, const int64_t* rearrange,
                           int64_t n, int32_t elem_bytes, int32_t is_float, int64_t width,
                           void* out, int64_t* bad, void* workspace, size_t workspace_bytes,
                           void* stream);

/* Replaces zero_copy_mask (zero_copy_reader.cu:235-265, kernel :54-80):
 * out[k] = row row_start + i for the k-th i < row_stop - row_start with
 * mask[i] != 0 (uint8, one a row); at most n_out rows are written and the
 * kept count is stored into *kept (device int64). */
size_t pf_zero_copy_mask_workspace_bytes(int64_t n);
int pf_zero_copy_mask(const int64_t* seg_ptrs, const int64_t* seg_rows, int32_t num_segs,
                      int64_t row_start, int64_t row_stop, const uint8_t* mask, int32_t elem_bytes,
                      int32_t is_float, int64_t width, void* out, int64_t n_out,
                      int64_t* kept, int64_t* bad, void* workspace, size_t workspace_bytes,
                      void* stream);

/* ---------------------------------------------------------- batching --
 * No counterpart in the reference (the batched tier).  A plan for njoins
 * independent inner joins on int32 keys (unique builds): nb[j] build and
 * np[j] probe rows, stored back to back in the key buffers.  tune: NULL or
 * 12 int32, per size class of the build (< 32K rows, < 512K, larger) the
 * table slots per build row (2 or 4), the probe rows per work item (a power
 * of two, 256 to 2^20), the probe/compaction workgroup size (256 or 512)
 * and the probe slice size in KiB (0, or 64 to 2^20): a join whose table is
 * larger has its probe rows partitioned by the top bits of their table slot
 * (at most 8 bits, one pass), its build rows too, and joined slice by
 * slice, a work item building the table of 2^14 slots of its slice in LDS
 * (a larger slice is taken in several such sub-ranges).  The plan owns its
 * device scratch (one allocation at creation).
 *
 * pf_batch_join writes join j's pairs (join-local int64 probe row, build
 * row; probe-row order, within each slice for a sliced join, slices
 * ascending) at out_probe/out_build + (sum of np[i], i < j),
 * their count at counts[j] (device).  mode 0: the whole batch as one
 * pipeline of persistent kernels fed by device work queues (no host wait);
 * mode 1: the same kernels one join at a time with one host wait a join
 * (the plain baseline).  pf_batch_capture records mode 0 on `stream` (not
 * the null stream) as a hipGraph bound to these buffers; pf_batch_replay
 * launches it.  pf_batch_plan_info: 0 build rows, 1 probe rows, 2 table
 * slots, 3 build items, 4 probe items, 5 scratch bytes, 6 sliced joins, 7
 * the last run's LDS tables whose build rows took more than one pass (waits
 * for the device). */
void* pf_batch_plan_create(int32_t njoins, const int64_t* nb, const int64_t* np,
                           const int32_t* tune);
int pf_batch_plan_destroy(void* plan);
int64_t pf_batch_plan_info(void* plan, int32_t what);
int pf_batch_join(void* plan, const int32_t* bkeys, const int32_t* pkeys, int64_t* out_probe,
                  int64_t* out_build, int64_t* counts, int32_t mode, void* stream);
int pf_batch_capture(void* plan, const int32_t* bkeys, const int32_t* pkeys, int64_t* out_probe,
                     int64_t* out_build, int64_t* counts, void* stream);
int pf_batch_replay(void* plan, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PAGEFLOW_JOIN_H */
