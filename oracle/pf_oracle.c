/*
 * CPU oracle for the equi-join hot path.  TEST INFRASTRUCTURE ONLY: the
 * tests, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it as
 * the checker; the product (pageflow_amd) never does.
 *
 * Each function restates, in plain sequential C, the reference program it
 * cites (paths relative to the reference repository).  Where the reference is
 * a GPU program whose result is order-free (atomics, CUB select), the
 * restatement runs the same algorithm in row order, which yields the one
 * result the reference documents (probe-row order, first build row).
 *
 * Build: oracle/Makefile (gcc -O2 -shared -fPIC).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define EMPTY_KEY INT64_MIN

/* Murmur3 fmix64: operators/kernels/kernel_utils.cuh:18-22 (numpy twin
 * tests/mix64.py:12-21). */
static inline uint64_t mix64(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdULL;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ULL;
  k ^= k >> 33;
  return k;
}

void o_mix64(const uint64_t* in, uint64_t* out, int64_t n) {
  for (int64_t i = 0; i < n; ++i) out[i] = mix64(in[i]);
}

/* A key's radix bucket: the top logB bits of mix64, 0 when B is 1
 * (operators/kernels/partition_sort.cu:110-112; tests/mix64.py:30-36). */
void o_bucket_of(const int64_t* keys, int64_t n, int logB, int32_t* out) {
  for (int64_t i = 0; i < n; ++i)
    out[i] = logB ? (int32_t)(mix64((uint64_t)keys[i]) >> (64 - logB)) : 0;
}

/* The reference's first slot of a hash and its successor
 * (operators/kernels/join_core.cu:56-64): multiply-shift of the bits below
 * the top 9 onto [0, cap); linear probing with wrap-around. */
static inline uint64_t slot_of(uint64_t h, uint64_t cap) {
  return (uint64_t)(((unsigned __int128)(h << 9) * cap) >> 64);
}
static inline uint64_t next_slot(uint64_t s, uint64_t cap) { return s + 1 == cap ? 0 : s + 1; }

/* join_core.cu:66-74 */
static inline uint64_t tuple_hash(int64_t k0, int64_t k1, int two) {
  uint64_t h = mix64((uint64_t)k0);
  if (two) {
    uint64_t h1 = mix64((uint64_t)k1);
    h ^= h1 + 0x9E3779B97F4A7C15ULL + (h << 6) + (h >> 2);
    h = mix64(h);
  }
  return h;
}

/* unique_hash_pairs: operators/join_core.py:97-113 over
 * operators/kernels/join_core.cu:105-146, 301-351.  Table int32[2 nb] of
 * build rows, built in row order; each probe row walks from its first slot
 * until an empty slot or a build row with its tuple; the matched probe rows
 * in ascending order with their build rows.  Returns the pair count. */
int64_t o_unique_hash_pairs(const int64_t* b0, const int64_t* b1, int64_t nb,
                            const int64_t* p0, const int64_t* p1, int64_t np, int two,
                            int64_t* build_idx, int64_t* probe_idx) {
  if (nb == 0 || np == 0) return 0;
  const uint64_t cap = 2ULL * (uint64_t)nb;
  int32_t* table = (int32_t*)malloc(cap * sizeof(int32_t));
  if (!table) return -1;
  memset(table, 0xff, cap * sizeof(int32_t));
  for (int64_t i = 0; i < nb; ++i) {
    uint64_t s = slot_of(tuple_hash(b0[i], two ? b1[i] : 0, two), cap);
    while (table[s] != -1) s = next_slot(s, cap);
    table[s] = (int32_t)i;
  }
  int64_t n = 0;
  for (int64_t i = 0; i < np; ++i) {
    const int64_t k0 = p0[i], k1 = two ? p1[i] : 0;
    uint64_t s = slot_of(tuple_hash(k0, k1, two), cap);
    for (;;) {
      const int32_t b = table[s];
      if (b == -1) break;
      if (b0[b] == k0 && (!two || b1[b] == k1)) {
        build_idx[n] = b;
        probe_idx[n] = i;
        ++n;
        break;
      }
      s = next_slot(s, cap);
    }
  }
  free(table);
  return n;
}

/* semi_anti: operators/join_core.py:116-143 over join_core.cu:154-212.  A
 * hash set [2 nb] of the build keys (empty slot = int64 min) with, per key,
 * the min and/or max of the build value the residual operator reads
 * (op 1 '<>': both, 2 '<' / 3 '<=': min, 4 '>' / 5 '>=': max).  A probe row
 * is kept when it has a match passing build.v OP probe.v (semi) or none
 * (anti); kept rows ascending into out.  *bad = 1 when a build key equals
 * int64 min (the caller then falls back).  Returns the kept count. */
int64_t o_semi_anti(const int64_t* bk, int64_t nb, const int64_t* pk, int64_t np,
                    const int64_t* bv, const int64_t* pv, int op, int anti, int64_t* out,
                    int32_t* bad) {
  *bad = 0;
  if (np == 0) return 0;
  if (nb == 0) {
    if (!anti) return 0;
    for (int64_t i = 0; i < np; ++i) out[i] = i;
    return np;
  }
  const uint64_t cap = 2ULL * (uint64_t)nb;
  int64_t* table = (int64_t*)malloc(cap * sizeof(int64_t));
  int64_t* mn = (int64_t*)malloc(cap * sizeof(int64_t));
  int64_t* mx = (int64_t*)malloc(cap * sizeof(int64_t));
  if (!table || !mn || !mx) { free(table); free(mn); free(mx); return -1; }
  for (uint64_t s = 0; s < cap; ++s) { table[s] = EMPTY_KEY; mn[s] = INT64_MAX; mx[s] = INT64_MIN; }
  for (int64_t i = 0; i < nb; ++i) {
    const int64_t k = bk[i];
    if (k == EMPTY_KEY) { *bad = 1; continue; }
    uint64_t s = slot_of(mix64((uint64_t)k), cap);
    while (table[s] != EMPTY_KEY && table[s] != k) s = next_slot(s, cap);
    table[s] = k;
    if (op != 0) {
      if (bv[i] < mn[s]) mn[s] = bv[i];
      if (bv[i] > mx[s]) mx[s] = bv[i];
    }
  }
  int64_t n = 0;
  for (int64_t i = 0; i < np; ++i) {
    const int64_t k = pk[i];
    int hit = 0;
    uint64_t s = slot_of(mix64((uint64_t)k), cap);
    if (k != EMPTY_KEY) {
      for (;;) {
        if (table[s] == k) { hit = 1; break; }
        if (table[s] == EMPTY_KEY) break;
        s = next_slot(s, cap);
      }
    }
    if (hit && op != 0) {
      const int64_t v = pv[i];
      switch (op) {
        case 1: hit = mn[s] != v || mx[s] != v; break;
        case 2: hit = mn[s] < v; break;
        case 3: hit = mn[s] <= v; break;
        case 4: hit = mx[s] > v; break;
        default: hit = mx[s] >= v; break;
      }
    }
    if (hit != anti) out[n++] = i;
  }
  free(table); free(mn); free(mx);
  return n;
}

/* expand: operators/join_core.py:54-61 over join_core.cu:86-101, 263-295.
 * Probe row i's run sorted_ids[start[i] .. start[i] + len[i]) as pairs, in
 * probe-row order then run order.  With cs == NULL only counts. */
int64_t o_expand(const int64_t* sorted_ids, const int64_t* start, const int64_t* len, int64_t m,
                 int64_t* cs, int64_t* cr) {
  int64_t n = 0;
  for (int64_t i = 0; i < m; ++i) {
    if (cs) {
      for (int64_t j = 0; j < len[i]; ++j) {
        cs[n + j] = sorted_ids[start[i] + j];
        cr[n + j] = i;
      }
    }
    n += len[i];
  }
  return n;
}

/* SUM and COUNT per group: operators/operator_utils.py:342-399 (the
 * bincount / two-step scatter_reduce of KeyedAggregate.pass1,
 * operators/aggregate.py:1511-1563), restated in exact int64.  COUNT counts
 * the values that are not constants.null (-64). */
void o_group_sum_count(const int32_t* gid, const int64_t* val, int64_t n, int32_t G,
                       int64_t* sum, int64_t* cnt) {
  for (int32_t g = 0; g < G; ++g) { sum[g] = 0; cnt[g] = 0; }
  for (int64_t i = 0; i < n; ++i) {
    sum[gid[i]] += val[i];
    cnt[gid[i]] += val[i] != -64;
  }
}

/* Row gather with a null row where the index is constants.null:
 * conversion.py:39-50 (index_with_null) and operators/hashjoin.py:359-377
 * (_gather_join_result's index_select), for 8-byte rows. */
void o_gather_i64(const int64_t* src, int64_t nsrc, const int64_t* idx, int64_t n, int64_t* out) {
  for (int64_t i = 0; i < n; ++i)
    out[i] = (idx[i] == -64 || nsrc == 0) ? -64 : src[idx[i]];
}
