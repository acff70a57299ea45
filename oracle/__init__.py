"""CPU oracle of the equi-join hot path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s cpu_baseline leg import this package, and only as the checker
or the timed CPU baseline; the product package ``pageflow_amd`` never imports
it and has no CPU fallback.

The restatement is plain C (``pf_oracle.c``, loaded through ctypes) for the
hash programs, and numpy for the sort/partition/layout arithmetic.  Every
function cites the reference file:line it follows (paths relative to the
reference repository).

Parity pinning (what checks each function besides the formula itself):

* ``mix64_np`` / ``mix64``: golden (key, mix64) vectors written from the
  reference's own numpy twin (``tests/mix64.py``) by
  ``tests/golden/make_mix64_golden.py``; the two restatements (numpy and C)
  agree; fmix64 is a bijection (checked against its inverse) and fmix64(0)=0.
* ``partition``: the reference test's stable-sort definition
  (``tests/operators/test_partition_sort.py:106-118``) and the padded layout
  invariants (``:93-103``), plus brute force per bucket.
* ``radix_partition``: each row's digit computed with Python ints from the
  golden hashes (hashed mode, every shift) or from the key read as an
  unsigned integer of its width (raw mode, int32 and int64), and the buckets
  filled by a brute-force scan in input order; at shift 0 it is the
  reference's stable sort by ``bucket_of``
  (``tests/operators/test_partition_sort.py:106-118``).
* ``unique_hash_pairs``: the known answers of
  ``tests/operators/test_unique_hash_join.py:86-151`` and a brute-force
  dense comparison.
* ``semi_anti``: the dense reference (``tests/operators/test_join_core.py:93-101``)
  and the known answers ``:121-144``.
* ``expand``: the ``repeat_interleave`` reference (``test_join_core.py:55-61``)
  and its hand cases (``:64-77``).
* ``sort_search_pairs``: the dense build x probe comparison
  (``test_join_core.py:263-281``).
* ``group_sum_count``: closed form (per-group sums by brute force).
* ``keyed_pass1``: a brute force over Python dicts keyed by the tuple, the
  hash computed with Python integers reduced mod 2^64, and the bucket by
  the golden-pinned ``bucket_of``.
* ``partition_outputs``: each bucket's rows in source order (brute force per
  bucket) at the starts ``padded_layout`` gives, padding unused.
* ``fmix32_np`` (the batch tier's table slots; placement only, no result
  depends on it): Murmur3's published 32-bit finalizer, pinned as a
  bijection by its algebraic inverse (the multipliers' inverses mod 2^32
  computed by Python), by fmix32(0) = 0, and against the same steps in
  Python integers.
* ``zero_copy_rearrange`` / ``zero_copy_mask``: the reference test's closed
  forms (``tests/IO/test_scan_regions.py:187-218``: ``x[lo:hi][idx]`` with
  null rows -64, ``x[lo:hi][mask]``) over the same column cut into
  segments of random sizes.
"""
from __future__ import annotations

import ctypes
import os
from typing import List, Optional, Sequence, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_I64_MIN = np.iinfo(np.int64).min
NULL = -64  # constants.py:9


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libpf_oracle.so")
        if not os.path.exists(path):
            import subprocess
            subprocess.run(["make", "-s", "-C", _HERE], check=True)
        lib = ctypes.CDLL(path)
        P = ctypes.c_void_p
        I64 = ctypes.c_int64
        I32 = ctypes.c_int32
        lib.o_mix64.argtypes = [P, P, I64]
        lib.o_bucket_of.argtypes = [P, I64, ctypes.c_int, P]
        lib.o_unique_hash_pairs.argtypes = [P, P, I64, P, P, I64, ctypes.c_int, P, P]
        lib.o_unique_hash_pairs.restype = I64
        lib.o_semi_anti.argtypes = [P, I64, P, I64, P, P, ctypes.c_int, ctypes.c_int, P, P]
        lib.o_semi_anti.restype = I64
        lib.o_expand.argtypes = [P, P, P, I64, P, P]
        lib.o_expand.restype = I64
        lib.o_group_sum_count.argtypes = [P, P, I64, I32, P, P]
        lib.o_gather_i64.argtypes = [P, I64, P, I64, P]
        _LIB = lib
    return _LIB


def _p(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _i64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a), dtype=np.int64)


# ---------------------------------------------------------------- mix64 --

_M1 = np.uint64(0xff51afd7ed558ccd)
_M2 = np.uint64(0xc4ceb9fe1a85ec53)
_S33 = np.uint64(33)


def mix64_np(keys) -> np.ndarray:
    """mix64 of int64 keys as uint64, in numpy uint64 arithmetic
    (kernel_utils.cuh:18-22; the reference's own twin tests/mix64.py:12-21)."""
    k = np.ascontiguousarray(np.asarray(keys), dtype=np.int64).view(np.uint64)
    with np.errstate(over="ignore"):
        k = k ^ (k >> _S33)
        k = k * _M1
        k = k ^ (k >> _S33)
        k = k * _M2
        k = k ^ (k >> _S33)
    return k


_F1 = np.uint32(0x85ebca6b)
_F2 = np.uint32(0xc2b2ae35)


def fmix32_np(keys) -> np.ndarray:
    """Murmur3's 32-bit finalizer of each key's 32 bits, as uint32 (the
    batch tier's table slot hash; the reference hashes with mix64 only, and
    no join result depends on the slot)."""
    h = np.ascontiguousarray(np.asarray(keys)).astype(np.int64).astype(np.uint32)
    with np.errstate(over="ignore"):
        h = h ^ (h >> np.uint32(16))
        h = h * _F1
        h = h ^ (h >> np.uint32(13))
        h = h * _F2
        h = h ^ (h >> np.uint32(16))
    return h


def mix64(keys) -> np.ndarray:
    """mix64 through the C restatement (uint64)."""
    k = _i64(keys)
    out = np.empty(k.shape[0], dtype=np.uint64)
    _lib().o_mix64(_p(k), _p(out), k.shape[0])
    return out


def bucket_of(keys, B: int) -> np.ndarray:
    """Each key's radix bucket among B: the top log2(B) bits of mix64
    (partition_sort.cu:110-112; tests/mix64.py:30-36)."""
    k = _i64(keys)
    log_b = B.bit_length() - 1
    out = np.empty(k.shape[0], dtype=np.int32)
    _lib().o_bucket_of(_p(k), k.shape[0], log_b, _p(out))
    return out


# ------------------------------------------------------------ partition --

def rows_per_unit(row_bytes: int) -> int:
    """Rows of one padding unit, lcm(row bytes, 512) / row bytes
    (partition_sort.cu:582; IO/codec/batched_compress.py:27-62)."""
    from math import gcd
    return 512 // gcd(row_bytes, 512)


def partition(cols: Sequence[np.ndarray], key_pos: int, mask: Optional[np.ndarray],
              goff: int, B: int) -> Tuple[List[int], List[List[np.ndarray]]]:
    """The bucket counts and, per output (each column, then the gidx), each
    bucket's rows: the kept rows stably ordered by bucket, gidx = goff + the
    pre-mask row (operators/partition_sort.py:68-143, partition_sort.cu:1-31;
    the reference test's definition tests/operators/test_partition_sort.py:106-118).
    ``cols`` are raw (pre-mask) columns."""
    n_pre = mask.shape[0] if mask is not None else cols[key_pos].shape[0]
    gidx = np.arange(goff, goff + n_pre, dtype=np.int64)
    if mask is not None:
        keep = mask.astype(bool)
        cols = [c[keep] for c in cols]
        gidx = gidx[keep]
    bucket = bucket_of(cols[key_pos], B)
    order = np.argsort(bucket, kind="stable")
    counts = np.bincount(bucket, minlength=B).astype(np.int64)
    splits = np.cumsum(counts)[:-1]
    outs = [list(np.split(c[order], splits)) for c in list(cols) + [gidx]]
    return counts.tolist(), outs


def partition_outputs(outs: Sequence[np.ndarray], bucket, B: int
                      ) -> Tuple[List[int], List[np.ndarray]]:
    """A partitioned aggregation's pass-1 tick in the padded layout with the
    buckets given (operators/aggregate.py:120-152): the bucket counts and,
    per output, the padded array (padding rows 0; the reference leaves them
    unwritten).  Restates the reference's steps: a stable sort of the
    buckets, each row's sorted position, and the padding of the buckets
    before its own added to it."""
    b = np.asarray(bucket, dtype=np.int64)
    n = b.shape[0]
    counts = np.bincount(b, minlength=B).astype(np.int64)
    order = np.argsort(b, kind="stable")
    pos = np.empty(n, np.int64)
    pos[order] = np.arange(n)
    padded = []
    for t in outs:
        rb = t.dtype.itemsize * int(np.prod(t.shape[1:], dtype=np.int64))
        unit = rows_per_unit(rb)
        pad = -(-counts // unit) * unit - counts
        shift = np.cumsum(pad) - pad
        dest = pos + shift[b]
        total = int(counts.sum() + pad.sum())
        out = np.zeros((total,) + t.shape[1:], dtype=t.dtype)
        out[dest] = t
        padded.append(out)
    return counts.tolist(), padded


def padded_layout(counts: Sequence[int], row_bytes: int) -> Tuple[List[int], int]:
    """Each bucket's first row and the padded total: every non-empty bucket
    starts on a whole lcm(row bytes, 512)-byte unit
    (IO/codec/batched_compress.py:27-62; IO/pipeline/common.py:36-45)."""
    unit = rows_per_unit(row_bytes)
    starts, at = [], 0
    for c in counts:
        starts.append(at)
        at += -(-int(c) // unit) * unit
    return starts, at


def radix_partition(keys: np.ndarray, vals: np.ndarray, bits: int, shift: int,
                    hashed: bool) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Stable partition of (key, value) rows by a digit of ``bits`` bits:
    the top bits of mix64(key) below the top ``shift`` bits when ``hashed``
    (partition_sort.cu:110-112 generalised to 2^12 buckets), else bits
    [shift, shift + bits) of the unsigned key.  Returns (keys, values,
    per-bucket counts)."""
    k = np.ascontiguousarray(keys)
    if hashed:
        h = mix64(k.astype(np.int64))
        d = ((h << np.uint64(shift)) >> np.uint64(64 - bits)).astype(np.int64) if bits else \
            np.zeros(k.shape[0], np.int64)
    else:
        u = k.astype(np.int64).view(np.uint64) if k.dtype == np.int64 else \
            k.astype(np.uint32).astype(np.uint64)
        d = ((u >> np.uint64(shift)) & np.uint64((1 << bits) - 1)).astype(np.int64)
    order = np.argsort(d, kind="stable")
    counts = np.bincount(d, minlength=1 << bits).astype(np.int64)
    return k[order], np.ascontiguousarray(vals)[order], counts


# ------------------------------------------------------------ join core --

def unique_hash_pairs(build_keys: Sequence[np.ndarray],
                      probe_keys: Sequence[np.ndarray]) -> Tuple[np.ndarray, np.ndarray]:
    """(build rows, probe rows) of an inner join on a unique build tuple of
    1-2 int64 keys, in probe-row order (join_core.py:97-113,
    join_core.cu:105-146, 301-351)."""
    two = len(build_keys) == 2
    b0, p0 = _i64(build_keys[0]), _i64(probe_keys[0])
    b1 = _i64(build_keys[1]) if two else b0
    p1 = _i64(probe_keys[1]) if two else p0
    nb, np_ = b0.shape[0], p0.shape[0]
    bi = np.empty(max(np_, 1), np.int64)
    pi = np.empty(max(np_, 1), np.int64)
    n = _lib().o_unique_hash_pairs(_p(b0), _p(b1), nb, _p(p0), _p(p1), np_, int(two),
                                   _p(bi), _p(pi))
    if n < 0:
        raise MemoryError("oracle: unique_hash_pairs table allocation failed")
    return bi[:n].copy(), pi[:n].copy()


SUMMARY_OPS = {'<>': 1, '<': 2, '<=': 3, '>': 4, '>=': 5}  # join_core.py:35


def semi_anti(build_key, probe_key, anti: bool, summary=None) -> Optional[np.ndarray]:
    """Probe rows, ascending, with (semi) or without (anti) a match; None
    when a build key is int64 min (join_core.py:116-143,
    join_core.cu:154-212, 358-400)."""
    bk, pk = _i64(build_key), _i64(probe_key)
    bv = pv = None
    op = 0
    if summary is not None:
        bv, pv, name = _i64(summary[0]), _i64(summary[1]), summary[2]
        op = SUMMARY_OPS[name]
    out = np.empty(max(pk.shape[0], 1), np.int64)
    bad = np.zeros(1, np.int32)
    n = _lib().o_semi_anti(_p(bk), bk.shape[0], _p(pk), pk.shape[0], _p(bv), _p(pv), op,
                           int(bool(anti)), _p(out), _p(bad))
    if n < 0:
        raise MemoryError("oracle: semi_anti table allocation failed")
    if bad[0]:
        return None
    return out[:n].copy()


def expand(sorted_ids, start, length) -> Tuple[np.ndarray, np.ndarray]:
    """(build rows, probe rows) of each probe row's run, probe-row then run
    order (join_core.py:54-61, join_core.cu:86-101, 263-295)."""
    s, st, ln = _i64(sorted_ids), _i64(start), _i64(length)
    m = ln.shape[0]
    n = _lib().o_expand(_p(s), _p(st), _p(ln), m, None, None)
    cs = np.empty(max(n, 1), np.int64)
    cr = np.empty(max(n, 1), np.int64)
    _lib().o_expand(_p(s), _p(st), _p(ln), m, _p(cs), _p(cr))
    return cs[:n].copy(), cr[:n].copy()


def sort_search_pairs(left, right, bucket_counts=None) -> Tuple[np.ndarray, np.ndarray]:
    """The sort/search core's (left rows, right rows) (hashjoin.py:380-441):
    per bucket, the left keys sorted (stably here; the reference's sort is
    unstable, so only the multiset is pinned), each right row's run found by
    two searchsorted, every run expanded."""
    lk, rk = _i64(left), _i64(right)
    nL, nR = lk.shape[0], rk.shape[0]
    lc, rc = bucket_counts if bucket_counts is not None else ([nL], [nR])
    sorted_ids = np.empty(nL, np.int64)
    start = np.zeros(nR, np.int64)
    length = np.zeros(nR, np.int64)
    ls = rs = 0
    for lcnt, rcnt in zip(lc, rc):
        le, re_ = ls + lcnt, rs + rcnt
        if lcnt > 0:
            perm = np.argsort(lk[ls:le], kind="stable")
            sk = lk[ls:le][perm]
            sorted_ids[ls:le] = perm + ls
            if rcnt > 0:
                rkb = rk[rs:re_]
                lo = np.searchsorted(sk, rkb, side="left")
                length[rs:re_] = np.searchsorted(sk, rkb, side="right") - lo
                start[rs:re_] = lo + ls
        ls, rs = le, re_
    return expand(sorted_ids, start, length)


def join_pairs_sorted(lk, rk) -> np.ndarray:
    """All (left row, right row) pairs of an inner equi-join, as an int64
    [n, 2] array sorted by (left, right) (the multiset of
    hashjoin.py:380-441)."""
    l, r = sort_search_pairs(lk, rk)
    pairs = np.stack([l, r], axis=1) if l.size else np.empty((0, 2), np.int64)
    if pairs.shape[0]:
        order = np.lexsort((pairs[:, 1], pairs[:, 0]))
        pairs = pairs[order]
    return pairs


# --------------------------------------------------------------- gather --

def gather_with_null(src, idx) -> np.ndarray:
    """conversion.py:39-50 for 1-D int64 rows."""
    s, i = _i64(src), _i64(idx)
    out = np.empty(max(i.shape[0], 1), np.int64)
    _lib().o_gather_i64(_p(s), s.shape[0], _p(i), i.shape[0], _p(out))
    return out[:i.shape[0]].copy()


# ------------------------------------------------------- zero-copy reads --

def _segment_row(parts: Sequence[np.ndarray], first_row: int, row: int):
    """The segment holding row-space row ``row`` and its row there, or None
    outside every segment (zero_copy_reader.cu:41-52: the last segment
    whose first row is <= row)."""
    start = first_row
    for p in parts:
        if start <= row < start + p.shape[0]:
            return p, row - start
        start += p.shape[0]
    return None


def zero_copy_rearrange(parts: Sequence[np.ndarray], first_row: int, rearrange,
                        row_start: int) -> np.ndarray:
    """zero_copy_reader.cu:87-120: row r of the result is row
    ``row_start + rearrange[r]`` of a column held as segments ``parts``
    (2-D numpy arrays, back to back from row ``first_row``), or -64 in every
    element where ``rearrange[r]`` is -64.  A row outside every segment
    raises (the reference traps).  Order of host reads is not observable."""
    idx = np.asarray(rearrange, dtype=np.int64)
    width, dtype = parts[0].shape[1], parts[0].dtype
    out = np.empty((idx.shape[0], width), dtype=dtype)
    # segments back to back: row-space row first_row + k is row k of their
    # concatenation, each segment holding [start, start + len)
    column = parts[0] if len(parts) == 1 else np.concatenate(parts)
    rows = row_start + idx
    valid = idx != NULL
    outside = valid & ((rows < first_row) | (rows >= first_row + column.shape[0]))
    if outside.any():
        r = int(rows[np.nonzero(outside)[0][0]])
        assert _segment_row(parts, first_row, r) is None
        raise IndexError(f"row {r} lies outside every segment")
    out[valid] = column[rows[valid] - first_row]
    # -64 converted to the column's type as C converts it (unsigned: wraps)
    out[~valid] = np.array(NULL, dtype=np.int64).astype(dtype)
    return out


def zero_copy_mask(parts: Sequence[np.ndarray], first_row: int, mask,
                   row_start: int) -> np.ndarray:
    """zero_copy_reader.cu:54-80: the rows ``row_start + i`` for each i
    with ``mask[i]`` true, in order."""
    keep = np.nonzero(np.asarray(mask, dtype=bool))[0].astype(np.int64)
    return zero_copy_rearrange(parts, first_row, keep, row_start)


# ------------------------------------------------------------- group-by --

def group_sum_count(gid, val, G: int) -> Tuple[np.ndarray, np.ndarray]:
    """Exact int64 SUM and COUNT per group (operator_utils.py:342-399)."""
    g = np.ascontiguousarray(np.asarray(gid), dtype=np.int32)
    v = _i64(val)
    s = np.empty(G, np.int64)
    c = np.empty(G, np.int64)
    _lib().o_group_sum_count(_p(g), _p(v), g.shape[0], G, _p(s), _p(c))
    return s, c


def keyed_pass1(keys: Sequence[np.ndarray], vals: Optional[np.ndarray],
                coeffs: Sequence[int], B: int = 1):
    """Pass 1 of a keyed aggregation for int64 keys (aggregate.py:1511-1563):
    the rows grouped by their exact key tuple; per group the linear grouping
    hash sum_i key_i * coeffs[i] in wrapping int64 (_hash_key_columns,
    aggregate.py:551-564), the key tuple, SUM(vals) and COUNT(vals != -64)
    (operator_utils.py:342-399), in ascending hash order as torch.unique
    orders them; and each group's bucket among B (bucket_of the hash) with
    the bucket counts.  Returns (hash, [key columns], sums, counts, buckets,
    bucket counts); sums and counts are None without vals."""
    ks = [_i64(k) for k in keys]
    n = ks[0].shape[0]
    with np.errstate(over="ignore"):
        h = np.zeros(n, np.int64)
        for k, c in zip(ks, coeffs):
            h += k * np.int64(c)
    if n == 0:
        e = np.empty(0, np.int64)
        return e, [e.copy() for _ in ks], (e if vals is not None else None), \
            (e.copy() if vals is not None else None), np.empty(0, np.int32), [0] * B
    table = np.stack(ks, axis=1)
    uniq, first, inv = np.unique(table, axis=0, return_index=True, return_inverse=True)
    inv = inv.reshape(-1)
    G = uniq.shape[0]
    gh = h[first]
    order = np.argsort(gh, kind="stable")
    sums = cnts = None
    if vals is not None:
        v = _i64(vals)
        sums = np.zeros(G, np.int64)
        np.add.at(sums, inv, v)
        cnts = np.bincount(inv, weights=(v != NULL), minlength=G).astype(np.int64)
        sums, cnts = sums[order], cnts[order]
    gh = gh[order]
    bucket = bucket_of(gh, B) if B > 1 else np.zeros(G, np.int32)
    return (gh, [uniq[order, i].copy() for i in range(len(ks))], sums, cnts, bucket,
            np.bincount(bucket, minlength=B).astype(np.int64).tolist())


# ------------------------------------------------- residual conditions --

_EPS = 1e-5  # constants.py:8
_REL = {'<=': np.less_equal, '>=': np.greater_equal, '<': np.less, '>': np.greater,
        '=': np.equal, '<>': np.not_equal}
_ARITH = {'+', '-', '*', '/'}


def _normalize(tree):
    while isinstance(tree, list) and len(tree) == 1:  # operator_utils.py:22-26
        tree = tree[0]
    return tree


def _col_id(token) -> Optional[int]:
    """parsing/expressions.py:17-32."""
    import re
    if not isinstance(token, str) or token.startswith('Brand#'):
        return None
    head = token.split(' as', 1)[0] if ' as' in token else token
    m = re.search(r'#(\d+)L?$', head)
    return int(m.group(1)) if m else None


def isclose_f64(a, b, rtol: float = 1e-5, atol: float = 1e-5) -> np.ndarray:
    """torch.isclose(a, b, rtol, atol) on float64 (equal_nan False): equal,
    or a finite |a - b| at most atol + |rtol * b|."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    with np.errstate(invalid="ignore", over="ignore"):
        actual = np.abs(a - b)
        allowed = atol + np.abs(rtol * b)
        return (a == b) | (np.isfinite(actual) & (actual <= allowed))


def _compare(lhs, rhs, op: str):
    """ops[op](lhs, rhs) with torch's operand rules
    (operator_utils.py:476-478): the left operand is a tensor; a Python int
    is cast to an integer tensor's dtype (wrapping) or to float64; a Python
    float against an integer tensor compares in float32 (the default
    dtype); two tensors compare in their promoted dtype."""
    if not isinstance(lhs, np.ndarray):
        raise TypeError(f"{op}: the left operand is a literal")
    if isinstance(rhs, np.ndarray):
        return _REL[op](lhs, rhs)
    if isinstance(rhs, (bool, np.bool_)):
        raise TypeError(f"{op}: a bool literal")
    if isinstance(rhs, int):
        if lhs.dtype.kind == 'f':
            return _REL[op](lhs, np.float64(rhs))
        if not -(1 << 63) <= rhs < (1 << 63):
            raise OverflowError(f"{op}: {rhs} outside int64")
        return _REL[op](lhs, np.array([rhs], dtype=np.int64).astype(lhs.dtype)[0])
    if isinstance(rhs, float):
        if lhs.dtype.kind == 'f':
            return _REL[op](lhs, np.float64(rhs))
        return _REL[op](lhs.astype(np.float32), np.float32(rhs))
    raise TypeError(f"{op}: operand {rhs!r}")


def residual_mask(left: dict, right: dict, tree, left_m, right_m) -> np.ndarray:
    """evaluate_join (hashjoin.py:235-242) through _evaluate_core in filter
    mode (operator_utils.py:235-478): a column id the left side holds is read
    at left_m, otherwise at right_m; the value of the condition per pair."""
    import datetime
    lm, rm = _i64(left_m), _i64(right_m)
    n = lm.shape[0]

    def leaf(cid):
        t = left[cid][lm] if cid in left else right[cid][rm]
        return np.squeeze(t)

    def ev(t):
        t = _normalize(t)
        if isinstance(t, str):  # :250-262
            if len(t) == 10 and t.count('-') == 2 and t.replace('-', '').isdigit():
                y, m, d = (int(x) for x in t.split('-'))  # conversion.py:24-27
                return (datetime.date(y, m, d) - datetime.date(1990, 1, 1)).days
            if t.replace('.', '').replace('-', '').replace('E', '').isdigit():
                return float(t) if '.' in t else int(t)
            cid = _col_id(t)
            return leaf(cid) if cid is not None else t
        if len(t) == 3 and t[1] == '#':  # :272-276
            return leaf(int(t[-1]))
        if len(t) == 2:
            if t[0] == 'NOT':  # :309-310
                return ~ev(t[1])
            if t[0] == 'isnotnull':  # :312-315
                return np.ones(n, dtype=bool)
            raise NotImplementedError(f"function {t[0]}")
        op = t[-2]
        if op == 'AS':  # :435-436
            return ev(t[0])
        if op in ('AND', 'OR'):
            a, b = ev(t[0]), ev(t[-1])
            return np.logical_and(a, b) if op == 'AND' else np.logical_or(a, b)
        if op in _ARITH or op in ('IN', 'LIKE'):
            raise NotImplementedError(f"operation {op}")
        if op not in _REL:
            raise ValueError(f"Unsupported operation: {op}")
        lhs, rhs = ev(t[0]), ev(t[-1])
        if isinstance(rhs, float) and op in ('<=', '>='):  # :458-462
            return _compare(lhs, rhs + _EPS if op == '<=' else rhs - _EPS, op)
        if isinstance(rhs, str):  # :464-472
            if len(rhs) == 1:
                return _compare(lhs, ord(rhs), op)
            raise NotImplementedError("string comparison")
        if op == '=' and isinstance(lhs, np.ndarray) and lhs.dtype == np.float64:  # :474-477
            if isinstance(rhs, float):
                rhs = np.full(lhs.shape, rhs, dtype=np.float64)
            if not isinstance(rhs, np.ndarray) or rhs.dtype != np.float64:
                raise TypeError("isclose: both operands float64 tensors")
            return isclose_f64(lhs, rhs)
        return _compare(lhs, rhs, op)

    out = ev(tree)
    return np.broadcast_to(np.asarray(out, dtype=bool), (n,)).copy()


def residual_filter(left: dict, right: dict, conditions, left_m, right_m
                    ) -> Tuple[np.ndarray, np.ndarray]:
    """hashjoin.py:276-290 / :428-439: the pairs every condition keeps, in
    pair order."""
    lm, rm = _i64(left_m), _i64(right_m)
    keep = np.ones(lm.shape[0], dtype=bool)
    for cond in conditions:
        keep &= residual_mask(left, right, cond, lm, rm)
    return lm[keep], rm[keep]
