"""The CPU oracle pinned before it is trusted: against the golden vectors
written from the reference's own mix64 twin, the known answers the
reference's tests hold (tests/operators/test_unique_hash_join.py,
tests/operators/test_join_core.py of the reference repository), brute force
on small inputs, and what the mathematics fixes (fmix64 is a bijection)."""
from __future__ import annotations

import os

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
I64 = np.iinfo(np.int64)


# ---------------------------------------------------------------- mix64 --

def test_mix64_matches_the_reference_twin_golden_vectors():
    d = np.load(os.path.join(HERE, "golden", "mix64_golden.npz"))
    assert np.array_equal(oracle.mix64(d["keys"]), d["mix64"])
    assert np.array_equal(oracle.mix64_np(d["keys"]), d["mix64"])
    assert np.array_equal(oracle.bucket_of(d["keys"], 512), d["bucket512"])
    assert np.array_equal(oracle.bucket_of(d["keys"], 1), d["bucket1"])


def _unmix64(h: np.ndarray) -> np.ndarray:
    """The inverse of fmix64: xorshift-33 is its own inverse on 64 bits, and
    each multiply is undone by the constant's inverse mod 2^64."""
    inv1 = np.uint64(pow(0xff51afd7ed558ccd, -1, 1 << 64))
    inv2 = np.uint64(pow(0xc4ceb9fe1a85ec53, -1, 1 << 64))
    s = np.uint64(33)
    k = h.copy()
    with np.errstate(over="ignore"):
        k ^= k >> s
        k *= inv2
        k ^= k >> s
        k *= inv1
        k ^= k >> s
    return k


def test_mix64_is_a_bijection_with_fixed_point_zero():
    keys = np.random.default_rng(3).integers(I64.min, I64.max, 50_000, dtype=np.int64)
    keys[:5] = [0, 1, -1, I64.min, I64.max]
    h = oracle.mix64(keys)
    assert np.array_equal(_unmix64(h).view(np.int64), keys)
    assert int(oracle.mix64(np.array([0]))[0]) == 0
    assert len(np.unique(h)) == len(np.unique(keys))


def test_buckets_spread_a_linear_hash():
    # the reference's spread check (test_agg_partition.py:70-80): 2^20
    # multiples of a coefficient over 256 buckets within a quarter of the mean
    h = np.arange(1, (1 << 20) + 1, dtype=np.int64) * 9_999_991
    counts = np.bincount(oracle.bucket_of(h, 256), minlength=256)
    assert counts.min() > 0.75 * 4096 and counts.max() < 1.25 * 4096


# ------------------------------------------------------- radix partition --
#
# oracle.radix_partition is the checker of the library's radix partition.  It
# is pinned here by an independent computation: the digit of each row from
# Python ints (the hash taken from the golden vectors, not recomputed), and
# the bucket contents by a brute-force scan in input order, which is the
# reference's stable-sort-by-bucket definition of a partition
# (tests/operators/test_partition_sort.py:106-118, buckets from
# tests/mix64.py:30-36).

def _golden():
    d = np.load(os.path.join(HERE, "golden", "mix64_golden.npz"))
    keys = d["keys"].astype(np.int64)
    hashes = [int(h) for h in d["mix64"].astype(np.uint64)]
    return keys, hashes


def _brute_partition(keys, vals, digits, nbuckets):
    """Bucket by bucket, the rows with that digit in input order."""
    out_k, out_v, counts = [], [], []
    for b in range(nbuckets):
        rows = [i for i in range(len(digits)) if digits[i] == b]
        out_k += [keys[i] for i in rows]
        out_v += [vals[i] for i in rows]
        counts.append(len(rows))
    return out_k, out_v, counts


@pytest.mark.parametrize("bits,shift", [(0, 0), (1, 0), (4, 0), (8, 0), (12, 0), (3, 5),
                                        (6, 6), (8, 4), (4, 60), (1, 63), (12, 52)])
def test_radix_partition_hashed_digits_by_python_ints(bits, shift):
    keys, hashes = _golden()
    rng = np.random.default_rng(bits * 64 + shift)
    order = rng.permutation(len(keys))
    keys = np.concatenate([keys[order], keys[order[:300]]])     # repeated keys too
    hashes = [hashes[i] for i in order] + [hashes[i] for i in order[:300]]
    vals = np.arange(len(keys), dtype=np.int32)[::-1].copy()
    # the digit: the `bits` bits of the hash just below its top `shift` bits
    digits = [(h >> (64 - shift - bits)) & ((1 << bits) - 1) for h in hashes]
    want_k, want_v, want_c = _brute_partition(keys.tolist(), vals.tolist(), digits, 1 << bits)
    got_k, got_v, got_c = oracle.radix_partition(keys, vals, bits, shift, hashed=True)
    assert got_k.tolist() == want_k and got_v.tolist() == want_v
    assert got_c.tolist() == want_c


@pytest.mark.parametrize("bits", [1, 4, 8, 9])
def test_radix_partition_at_shift_zero_is_the_reference_bucket_sort(bits):
    # bucket_of is pinned by the golden vectors; at shift 0 the partition is
    # the reference's stable sort by it
    keys, _ = _golden()
    keys = np.concatenate([keys, keys[:123]])
    vals = np.arange(len(keys), dtype=np.int32)
    buckets = oracle.bucket_of(keys, 1 << bits)
    order = np.argsort(buckets, kind="stable")
    got_k, got_v, got_c = oracle.radix_partition(keys, vals, bits, 0, hashed=True)
    assert np.array_equal(got_k, keys[order]) and np.array_equal(got_v, vals[order])
    assert np.array_equal(got_c, np.bincount(buckets, minlength=1 << bits))


@pytest.mark.parametrize("dtype", [np.int64, np.int32])
@pytest.mark.parametrize("bits,shift", [(1, 0), (8, 0), (8, 8), (4, 28), (12, 20), (8, 24),
                                        (8, 28), (12, 26)])
def test_radix_partition_raw_digits_by_python_ints(dtype, bits, shift):
    rng = np.random.default_rng(bits + 100 * shift)
    info = np.iinfo(dtype)
    keys = rng.integers(info.min, info.max, 3000, dtype=dtype, endpoint=True)
    keys[:6] = [info.min, info.max, -1, 0, 1, -2]
    vals = rng.permutation(len(keys)).astype(np.int32)
    width = 64 if dtype == np.int64 else 32
    # raw mode reads the key as an unsigned integer of its own width
    digits = [((int(k) % (1 << width)) >> shift) & ((1 << bits) - 1) for k in keys]
    want_k, want_v, want_c = _brute_partition(keys.tolist(), vals.tolist(), digits, 1 << bits)
    got_k, got_v, got_c = oracle.radix_partition(keys, vals, bits, shift, hashed=False)
    assert got_k.dtype == dtype
    assert got_k.tolist() == want_k and got_v.tolist() == want_v
    assert got_c.tolist() == want_c


# ------------------------------------------------------------ unique hash --

def _first_slots(keys, cap):
    """The reference's slot function (join_core.cu:56-59), as its test
    computes it (test_unique_hash_join.py:80-83)."""
    return [((int(h) << 9) % (1 << 64)) * cap >> 64 for h in oracle.mix64_np(keys)]


def test_unique_hash_colliding_keys_known_answer():
    cand = np.arange(1, 2048, dtype=np.int64)
    slots = _first_slots(cand, 12)
    bucket = next(s for s in range(12) if slots.count(s) >= 6)
    coll = [int(cand[i]) for i, s in enumerate(slots) if s == bucket][:6]
    probe = [coll[5], -999, coll[0], coll[3], coll[0]]
    bi, pi = oracle.unique_hash_pairs([np.array(coll)], [np.array(probe)])
    assert pi.tolist() == [0, 2, 3, 4]
    assert bi.tolist() == [5, 0, 3, 0]


def test_unique_hash_two_key_known_answer():
    bi, pi = oracle.unique_hash_pairs(
        [np.array([1, 1, 2, 2]), np.array([7, 8, 7, 8])],
        [np.array([2, 1, 2, 3, 1, 1]), np.array([8, 8, 9, 7, 7, 7])])
    assert pi.tolist() == [0, 1, 4, 5]
    assert bi.tolist() == [3, 1, 0, 0]


@pytest.mark.parametrize("b,p,eb,ep", [
    ([], [1, 2], [], []), ([1, 2], [], [], []), ([1, 2], [3, 4], [], []),
    ([1, 2, 3], [3, 2, 1], [2, 1, 0], [0, 1, 2])])
def test_unique_hash_edges(b, p, eb, ep):
    bi, pi = oracle.unique_hash_pairs([np.array(b, np.int64)], [np.array(p, np.int64)])
    assert bi.tolist() == eb and pi.tolist() == ep


def test_unique_hash_capacity_boundaries():
    for nb in (7, 8, 9, 15, 16, 17, 31, 32, 33):
        build = np.arange(nb, dtype=np.int64) * 17 + 3
        probe = np.concatenate((build[::-1], [-1], build[:3]))
        bi, pi = oracle.unique_hash_pairs([build], [probe])
        assert bi.tolist() == list(range(nb - 1, -1, -1)) + [0, 1, 2]
        assert pi.tolist() == list(range(nb)) + [nb + 1, nb + 2, nb + 3]


def test_unique_hash_brute_force():
    rng = np.random.default_rng(5)
    build = rng.permutation(4000)[:1500].astype(np.int64) - 2000
    probe = rng.integers(-2500, 2500, 3000)
    bi, pi = oracle.unique_hash_pairs([build], [probe])
    where = {int(k): i for i, k in enumerate(build)}
    want = [(where[int(k)], i) for i, k in enumerate(probe) if int(k) in where]
    assert list(zip(bi.tolist(), pi.tolist())) == want


# ------------------------------------------------------------ semi / anti --

_OPS = {'<>': np.not_equal, '<': np.less, '<=': np.less_equal, '>': np.greater,
        '>=': np.greater_equal}


def _dense_semi_anti(bk, pk, anti, bv=None, pv=None, op=None):
    hit = bk[:, None] == pk[None, :]
    if op is not None:
        hit &= _OPS[op](bv[:, None], pv[None, :])
    keep = hit.any(axis=0) if bk.size else np.zeros(pk.size, bool)
    if anti:
        keep = ~keep
    return np.nonzero(keep)[0]


@pytest.mark.parametrize("anti", [False, True])
@pytest.mark.parametrize("op", [None, '<>', '<', '<=', '>', '>='])
@pytest.mark.parametrize("nb,np_", [(0, 5), (5, 0), (1, 1), (7, 300), (2000, 1500)])
def test_semi_anti_dense(anti, op, nb, np_):
    rng = np.random.default_rng(nb * 7 + np_)
    bk = rng.integers(-50, max(nb // 3, 1), nb)
    pk = rng.integers(-60, 2 * max(nb // 3, 1), np_)
    bv = rng.integers(0, 4, nb)
    pv = rng.integers(0, 4, np_)
    got = oracle.semi_anti(bk, pk, anti, (bv, pv, op) if op else None)
    assert np.array_equal(got, _dense_semi_anti(bk, pk, anti, bv, pv, op))


@pytest.mark.parametrize("anti", [False, True])
def test_semi_anti_summary_known_answers(anti):
    bk = np.array([1, 1, 1, 2, 2, 3])
    pk = np.array([1, 2, 3, 4, 1])
    same = np.full(6, 9)
    got = oracle.semi_anti(bk, pk, anti, (same, np.full(5, 9), '<>'))
    assert got.tolist() == ([0, 1, 2, 3, 4] if anti else [])
    got = oracle.semi_anti(bk, pk, anti, (same, np.full(5, 5), '<>'))
    assert got.tolist() == ([3] if anti else [0, 1, 2, 4])


def test_semi_anti_extreme_keys_and_empty_slot_key():
    keys = [I64.min + 1, -1, 0, -64, I64.max, 42]
    bk = np.array(keys[:4], np.int64)
    pk = np.array(keys + [I64.min], np.int64)
    assert oracle.semi_anti(bk, pk, False).tolist() == [0, 1, 2, 3]
    assert oracle.semi_anti(bk, pk, True).tolist() == [4, 5, 6]
    assert oracle.semi_anti(np.array([3, I64.min]), np.array([3]), False) is None


# ------------------------------------------------------------------ expand --

def _expand_reference(sorted_ids, start, length):
    rows = np.arange(length.size)
    cr = np.repeat(rows, length)
    first = np.repeat(start, length)
    run_start = np.repeat(np.cumsum(length) - length, length)
    within = np.arange(cr.size) - run_start
    return sorted_ids[first + within], cr


@pytest.mark.parametrize("lengths", [[], [0], [0, 0, 0], [1], [3], [1, 0, 2, 0, 0, 3, 0],
                                     [0, 0, 5], [5, 0, 0], [2, 2, 2, 2]])
def test_expand_hand_cases(lengths):
    length = np.array(lengths, np.int64)
    nb = max(8, int(length.sum()) + 3)
    sorted_ids = np.random.default_rng(1).permutation(nb).astype(np.int64)
    start = np.array([(3 * i) % (nb - n) if nb > n else 0 for i, n in enumerate(lengths)],
                     np.int64)
    cs, cr = oracle.expand(sorted_ids, start, length)
    rcs, rcr = _expand_reference(sorted_ids, start, length)
    assert np.array_equal(cs, rcs) and np.array_equal(cr, rcr)


def test_sort_search_pairs_is_the_dense_join():
    rng = np.random.default_rng(9)
    lk = rng.integers(0, 40, 300)
    rk = rng.integers(0, 60, 200)
    got = oracle.join_pairs_sorted(lk, rk)
    li, ri = np.nonzero(lk[:, None] == rk[None, :])
    want = np.stack([li, ri], 1)
    want = want[np.lexsort((want[:, 1], want[:, 0]))]
    assert np.array_equal(got, want)


@pytest.mark.parametrize("buckets", [None, 4])
def test_torch_cpu_sort_search_is_the_dense_join(buckets):
    import torch
    from oracle import torch_cpu
    rng = np.random.default_rng(10)
    lk = rng.integers(0, 40, 3000).astype(np.int32)
    rk = rng.integers(0, 60, 2000).astype(np.int32)
    counts = None
    if buckets:
        # a tick in bucket order: rows sorted by key mod 4, the runs never
        # cross a bucket
        lo, ro = np.argsort(lk % buckets, kind="stable"), np.argsort(rk % buckets, kind="stable")
        lk, rk = lk[lo], rk[ro]
        counts = (np.bincount(lk % buckets, minlength=buckets).tolist(),
                  np.bincount(rk % buckets, minlength=buckets).tolist())
    li, ri = torch_cpu.sort_search_pairs(torch.from_numpy(lk), torch.from_numpy(rk), counts)
    assert li.dtype == ri.dtype == torch.int64
    # right-row order, and each right row's left rows hold its key
    assert bool((ri[1:] >= ri[:-1]).all())
    got = sorted(zip(li.tolist(), ri.tolist()))
    dl, dr = np.nonzero(lk[:, None] == rk[None, :])
    assert got == sorted(zip(dl.tolist(), dr.tolist()))
    nli, nri = oracle.sort_search_pairs(lk.astype(np.int64), rk.astype(np.int64), counts)
    assert got == sorted(zip(nli.tolist(), nri.tolist()))


def _brute_keyed(keys, vals, coeffs):
    """Python dicts: tuple -> (hash, sum, count), the hash from Python ints
    reduced mod 2^64 and read as signed."""
    groups = {}
    for r in range(len(keys[0])):
        t = tuple(int(k[r]) for k in keys)
        h = sum(a * c for a, c in zip(t, coeffs)) % (1 << 64)
        h = h - (1 << 64) if h >= 1 << 63 else h
        g = groups.setdefault(t, [h, 0, 0])
        v = int(vals[r])
        g[1] += v
        g[2] += v != -64
    return sorted(((g[0], t, g[1], g[2]) for t, g in groups.items()))


@pytest.mark.parametrize("nkeys,n,span", [(1, 0, 5), (1, 1, 5), (1, 3000, 40), (1, 5000, 10**6),
                                          (2, 4000, 7), (2, 5000, 300)])
def test_keyed_pass1_is_the_brute_force(nkeys, n, span):
    rng = np.random.default_rng(n + nkeys)
    keys = [rng.integers(-span, span, n) for _ in range(nkeys)]
    if n > 10:
        keys[0][:4] = [I64.min, I64.max, -1, 0]
    vals = rng.integers(-10**12, 10**12, n)
    vals[::7] = -64
    coeffs = [int(c) for c in rng.integers(2, 10_000_000, nkeys)]
    h, ks, sums, cnts, bucket, bcounts = oracle.keyed_pass1(keys, vals, coeffs, 16)
    want = _brute_keyed(keys, vals, coeffs)
    assert h.tolist() == [w[0] for w in want]
    assert [tuple(int(k[i]) for k in ks) for i in range(len(h))] == [w[1] for w in want]
    assert sums.tolist() == [w[2] for w in want] and cnts.tolist() == [w[3] for w in want]
    assert np.array_equal(bucket, oracle.bucket_of(h, 16)) if len(h) else True
    assert sum(bcounts) == len(h)


# ------------------------------------------------------- partition/layout --

def test_partition_is_a_stable_bucket_sort():
    rng = np.random.default_rng(2)
    keys = rng.integers(-(1 << 40), 1 << 40, 5000)
    pay = rng.integers(0, 100, 5000)
    mask = rng.random(5000) < 0.4
    counts, outs = oracle.partition([keys, pay], 0, mask, 77, 16)
    assert sum(counts) == mask.sum()
    gidx = np.concatenate(outs[2])
    kept = np.nonzero(mask)[0]
    b = oracle.bucket_of(keys[kept], 16)
    for bucket, rows in enumerate(outs[2]):
        src = rows - 77
        assert np.all(np.diff(src) > 0)
        assert np.all(oracle.bucket_of(keys[src], 16) == bucket)
    assert sorted(gidx.tolist()) == (kept + 77).tolist()
    assert np.array_equal(np.bincount(b, minlength=16), counts)


def test_padded_layout():
    starts, total = oracle.padded_layout([0, 3, 64, 65, 0], 8)
    assert starts == [0, 0, 64, 128, 256] and total == 256
    starts, total = oracle.padded_layout([1, 1], 25)
    assert starts == [0, 512] and total == 1024


# --------------------------------------------------------------- group-by --

def test_group_sum_count_brute_force():
    rng = np.random.default_rng(4)
    g = rng.integers(0, 25, 10_000).astype(np.int32)
    v = rng.integers(100, 10_000_000, 10_000)
    v[::97] = -64
    s, c = oracle.group_sum_count(g, v, 25)
    for k in range(25):
        assert s[k] == int(sum(int(x) for x, gg in zip(v, g) if gg == k))
        assert c[k] == sum(1 for x, gg in zip(v, g) if gg == k and x != -64)


def test_gather_with_null():
    src = np.array([10, 11, 12], np.int64)
    assert oracle.gather_with_null(src, [2, -64, 0]).tolist() == [12, -64, 10]
    assert oracle.gather_with_null(np.array([], np.int64), [-64, -64]).tolist() == [-64, -64]


def test_geometric_duplicates_reach_the_exact_row_count():
    from pageflow_amd.workloads import geometric_duplicates
    for nd, nb, seed in ((2500, 10_000, 3), (10_000, 10_000, 1), (1000, 9000, 7), (3, 50, 2)):
        c = geometric_duplicates(nd, nb, 0.25, np.random.default_rng(seed))
        assert c.sum() == nb and c.min() >= 1 and c.size == nd
        again = geometric_duplicates(nd, nb, 0.25, np.random.default_rng(seed))
        assert np.array_equal(c, again)


def _cut(x, sizes):
    out, a = [], 0
    for s in sizes:
        out.append(x[a:a + s])
        a += s
    return out


def test_zero_copy_rearrange_matches_the_reference_closed_form():
    # tests/IO/test_scan_regions.py:187-201: indices count from the slice's
    # first row; a null index stays null (want = x[lo:hi][idx], nulls -64)
    rng = np.random.default_rng(5)
    x = rng.integers(-10**12, 10**12, size=(3000, 2))
    parts = _cut(x, [700, 1, 1299, 1000])
    lo, hi = 701, 2307
    idx = rng.integers(0, hi - lo, 500)
    idx[::97] = oracle.NULL
    want = x[lo:hi][np.clip(idx, 0, None)]
    want[idx == oracle.NULL] = oracle.NULL
    assert np.array_equal(oracle.zero_copy_rearrange(parts, 0, idx, lo), want)
    # a row space that starts past 0, and a row outside every segment
    assert np.array_equal(oracle.zero_copy_rearrange(parts, 50, idx, lo + 50), want)
    with pytest.raises(IndexError):
        oracle.zero_copy_rearrange(parts, 0, np.array([3000 - lo]), lo)


def test_zero_copy_mask_matches_the_reference_closed_form():
    # tests/IO/test_scan_regions.py:204-218: the kept rows of a slice, in order
    x = np.arange(3000, dtype=np.int32).reshape(-1, 1) * 3
    parts = _cut(x, [1000, 999, 1001])
    lo, hi = 123, 2911
    mask = (np.arange(3000) % 7 == 3)[lo:hi]
    assert np.array_equal(oracle.zero_copy_mask(parts, 0, mask, lo), x[lo:hi][mask])


def test_partition_outputs_is_the_stable_padded_layout():
    # operators/aggregate.py:120-152: bucket b's rows in source order from a
    # whole lcm(row bytes, 512)-byte unit; checked per bucket by brute force
    rng = np.random.default_rng(3)
    n, B = 5000, 16
    b = rng.integers(0, B, n)
    b[:700] = 5  # a large bucket
    outs = [rng.integers(-10**9, 10**9, n), rng.integers(0, 255, (n, 3)).astype(np.uint8),
            rng.standard_normal(n).astype(np.float32)]
    counts, padded = oracle.partition_outputs(outs, b, B)
    assert counts == np.bincount(b, minlength=B).tolist()
    for t, p in zip(outs, padded):
        rb = t.dtype.itemsize * int(np.prod(t.shape[1:]))
        starts, total = oracle.padded_layout(counts, rb)
        assert p.shape[0] == total
        for k in range(B):
            want = t[[i for i in range(n) if b[i] == k]]
            assert np.array_equal(p[starts[k]:starts[k] + counts[k]], want), (rb, k)


# ------------------------------------------------- residual conditions --

def _raises(fn):
    try:
        return False, fn()
    except (TypeError, RuntimeError, OverflowError, ValueError, NotImplementedError,
            AttributeError) as e:
        return True, e


@pytest.mark.parametrize("seed", range(6))
def test_residual_mask_is_the_torch_evaluation(seed):
    """The oracle's comparison rules (numpy with torch's scalar casts,
    promotions and isclose restated) give torch's own answers on random
    condition trees over every column dtype and literal kind; where torch
    raises, the oracle raises."""
    from residual_cases import condition, sides, torch_mask
    rng = np.random.default_rng(seed)
    left, right = sides(rng, 300, 200)
    n = 2000
    lm = rng.integers(0, 300, n)
    rm = rng.integers(0, 200, n)
    agreed = raised = 0
    for _ in range(150):
        tree = condition(rng, allow_unsupported=True)
        t_err, want = _raises(lambda: torch_mask(left, right, tree, lm, rm))
        o_err, got = _raises(lambda: oracle.residual_mask(left, right, tree, lm, rm))
        assert t_err == o_err, (tree, want, got)
        if t_err:
            raised += 1
            continue
        assert got.dtype == bool and got.shape == (n,)
        assert np.array_equal(got, want), tree
        agreed += 1
    assert agreed > 80 and raised > 3


def test_isclose_is_torch_isclose_at_the_tolerance_boundary():
    import torch
    rng = np.random.default_rng(3)
    b = np.concatenate([rng.standard_normal(2000) * 10.0 ** rng.integers(-8, 8, 2000),
                        [0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, np.nan, 1e308, 5e-324]])
    with np.errstate(invalid="ignore", over="ignore"):
        tol = 1e-5 + np.abs(1e-5 * b)
        cols = [b, b + tol, b - tol, np.nextafter(b + tol, np.inf),
                np.nextafter(b - tol, -np.inf), b + 0.5 * tol, b * (1 + 2e-5),
                np.full_like(b, np.inf), np.full_like(b, np.nan)]
    for a in cols:
        with np.errstate(invalid="ignore", over="ignore"):
            got = oracle.isclose_f64(a, b)
        want = torch.isclose(torch.from_numpy(a), torch.from_numpy(b), atol=1e-5).numpy()
        assert np.array_equal(got, want)
    # closed forms: the tolerance is 1e-5 + 1e-5 |b|, infinities equal themselves
    assert oracle.isclose_f64(0.0, 1e-5) and not oracle.isclose_f64(0.0, 1.1e-5)
    assert oracle.isclose_f64(1.0 + 2e-5, 1.0) and not oracle.isclose_f64(1.0 + 2.1e-5, 1.0)
    assert oracle.isclose_f64(np.inf, np.inf) and not oracle.isclose_f64(np.inf, -np.inf)
    assert not oracle.isclose_f64(np.nan, np.nan)


def test_residual_filter_known_answers():
    # tests/operators/test_unique_hash_join.py:251-272: #5 > #6 on the pairs
    # (0,0), (1,1), (2,2) keeps (1,1)
    left = {1: np.array([1, 2, 3]), 5: np.array([10, 20, 30])}
    right = {2: np.array([1, 2, 3]), 6: np.array([15, 15, 35])}
    lo, ro = oracle.residual_filter(left, right, [["#5", ">", "#6"]], [0, 1, 2], [0, 1, 2])
    assert lo.tolist() == [1] and ro.tolist() == [1]
    # Q21's residual NOT (l_suppkey#3 = l_suppkey#4), every pair of a 4 x 3
    left = {3: np.array([7, 8, 7, 9])}
    right = {4: np.array([7, 9, 1])}
    lm, rm = np.repeat(np.arange(4), 3), np.tile(np.arange(3), 4)
    q21 = ['NOT', ['l_suppkey#3', '=', 'l_suppkey#4']]
    lo, ro = oracle.residual_filter(left, right, [q21], lm, rm)
    want = [(i, j) for i in range(4) for j in range(3) if left[3][i] != right[4][j]]
    assert list(zip(lo.tolist(), ro.tolist())) == want
    # several conditions AND, pair order kept; an id on both sides reads left
    both = {5: np.array([1, 2, 3])}
    lo, ro = oracle.residual_filter(both, {5: np.array([9, 9, 9]), 6: np.array([0, 5, 1])},
                                    [['#5', '>', '1'], ['#6', '<', '#5']],
                                    [2, 1, 0, 2], [2, 2, 2, 0])
    assert list(zip(lo.tolist(), ro.tolist())) == [(2, 2), (1, 2), (2, 0)]


def test_residual_literal_rules():
    """Closed forms of the literal rules: dates are days since 1990-01-01,
    an integer literal wraps to the column dtype, a float literal under <=
    / >= is shifted by 1e-5, and a one-character string is its code
    point."""
    i32 = {1: np.array([1898, 1899, 1900], dtype=np.int32)}
    r = {2: np.zeros(3, np.int64)}
    idx = np.arange(3)
    assert oracle.residual_mask(i32, r, ['#1', '=', '1995-03-15'], idx, idx).tolist() == \
        [False, True, False]
    assert oracle.residual_mask(i32, r, ['#1', '>', str(2 ** 32 + 1899)], idx, idx).tolist() == \
        [False, False, True]
    f = {1: np.array([1.0, 1.0 + 0.9e-5, 1.0 + 1.1e-5])}
    assert oracle.residual_mask(f, r, ['#1', '<=', '1.0'], idx, idx).tolist() == \
        [True, True, False]
    assert oracle.residual_mask(f, r, ['#1', '<', '1.0'], idx, idx).tolist() == \
        [False, False, False]
    s = {1: np.array([64, 65, 66], dtype=np.int8)}
    assert oracle.residual_mask(s, r, ['#1', '=', 'A'], idx, idx).tolist() == \
        [False, True, False]
    u8 = {1: np.array([0, 100, 255], dtype=np.uint8)}
    assert oracle.residual_mask(u8, r, ['#1', '<', '-1'], idx, idx).tolist() == \
        [True, True, False]


def test_residual_constants_mirror_the_header():
    import re
    from pageflow_amd.operators import residual
    text = open(os.path.join(os.path.dirname(__file__), "..", "include",
                             "pageflow_join.h")).read()
    d = {k: v for k, v in re.findall(r"#define (PF_(?:RES|REL|COL)_\w+) \(?(\d+)", text)}
    assert int(d["PF_RES_CMP"]) == residual.CMP and int(d["PF_RES_AND"]) == residual.AND
    assert int(d["PF_RES_OR"]) == residual.OR and int(d["PF_RES_NOT"]) == residual.NOT
    assert int(d["PF_RES_TRUE"]) == residual.TRUE
    for name, rel in {"EQ": '=', "NE": '<>', "LT": '<', "LE": '<=', "GT": '>',
                      "GE": '>='}.items():
        assert int(d[f"PF_REL_{name}"]) == residual.REL[rel]
    assert int(d["PF_REL_CLOSE"]) == residual.REL_CLOSE
    assert re.search(r"PF_RES_F64 \(1LL << 16\)", text) and residual.F64 == 1 << 16
    assert re.search(r"PF_RES_A_COL \(1LL << 17\)", text) and residual.A_COL == 1 << 17
    assert re.search(r"PF_RES_B_COL \(1LL << 18\)", text) and residual.B_COL == 1 << 18
    import torch
    for name, dt in {"I64": torch.int64, "I32": torch.int32, "I16": torch.int16,
                     "I8": torch.int8, "U8": torch.uint8, "F64": torch.float64}.items():
        assert int(d[f"PF_COL_{name}"]) == residual.COL_TYPE[dt]
    assert int(d["PF_RES_MAX_COLS"]) == residual.MAX_COLS
    assert int(d["PF_RES_MAX_INS"]) == residual.MAX_INS


def test_fmix32_is_murmur3s_finalizer_and_a_bijection():
    """fmix32 against its steps in Python integers, its inverse (the
    multipliers' inverses mod 2^32), and 0 -> 0."""
    rng = np.random.default_rng(32)
    x = rng.integers(-(1 << 31), 1 << 31, 5000, dtype=np.int64)
    h = oracle.fmix32_np(x)
    M = (1 << 32) - 1

    def steps(v):
        v &= M
        v ^= v >> 16
        v = (v * 0x85ebca6b) & M
        v ^= v >> 13
        v = (v * 0xc2b2ae35) & M
        return v ^ (v >> 16)

    def inverse(v):
        v ^= v >> 16
        v = (v * pow(0xc2b2ae35, -1, 1 << 32)) & M
        v ^= (v >> 13) ^ (v >> 26)
        v = (v * pow(0x85ebca6b, -1, 1 << 32)) & M
        return v ^ (v >> 16)

    assert [int(v) for v in h] == [steps(int(v)) for v in x]
    assert [inverse(int(v)) for v in h] == [int(v) & M for v in x]
    assert int(oracle.fmix32_np(np.array([0]))[0]) == 0
    assert np.unique(oracle.fmix32_np(np.arange(-50000, 50000))).size == 100000
