"""The join core kernels (unique hash, semi/anti, expand) against the oracle
and the known answers of the reference's tests
(tests/operators/test_join_core.py, tests/operators/test_unique_hash_join.py
of the reference repository).  Unique-hash and semi/anti results are ordered
(probe rows ascending), so they are compared exactly."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
import pageflow_amd.operators.join_core as join_core
from pageflow_amd.operators.join_core import unique_hash_pairs

pytestmark = pytest.mark.gpu

I64_MIN = -(1 << 63)
I64_MAX = (1 << 63) - 1


def _dev(values):
    return torch.as_tensor(np.asarray(values, dtype=np.int64)).cuda()


def _ints(n, lo, hi, seed):
    """n seeded int64 values in [lo, hi) on the GPU."""
    return _dev(np.random.default_rng(seed).integers(lo, hi, n, dtype=np.int64))


def _np(t):
    return t.cpu().numpy()


# ------------------------------------------------------------------ expand --

@pytest.mark.parametrize("lengths", [[], [0], [0, 0, 0], [1], [3], [1, 0, 2, 0, 0, 3, 0],
                                     [0, 0, 5], [5, 0, 0], [2, 2, 2, 2]])
def test_expand_hand_cases(lengths):
    length = _dev(lengths)
    nb = max(8, sum(lengths) + 3)
    sorted_ids = torch.randperm(nb, device="cuda")
    start = _dev([(3 * i) % (nb - n) if nb > n else 0 for i, n in enumerate(lengths)])
    cs, cr = join_core.expand(sorted_ids, start, length)
    rcs, rcr = oracle.expand(_np(sorted_ids), _np(start), _np(length))
    assert np.array_equal(_np(cs), rcs) and np.array_equal(_np(cr), rcr)


@pytest.mark.parametrize("n_probe,n_build", [(60_000, 200_000), (3_000_000, 1_000_000)])
def test_expand_random_runs(n_probe, n_build):
    length = _ints(n_probe, 0, 7, 1)
    length[_ints(n_probe, 0, 10, 2) < 4] = 0
    start = torch.minimum(_ints(n_probe, 0, n_build, 3), n_build - length)
    sorted_ids = torch.randperm(n_build, device="cuda")
    cs, cr = join_core.expand(sorted_ids, start, length)
    rcs, rcr = oracle.expand(_np(sorted_ids), _np(start), _np(length))
    assert np.array_equal(_np(cs), rcs) and np.array_equal(_np(cr), rcr)


def test_expand_one_long_run():
    # one probe row with a 100K-pair run between empty ones (skew)
    length = _dev([0, 3, 100_000, 0, 1])
    start = _dev([0, 5, 0, 0, 7])
    sorted_ids = torch.randperm(100_005, device="cuda")
    cs, cr = join_core.expand(sorted_ids, start, length)
    rcs, rcr = oracle.expand(_np(sorted_ids), _np(start), _np(length))
    assert np.array_equal(_np(cs), rcs) and np.array_equal(_np(cr), rcr)


# ------------------------------------------------------------- semi / anti --

@pytest.mark.parametrize("anti", [False, True])
@pytest.mark.parametrize("op", [None, '<>', '<', '<=', '>', '>='])
@pytest.mark.parametrize("n_build,n_probe", [(0, 5), (5, 0), (1, 1), (7, 300), (2000, 1500),
                                             (40_000, 3000)])
def test_semi_anti_matches_oracle(anti, op, n_build, n_probe):
    seed = n_build * 7 + n_probe
    bk = _ints(n_build, -50, max(n_build // 3, 1), seed)
    pk = _ints(n_probe, -60, 2 * max(n_build // 3, 1), seed + 1)
    bv = _ints(n_build, 0, 4, seed + 2)
    pv = _ints(n_probe, 0, 4, seed + 3)
    summary = (bv, pv, op) if op is not None else None
    got = join_core.semi_anti(bk, pk, anti, summary)
    want = oracle.semi_anti(_np(bk), _np(pk), anti,
                            (_np(bv), _np(pv), op) if op else None)
    assert np.array_equal(_np(got), want)
    mask = join_core.semi_anti_mask(bk, pk, anti, summary)
    assert np.array_equal(np.nonzero(_np(mask))[0], want)


@pytest.mark.parametrize("anti", [False, True])
def test_summary_all_true_and_all_false(anti):
    bk = _dev([1, 1, 1, 2, 2, 3])
    pk = _dev([1, 2, 3, 4, 1])
    same = _dev([9] * 6)
    got = join_core.semi_anti(bk, pk, anti, (same, _dev([9] * 5), '<>'))
    assert got.tolist() == ([0, 1, 2, 3, 4] if anti else [])
    got = join_core.semi_anti(bk, pk, anti, (same, _dev([5] * 5), '<>'))
    assert got.tolist() == ([3] if anti else [0, 1, 2, 4])


def test_semi_anti_extreme_keys_and_null_sentinel():
    keys = [I64_MIN + 1, -1, 0, -64, I64_MAX, 42]
    bk = _dev(keys[:4])
    pk = _dev(keys + [I64_MIN])
    assert join_core.semi_anti(bk, pk, False).tolist() == [0, 1, 2, 3]
    assert join_core.semi_anti(bk, pk, True).tolist() == [4, 5, 6]


def test_semi_anti_refuses_the_empty_slot_key():
    assert join_core.semi_anti(_dev([3, I64_MIN]), _dev([3]), False) is None
    assert join_core.semi_anti_mask(_dev([3, I64_MIN]), _dev([3]), False) is None


def _one_bucket_keys(n, B=128, bucket=5, seed=0):
    cand = torch.randperm(n * B * 2, generator=torch.Generator().manual_seed(seed))
    cand = cand[torch.from_numpy(oracle.bucket_of(cand.numpy(), B)) == bucket][:n]
    assert cand.numel() == n
    return cand.cuda()


def test_tables_with_keys_of_one_radix_bucket():
    keys = _one_bucket_keys(20_000)
    probe = torch.cat((keys[::3], keys[1::7] + (1 << 40)))
    got = join_core.semi_anti(keys, probe, False)
    assert np.array_equal(_np(got), oracle.semi_anti(_np(keys), _np(probe), False))
    build_idx, probe_idx = unique_hash_pairs([keys], [probe])
    assert torch.equal(probe[probe_idx], keys[build_idx])
    assert probe_idx.tolist() == list(range(keys[::3].numel()))


# -------------------------------------------------------------- unique hash --

@pytest.mark.parametrize("two", [False, True])
def test_unique_hash_random_tuples(two):
    n_build, n_probe = 30_000, 50_000
    k0 = torch.randperm(n_build * 4, device="cuda")[:n_build]
    build = [k0 // 3, k0 % 3] if two else [k0]
    picks = _ints(n_probe, 0, n_build, 11)
    miss = _ints(n_probe, 0, 4, 12) == 0
    p0 = torch.where(miss, k0[picks] + n_build * 8, k0[picks])
    probe = [p0 // 3, p0 % 3] if two else [p0]
    build_idx, probe_idx = unique_hash_pairs(build, probe)
    assert torch.equal(probe_idx, torch.nonzero(~miss).view(-1))
    assert torch.equal(build_idx, picks[probe_idx])
    bi, pi = oracle.unique_hash_pairs([_np(b) for b in build], [_np(p) for p in probe])
    assert np.array_equal(_np(build_idx), bi) and np.array_equal(_np(probe_idx), pi)


def _colliding_keys(cap, n):
    """n distinct keys sharing their first slot in a table of cap slots
    under this library's slot function (low 32 bits of mix64,
    multiply-shift)."""
    cand = np.arange(1, 1 << 16, dtype=np.int64)
    h = oracle.mix64_np(cand) & np.uint64(0xffffffff)
    slots = (h * np.uint64(cap)) >> np.uint64(32)
    for s in range(cap):
        hit = cand[slots == s]
        if hit.size >= n:
            return hit[:n].tolist()
    raise AssertionError("no slot with enough colliders")


def test_single_key_pairs_are_stable_and_collision_safe():
    coll = _colliding_keys(12, 6)
    build = _dev(coll)
    probe = _dev([coll[5], -999, coll[0], coll[3], coll[0]])
    build_idx, probe_idx = unique_hash_pairs([build], [probe])
    assert probe_idx.tolist() == [0, 2, 3, 4]
    assert build_idx.tolist() == [5, 0, 3, 0]
    b2, p2 = unique_hash_pairs([build], [probe])
    assert torch.equal(build_idx, b2) and torch.equal(probe_idx, p2)


def test_two_column_tuple_comparison_and_repeated_probe_keys():
    build_idx, probe_idx = unique_hash_pairs(
        [_dev([1, 1, 2, 2]), _dev([7, 8, 7, 8])],
        [_dev([2, 1, 2, 3, 1, 1]), _dev([8, 8, 9, 7, 7, 7])])
    assert probe_idx.tolist() == [0, 1, 4, 5]
    assert build_idx.tolist() == [3, 1, 0, 0]


@pytest.mark.parametrize("b,p,eb,ep", [
    ([], [1, 2], [], []), ([1, 2], [], [], []), ([1, 2], [3, 4], [], []),
    ([1, 2, 3], [3, 2, 1], [2, 1, 0], [0, 1, 2])])
def test_empty_no_match_and_all_match_edges(b, p, eb, ep):
    build_idx, probe_idx = unique_hash_pairs([_dev(b)], [_dev(p)])
    assert build_idx.tolist() == eb and probe_idx.tolist() == ep


def test_capacity_boundaries_match_a_reference_lookup():
    for nb in (7, 8, 9, 15, 16, 17, 31, 32, 33):
        build = torch.arange(nb, dtype=torch.int64, device="cuda") * 17 + 3
        probe = torch.cat((build.flip(0), _dev([-1]), build[:3]))
        build_idx, probe_idx = unique_hash_pairs([build], [probe])
        assert build_idx.tolist() == list(range(nb - 1, -1, -1)) + [0, 1, 2]
        assert probe_idx.tolist() == list(range(nb)) + [nb + 1, nb + 2, nb + 3]


def test_extreme_keys_and_the_empty_slot_value_in_a_unique_build():
    build = _dev([I64_MIN, 5, I64_MAX, -64, 0])
    probe = _dev([0, I64_MIN, 7, I64_MAX, I64_MIN, -64])
    build_idx, probe_idx = unique_hash_pairs([build], [probe])
    assert probe_idx.tolist() == [0, 1, 3, 4, 5]
    assert build_idx.tolist() == [4, 0, 2, 0, 3]


def test_a_repeated_build_key_takes_the_lowest_row():
    # outside the uniqueness proof: the documented "first build row"
    build = _dev([4, 9, 4, 4, 9])
    probe = _dev([9, 4, 1])
    build_idx, probe_idx = unique_hash_pairs([build], [probe])
    assert probe_idx.tolist() == [0, 1] and build_idx.tolist() == [1, 0]


def test_large_unique_hash_matches_oracle():
    nb, np_ = 1_000_000, 6_000_000
    g = torch.Generator().manual_seed(1)
    bk = torch.randperm(nb, generator=g) + 1
    pk = bk[torch.randint(0, nb, (np_,), generator=g)]
    bi, pi = unique_hash_pairs([bk.cuda()], [pk.cuda()])
    rbi, rpi = oracle.unique_hash_pairs([bk.numpy()], [pk.numpy()])
    assert np.array_equal(_np(bi), rbi) and np.array_equal(_np(pi), rpi)
