"""The keyed group-by of a keyed aggregation's pass 1 (pf_group_by_keys,
``operators.aggregate.keyed_pass1``) against ``oracle.keyed_pass1``: every
group's hash, key tuple, exact int64 SUM and COUNT, the group order
(ascending hash) and the buckets, bit for bit, for one and two int64 keys,
from one group to millions, with int64 extremes, null values and a
duplicate-free check of the hash."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from pageflow_amd.operators import aggregate as agg

pytestmark = pytest.mark.gpu


def _case(nkeys, n, span, seed, extremes=False):
    rng = np.random.default_rng(seed)
    keys = [rng.integers(-span, span, n, dtype=np.int64) for _ in range(nkeys)]
    if extremes and n >= 8:
        keys[0][:6] = [np.iinfo(np.int64).min, np.iinfo(np.int64).max, -1, 0,
                       np.iinfo(np.int64).min, 1]
    vals = rng.integers(-10**13, 10**13, n, dtype=np.int64)
    vals[rng.random(n) < 0.05] = -64
    # odd coefficients: a product with an odd number is a bijection of the
    # int64 ring, so one key's hash never merges two keys (an even one maps
    # 0 and int64 min together)
    coeffs = [int(c) | 1 for c in rng.integers(2, 10_000_000, nkeys)]
    return keys, vals, coeffs


def _check(keys, vals, coeffs, B):
    cols = {10 + i: torch.from_numpy(k).cuda() for i, k in enumerate(keys)}
    cols[99] = torch.from_numpy(vals).cuda()
    outs, buckets = agg.keyed_pass1(cols, [10 + i for i in range(len(keys))], 99, B, coeffs)
    h, ks, sums, cnts, wb, wcounts = oracle.keyed_pass1(keys, vals, coeffs, B)
    assert np.array_equal(outs[0].cpu().numpy(), h)
    for i in range(len(keys)):
        assert np.array_equal(outs[1 + i].cpu().numpy(), ks[i])
    assert np.array_equal(outs[-2].cpu().numpy(), sums)
    assert np.array_equal(outs[-1].cpu().numpy(), cnts)
    if B == 1 or h.size == 0:
        assert buckets is None
    else:
        bucket, counts, counts_t = buckets
        assert bucket.dtype == torch.int32
        assert np.array_equal(bucket.cpu().numpy(), wb)
        assert counts == wcounts and counts_t.tolist() == wcounts


@pytest.mark.parametrize("nkeys,n,span,B", [
    (1, 1, 5, 1), (1, 1000, 3, 16), (1, 100_000, 50_000, 64), (1, 2_000_000, 1_000_000, 512),
    (1, 10_000_000, 13, 1), (2, 300_000, 100, 8), (2, 3_000_000, 1200, 256),
    (2, 50_000, 1 << 62, 4096)])
def test_keyed_pass1_matches_the_oracle(nkeys, n, span, B):
    keys, vals, coeffs = _case(nkeys, n, span, seed=n + nkeys, extremes=True)
    _check(keys, vals, coeffs, B)


def test_millions_of_groups():
    # 4M rows, every key distinct: 4M groups, one row each
    rng = np.random.default_rng(3)
    keys = [rng.permutation(4_000_000).astype(np.int64) * 977 - 123]
    vals = rng.integers(0, 10**9, 4_000_000, dtype=np.int64)
    _check(keys, vals, [3], 64)


def test_empty_chunk_and_the_reference_coefficients():
    e = torch.empty(0, dtype=torch.int64, device="cuda")
    outs, buckets = agg.keyed_pass1({1: e, 2: e}, [1], 2, 64, [5])
    assert [t.numel() for t in outs] == [0, 0, 0, 0] and buckets is None
    k = torch.arange(10, device="cuda")
    c1 = agg._make_key_coeffs([k, k.view(-1, 1)], k.device)
    c2 = agg._make_key_coeffs([k, k.view(-1, 1)], k.device)
    assert [c.tolist() for c in c1] == [c.tolist() for c in c2]
    assert all(c.numel() == 1 and 2 <= int(c) < 10_000_000 for c in c1)
    outs, _ = agg.keyed_pass1({1: k, 2: k}, [1], 2, 1)
    assert sorted(outs[1].tolist()) == list(range(10)) and outs[-1].tolist() == [1] * 10


def test_a_hash_collision_is_reported():
    # coefficient 2: keys 0 and 2^63 share the hash 0
    k = torch.tensor([0, -(1 << 63), 5], dtype=torch.int64, device="cuda")
    with pytest.raises(RuntimeError, match="share a grouping hash"):
        agg.group_by_keys([k], k, [2], 1)
