"""The partitioned join's feed (``operators.tick_feed.streamed_join``): B
from the partition budget, each side partitioned in several sort ticks into
the padded layout, whole bucket pairs packed into core calls, each call a
``join_kernel_new`` with its bucket counts and gidx columns.  The union of
the calls' (left gidx, right gidx) rows must be the whole join's (a numpy
brute force), for every join type, through the unique-hash route and the
pair route, and the schedule must be the packer's."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from pageflow_amd.operators import tick_feed
from pageflow_amd.operators.partitions import pack_core_calls

pytestmark = pytest.mark.gpu

NULL = -64


def _sides(seed, n_left=300_000, n_right=1_200_000, unique_left=True):
    rng = np.random.default_rng(seed)
    lk = (rng.permutation(n_left) * 3 + 7).astype(np.int64) if unique_left else \
        rng.integers(0, n_left // 3, n_left).astype(np.int64)
    rk = lk[rng.integers(0, n_left, n_right)]
    rk[rng.random(n_right) < 0.2] = -5                    # absent
    lp = rng.integers(0, 1 << 40, n_left).astype(np.int64)
    return lk, lp, rk


def _expected(join_type, lk, rk):
    pairs = oracle.join_pairs_sorted(lk, rk)              # (left row, right row)
    matched = np.zeros(rk.size, bool)
    matched[pairs[:, 1]] = True
    if join_type == 'inner':
        return sorted(map(tuple, pairs.tolist()))
    if join_type == 'right-semi':
        return np.nonzero(matched)[0].tolist()
    if join_type == 'right-anti':
        return np.nonzero(~matched)[0].tolist()
    out = list(map(tuple, pairs.tolist())) + [(NULL, int(r)) for r in np.nonzero(~matched)[0]]
    return sorted(out)


@pytest.mark.parametrize("join_type", ['inner', 'right-semi', 'right-anti', 'right-outer'])
@pytest.mark.parametrize("proven", [True, False])
def test_streamed_join_equals_the_whole_join(join_type, proven):
    lk, lp, rk = _sides(5, unique_left=True)
    left = {1: torch.from_numpy(lk).cuda(), 3: torch.from_numpy(lp).cuda()}
    right = {2: torch.from_numpy(rk).cuda()}
    out, (B, calls) = tick_feed.streamed_join(
        left, right, 1, 2, [], join_type, [3], unique_key_sets=([[1]] if proven else [], []),
        partition_budget_b=2 << 20, core_budget_b=6 << 20, tick_rows=250_000)
    assert B > 1 and len(calls) > 1
    assert calls[0][0] == 0 and calls[-1][1] == B
    assert all(a[1] == b[0] for a, b in zip(calls, calls[1:]))
    lg, rg = out[-2].cpu().numpy(), out[-1].cpu().numpy()
    if join_type in ('right-semi', 'right-anti'):
        assert lg.size == 0
        assert sorted(rg.tolist()) == _expected(join_type, lk, rk)
        return
    assert sorted(zip(lg.tolist(), rg.tolist())) == _expected(join_type, lk, rk)
    # the payload column gathered at the left rows (null rows for -64)
    pay = out[0].cpu().numpy()
    want = np.where(lg == NULL, NULL, lp[np.where(lg == NULL, 0, lg)])
    assert np.array_equal(pay, want)


def test_multi_match_build_and_the_schedule():
    lk, lp, rk = _sides(6, n_left=200_000, n_right=500_000, unique_left=False)
    left = {1: torch.from_numpy(lk).cuda(), 3: torch.from_numpy(lp).cuda()}
    right = {2: torch.from_numpy(rk).cuda()}
    budget = 4 << 20
    out, (B, calls) = tick_feed.streamed_join(left, right, 1, 2, [], 'inner', [],
                                              partition_budget_b=1 << 20,
                                              core_budget_b=budget, tick_rows=70_000)
    assert sorted(zip(out[-2].tolist(), out[-1].tolist())) == _expected('inner', lk, rk)
    # the calls are the packer's over the buckets' rows
    import oracle as o
    lb = np.bincount(o.bucket_of(lk, B), minlength=B)
    rb = np.bincount(o.bucket_of(rk, B), minlength=B)
    ends, _ = pack_core_calls(lb.tolist(), rb.tolist(), (8 + 8 + 8, 8 + 8), budget // 2)
    assert [e for _, e in calls] == ends


def test_b_follows_the_partition_budget():
    rows = [(10_000_000, 24), (2_000_000, 16)]
    b_small, _, _ = tick_feed.join_partition_count(rows, [False, True], 64 << 20)
    b_large, _, _ = tick_feed.join_partition_count(rows, [False, True], 1 << 30)
    assert b_small > b_large >= 1
