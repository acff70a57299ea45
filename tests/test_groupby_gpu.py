"""Exact int64 group-by, the resident dimension table and the fused
fact->dimension join + group-by, and the payload gather, against the
oracle."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from pageflow_amd import ops

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("G,n", [(1, 1000), (25, 2_000_000), (1024, 300_000), (5000, 200_000)])
def test_group_sum_count(G, n):
    rng = np.random.default_rng(G)
    g = rng.integers(0, G, n).astype(np.int32)
    v = rng.integers(100, 10_000_000, n)
    v[::101] = -64
    s, c = ops.group_sum_count(torch.from_numpy(g).cuda(), torch.from_numpy(v).cuda(), G)
    ws, wc = oracle.group_sum_count(g, v, G)
    assert np.array_equal(s.cpu().numpy(), ws) and np.array_equal(c.cpu().numpy(), wc)


@pytest.mark.parametrize("nd,nf", [(1, 10), (1000, 100_000), (2_000_000, 5_000_000)])
def test_fact_join_groupby_matches_join_then_group(nd, nf):
    rng = np.random.default_rng(nd)
    dkey = rng.permutation(4 * nd)[:nd].astype(np.int32) - 7
    dgrp = rng.integers(0, 25, nd).astype(np.int32)
    fk = dkey[rng.integers(0, nd, nf)]
    fk[rng.random(nf) < 0.2] = 4 * nd + 100  # absent keys
    amt = rng.integers(100, 10_000_001, nf)
    table = ops.DimTable(torch.from_numpy(dkey).cuda(), torch.from_numpy(dgrp).cuda(), 25)
    table.check()
    s, c = table.join_groupby(torch.from_numpy(fk).cuda(), torch.from_numpy(amt).cuda())
    b, p = oracle.sort_search_pairs(dkey.astype(np.int64), fk.astype(np.int64))
    ws, wc = oracle.group_sum_count(dgrp[b], amt[p], 25)
    assert np.array_equal(s.cpu().numpy(), ws) and np.array_equal(c.cpu().numpy(), wc)


def test_dim_table_reports_bad_groups_and_repeated_keys():
    t = ops.DimTable(torch.tensor([1, 2, 2], dtype=torch.int32, device="cuda"),
                     torch.tensor([0, 30, 1], dtype=torch.int32, device="cuda"), 25)
    with pytest.raises(ValueError):
        t.check()


def test_gather_rows_and_null_rows():
    src = torch.arange(30, dtype=torch.int64, device="cuda").view(10, 3) * 7
    idx = torch.tensor([3, -64, 0, 9], dtype=torch.int64, device="cuda")
    got = ops.gather_rows(src, idx).cpu().numpy()
    want = np.stack([oracle.gather_with_null(src[:, q].cpu().numpy(), idx.cpu().numpy())
                     for q in range(3)], 1)
    assert np.array_equal(got, want)
    assert ops.gather_rows(torch.empty(0, dtype=torch.int32, device="cuda"), idx).tolist() == \
        [-64] * 4


def test_matched_rows():
    idx = torch.tensor([5, 1, 5, 7], dtype=torch.int64, device="cuda")
    assert ops.matched_rows(idx, 9, True).tolist() == [1, 5, 7]
    assert ops.matched_rows(idx, 9, False).tolist() == [0, 2, 3, 4, 6, 8]


@pytest.mark.parametrize("layout", ["dense", "sparse", "dense_keys_many_groups"])
def test_dimension_table_layouts(layout):
    rng = np.random.default_rng(7)
    nd, nf = 30_000, 400_000
    if layout == "sparse":  # keys spread far wider than 4n values: hashed
        dkey = rng.choice(2**30, nd, replace=False).astype(np.int32) - 2**29
    else:  # a permutation of a dense range: a byte array unless G > 255
        dkey = (rng.permutation(nd) - 1000).astype(np.int32)
    G = 300 if layout == "dense_keys_many_groups" else 25
    dgrp = rng.integers(0, G, nd).astype(np.int32)
    fk = dkey[rng.integers(0, nd, nf)]
    fk[rng.random(nf) < 0.2] = np.int32(2**31 - 5)  # absent, outside any key range
    fk[::97] = np.int32(-2**31)
    amt = rng.integers(100, 10_000_001, nf)
    amt[::13] = -64
    t = ops.DimTable(torch.from_numpy(dkey).cuda(), torch.from_numpy(dgrp).cuda(), G)
    t.check()
    assert t.dense == (layout == "dense")
    s, c = t.join_groupby(torch.from_numpy(fk).cuda(), torch.from_numpy(amt).cuda())
    b, p = oracle.sort_search_pairs(dkey.astype(np.int64), fk.astype(np.int64))
    ws, wc = oracle.group_sum_count(dgrp[b], amt[p], G)
    assert np.array_equal(s.cpu().numpy(), ws) and np.array_equal(c.cpu().numpy(), wc)


@pytest.mark.parametrize("keys", [[1, 2, 2, 3], [1, 2**30, 2**30, -2**30]])
def test_repeated_keys_in_either_layout(keys):
    t = ops.DimTable(torch.tensor(keys, dtype=torch.int32, device="cuda"),
                     torch.tensor([0, 1, 1, 2], dtype=torch.int32, device="cuda"), 25)
    assert int(t.info[1]) == 1 and int(t.info[0]) == 0
    assert t.dense == (keys[1] == 2)
    with pytest.raises(ValueError):
        t.check()
