"""The extension-level module (pageflow_amd.tqpkernels, the four functions
of the reference's TQPlib.tqpkernels) called the way the reference's
operators call it: positional arguments, one shared pinned host_out,
outputs at their upper bounds narrowed by host_out."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from pageflow_amd import tqpkernels

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def host():
    return torch.empty(512 + 5, dtype=torch.int64, pin_memory=True)


def test_join_unique_hash_and_semi_anti(host):
    rng = np.random.default_rng(5)
    bk = rng.permutation(50_000).astype(np.int64) * 3 - 7
    pk = np.concatenate([bk[rng.integers(0, bk.size, 200_000)], rng.integers(10**9, 10**10, 5000)])
    rng.shuffle(pk)
    b, p = torch.from_numpy(bk).cuda(), torch.from_numpy(pk).cuda()
    bi, pi = tqpkernels.join_unique_hash([b], [p], host)
    n = int(host[0])
    assert bi.numel() == pk.size
    wb, wp = oracle.unique_hash_pairs([bk], [pk])
    assert np.array_equal(bi[:n].cpu().numpy(), wb) and np.array_equal(pi[:n].cpu().numpy(), wp)
    for anti in (False, True):
        out = tqpkernels.join_semi_anti(b, p, None, None, 0, anti, host)
        n, bad = host[:2].tolist()
        assert not bad
        assert np.array_equal(out[:n].cpu().numpy(), oracle.semi_anti(bk, pk, anti))
    bv = torch.from_numpy(rng.integers(0, 100, bk.size)).cuda()
    pv = torch.from_numpy(rng.integers(0, 100, pk.size)).cuda()
    out = tqpkernels.join_semi_anti(b, p, bv, pv, 3, False, host)  # build.v <= probe.v
    n = int(host[0])
    want = oracle.semi_anti(bk, pk, False, (bv.cpu().numpy(), pv.cpu().numpy(), "<="))
    assert np.array_equal(out[:n].cpu().numpy(), want)


def test_join_expand(host):
    sorted_ids = torch.arange(10, dtype=torch.int64, device="cuda").flip(0).contiguous()
    start = torch.tensor([0, 3, 9, 5], dtype=torch.int64, device="cuda")
    length = torch.tensor([2, 0, 1, 3], dtype=torch.int64, device="cuda")
    cs, cr = tqpkernels.join_expand(sorted_ids, start, length, host)
    assert int(host[0]) == 6
    ws, wr = oracle.expand(sorted_ids.cpu().numpy(), start.cpu().numpy(), length.cpu().numpy())
    assert np.array_equal(cs.cpu().numpy(), ws) and np.array_equal(cr.cpu().numpy(), wr)


def test_partition_sort(host):
    rng = np.random.default_rng(9)
    n, B = 100_000, 64
    key = rng.integers(-2**40, 2**40, n)
    pay = rng.integers(0, 2**31, (n, 2)).astype(np.int32)
    outs = tqpkernels.partition_sort([torch.from_numpy(key).cuda(), torch.from_numpy(pay).cuda()],
                                     [0, 0], 0, None, n, 1000, B, False, host)
    want_counts, want_outs = oracle.partition([key, pay], 0, None, 1000, B)
    assert host[:B].tolist() == want_counts
    assert int(host[B + 4]) == -1  # no mask prologue ran
    # each output (key, payload, gidx) in the padded layout of its row width
    for got, want, rb in zip(outs, want_outs, (8, 8, 8)):
        starts, total = oracle.padded_layout(want_counts, rb)
        assert got.shape[0] >= total
        g = got.cpu().numpy()
        for st, rows in zip(starts, want):
            assert np.array_equal(g[st:st + rows.shape[0]], rows)
