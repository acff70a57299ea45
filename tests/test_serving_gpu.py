"""The serving engine (pf_serve_step + pageflow_amd.serving): every request's
SUM/COUNT per group equals the oracle's join-then-group-by, and is the same
served alone, in a Poisson mix of overlapping requests, under an HBM budget
that forces dimension tables out and back in, and with steps small enough
to split each request over many launches."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from pageflow_amd import serving

pytestmark = pytest.mark.gpu


def _expected(wl, req):
    keys, grp = wl.dims[req.dim]
    dk = keys.cpu().numpy().astype(np.int64)
    fk = req.fk.cpu().numpy().astype(np.int64)
    b, p = oracle.sort_search_pairs(dk, fk)
    return oracle.group_sum_count(grp.cpu().numpy()[b], req.amt.cpu().numpy()[p], wl.G)


def _engine(wl, **kw):
    eng = serving.ServingEngine(**kw)
    for d, (k, g) in wl.dims.items():
        eng.register_dim(d, k, g, wl.G)
    return eng


@pytest.fixture(scope="module")
def small():
    # 4 dims, 32 requests of 1K-300K fact rows arriving faster than served
    return serving.make_workload(32, rate=20000.0, seed=11, n_dims=4, fact_lo=1e3,
                                 fact_hi=3e5, dim_lo=50, dim_hi=5e4, slices_per_dim=3)


def test_results_match_oracle_alone_and_in_the_mix(small):
    eng = _engine(small, max_active=8, step_rows=1 << 19, rows_per_wg=1 << 14)
    mixed = eng.run(small.requests)
    eng.check_tables()
    assert eng.stats["max_items"] > 1  # requests shared steps
    for req, res in zip(small.requests, mixed):
        ws, wc = _expected(small, req)
        assert np.array_equal(res.sum, ws), req.rid
        assert np.array_equal(res.cnt, wc), req.rid
        alone = eng.run_alone(req)
        assert np.array_equal(alone.sum, res.sum) and np.array_equal(alone.cnt, res.cnt)
        assert res.done >= res.admitted >= res.arrival - 1e-9


def test_budget_evicts_and_readmits(small):
    sizes = [int(np.int64(1) << int(np.ceil(np.log2(max(2 * n, 16))))) * 8 + 16
             for n in small.dim_rows]
    budget = max(sizes) + sorted(sizes)[-2] + 4096  # about two tables at once
    eng = _engine(small, hbm_budget=budget, max_active=4, step_rows=1 << 18,
                  rows_per_wg=1 << 13)
    res = eng.run(small.requests)
    assert eng.stats["evictions"] > 0
    assert eng.peak_used <= budget
    for req, r in zip(small.requests, res):
        ws, wc = _expected(small, req)
        assert np.array_equal(r.sum, ws) and np.array_equal(r.cnt, wc), req.rid


def test_tiny_steps_split_requests_without_changing_results(small):
    reqs = small.requests[:8]
    eng = _engine(small, max_active=3, step_rows=4096, rows_per_wg=512)
    res = eng.run(reqs, realtime=False)
    assert eng.stats["steps"] > 8
    for req, r in zip(reqs, res):
        ws, wc = _expected(small, req)
        assert np.array_equal(r.sum, ws) and np.array_equal(r.cnt, wc), req.rid


def test_empty_request_and_absent_keys():
    dev = torch.device("cuda")
    keys = torch.arange(1, 101, dtype=torch.int32, device=dev)
    grp = (keys % 25).to(torch.int32)
    eng = serving.ServingEngine(max_active=2)
    eng.register_dim(0, keys, grp, 25)
    empty = serving.Request(0, 0, torch.empty(0, dtype=torch.int32, device=dev),
                            torch.empty(0, dtype=torch.int64, device=dev))
    fk = torch.tensor([5, 500, -1, 5, 100], dtype=torch.int32, device=dev)
    amt = torch.tensor([10, 20, 30, -64, 7], dtype=torch.int64, device=dev)
    r = eng.run([empty, serving.Request(1, 0, fk, amt)], realtime=False)
    assert not r[0].sum.any() and not r[0].cnt.any()
    want_s = np.zeros(25, np.int64)
    want_c = np.zeros(25, np.int64)
    want_s[5] = 10 - 64
    want_c[5] = 1          # -64 is the null amount: summed, not counted
    want_s[0] = 7
    want_c[0] = 1
    assert np.array_equal(r[1].sum, want_s) and np.array_equal(r[1].cnt, want_c)


def test_dimension_layouts_staged_l2_and_hashed():
    # dense dimensions staged in LDS (<= 65536 key values, one at exactly
    # that many, one with keys from 1000 on), a dense one looked up in L2
    # (100K key values) and a hashed one (sparse keys)
    dev = torch.device("cuda")
    g = torch.Generator(device=dev).manual_seed(21)
    dims = {
        0: torch.randperm(300, generator=g, device=dev) + 1000,
        1: torch.randperm(65536, generator=g, device=dev) + 1,
        2: torch.randperm(100_000, generator=g, device=dev) + 1,
        3: (torch.randperm(20_000, generator=g, device=dev) + 1) * 7919,
    }
    eng = serving.ServingEngine(max_active=6, step_rows=1 << 18, rows_per_wg=1 << 14)
    G = 25
    groups = {}
    for d, k in dims.items():
        k = k.to(torch.int32)
        grp = torch.randint(0, G, (k.numel(),), generator=g, device=dev, dtype=torch.int32)
        dims[d], groups[d] = k, grp
        eng.register_dim(d, k, grp, G)
    reqs = []
    for r in range(12):
        d = r % 4
        k = dims[d]
        nf = 50_000 + 37_000 * r
        fk = k[torch.randint(0, k.numel(), (nf,), generator=g, device=dev)]
        fk[::9] = -5  # absent
        amt = torch.randint(100, 10**7, (nf,), generator=g, device=dev)
        amt[::13] = -64
        reqs.append(serving.Request(r, d, fk, amt))
    res = eng.run(reqs, realtime=False)
    eng.check_tables()
    assert [eng.dims[d].table.dense for d in range(4)] == [True, True, True, False]
    for req, r in zip(reqs, res):
        dk = dims[req.dim].cpu().numpy().astype(np.int64)
        b, p = oracle.sort_search_pairs(dk, req.fk.cpu().numpy().astype(np.int64))
        ws, wc = oracle.group_sum_count(groups[req.dim].cpu().numpy()[b],
                                        req.amt.cpu().numpy()[p], G)
        assert np.array_equal(r.sum, ws) and np.array_equal(r.cnt, wc), req.rid


def test_step_rejects_bad_items():
    it = (serving.ServeItem * 1)()
    it[0].n = 10
    it[0].G = 25
    it[0].slots = 24  # not a power of two
    from pageflow_amd import _lib
    import ctypes
    rc = _lib.load().pf_serve_step(ctypes.cast(it, ctypes.c_void_p), 1, 1 << 14,
                                   ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc != 0 and b"power-of-two" in _lib.load().pf_last_error()
