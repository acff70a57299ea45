"""Every benchmark configuration at its full size, checked on the GPU.

At full size the oracle cannot join the whole input in test time, so each
case is checked two ways: through size-independent properties of the whole
result (pair counts equal to the generator's exact counts, key equality of
every pair, no pair twice, masks equal to key presence, per-group sums by
an int64 scatter-add), and bit-exactly against the oracle on a slice of the
input (every pair of a bounded set of probe rows, or whole joins and
requests of the batch and the serving stream).

  C2   16M unique int64 build (+ payload) x 640M probe, 10% absent
  C2b  the standalone radix partition of C2's 640M keys at 8 and 12 bits
  C3   10M build over 2.5M keys (geometric duplicates) x 400M Zipf(1.1)
       probe, 20% absent: inner pairs, semi and anti masks, and a residual
       condition evaluated inside the join
  C4   256 uneven joins (1K-4M unique int32 build, 8x probe, 15% absent)
  C5   the 2048-request serving stream over 8 resident dimension tables
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from pageflow_amd import radix, serving
from pageflow_amd.batch import BatchJoin, make_batch
from pageflow_amd.join import HashJoin
from pageflow_amd.operators.join_core import semi_anti_mask
from pageflow_amd.workloads import make_large_join, make_skewed_join

pytestmark = pytest.mark.gpu

CHUNK = 1 << 27


def say(msg):
    print(f"[full-size] {msg}", flush=True)


def _all_equal_in_chunks(fn, n):
    return all(bool(fn(a, min(n, a + CHUNK)).all()) for a in range(0, n, CHUNK))


@pytest.fixture(scope="module")
def c2():
    dev = torch.device("cuda")
    build, pays, probe, expect = make_large_join(16_000_000, 640_000_000, 2, dev)
    yield build, pays, probe, expect
    del build, pays, probe
    torch.cuda.empty_cache()


@pytest.mark.parametrize("ordered", [True, False])
def test_c2_large_join_at_full_size(c2, ordered):
    build, pays, probe, expect = c2
    nb, np_ = build.numel(), probe.numel()
    say(f"C2 join {nb} x {np_} ordered={ordered}")
    plan = HashJoin(nb, np_, unique=True, device=build.device, ordered=ordered)
    n = plan.prepare(build, probe)
    assert n == (expect if ordered else None)
    rows = n if ordered else np_
    op = torch.empty(rows, dtype=torch.int64, device=build.device)
    ob = torch.empty_like(op)
    pay = torch.empty_like(op)
    n = plan.write(op, ob, [pays[0]], [pay])
    assert n == expect
    op, ob, pay = op[:n], ob[:n], pay[:n]
    torch.cuda.synchronize()
    del plan
    say("C2 properties")
    assert _all_equal_in_chunks(lambda a, b: build[ob[a:b]] == probe[op[a:b]], n)
    assert _all_equal_in_chunks(lambda a, b: pay[a:b] == pays[0][ob[a:b]], n)
    assert int(torch.unique(op).numel()) == n            # a probe row matches at most once
    # bit-exact on a slice: every pair of the first 4M probe rows
    say("C2 oracle slice")
    m = 4_000_000
    sel = op < m
    got = torch.stack([op[sel], ob[sel]], 1).cpu().numpy()
    got = got[np.lexsort((got[:, 1], got[:, 0]))]
    wb, wp = oracle.unique_hash_pairs([build.cpu().numpy()], [probe[:m].cpu().numpy()])
    assert np.array_equal(got[:, 0], wp) and np.array_equal(got[:, 1], wb)


@pytest.mark.parametrize("bits", [8, 12])
def test_c2b_partition_at_full_size(c2, bits):
    _, _, probe, _ = c2
    n = probe.numel()
    dev = probe.device
    say(f"C2b partition of {n} keys into 2^{bits}")
    ids = torch.arange(n, dtype=torch.int32, device=dev)
    keys_out, ids_out, offs = radix.radix_partition(probe, bits, ids)
    torch.cuda.synchronize()
    offs_h = offs.cpu().numpy()
    assert offs_h[0] == 0 and offs_h[-1] == n and bool((np.diff(offs_h) >= 0).all())
    # every row is its source row, in its own bucket, and rows keep their
    # input order inside a bucket
    assert _all_equal_in_chunks(lambda a, b: keys_out[a:b] == probe[ids_out[a:b].long()], n)
    bucket_at = torch.repeat_interleave(torch.arange(1 << bits, device=dev),
                                        torch.from_numpy(np.diff(offs_h)).to(dev))
    assert _all_equal_in_chunks(
        lambda a, b: oracle_buckets(keys_out[a:b], bits) == bucket_at[a:b], n)
    step = ids_out[1:].long() - ids_out[:-1].long()
    new_bucket = bucket_at[1:] != bucket_at[:-1]
    assert bool(((step > 0) | new_bucket).all())
    # bit-exact against the oracle on a 2M-key slice
    m = 2_000_000
    sample = probe[:m].cpu().numpy()
    wk, wv, wc = oracle.radix_partition(sample, np.arange(m, dtype=np.int32), bits, 0,
                                        hashed=True)
    gk, gv, goffs = radix.radix_partition(probe[:m].contiguous(), bits,
                                          torch.arange(m, dtype=torch.int32, device=dev))
    assert np.array_equal(gk.cpu().numpy(), wk) and np.array_equal(gv.cpu().numpy(), wv)
    assert np.array_equal(np.diff(goffs.cpu().numpy()), wc)
    head = torch.bincount(oracle_buckets(probe[:m], bits), minlength=1 << bits)
    assert np.array_equal(head.cpu().numpy(), wc)


def oracle_buckets(keys: torch.Tensor, bits: int) -> torch.Tensor:
    """The top ``bits`` bits of mix64 of each key, on the GPU in plain torch
    arithmetic (uint64 wrap-around written with int64 and masks)."""
    k = keys.clone()
    m1, m2 = -49064778989728563, -4265267296055464877   # 0xff51afd7ed558ccd, 0xc4ceb9fe1a85ec53

    def xs33(x):
        return x ^ ((x >> 33) & ((1 << 31) - 1))
    k = xs33(k)
    k = k * m1
    k = xs33(k)
    k = k * m2
    k = xs33(k)
    return (k >> (64 - bits)) & ((1 << bits) - 1)


def test_mix64_torch_twin_matches_the_golden_vectors():
    import os
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "mix64_golden.npz"))
    got = oracle_buckets(torch.from_numpy(d["keys"].astype(np.int64)).cuda(), 9)
    assert np.array_equal(got.cpu().numpy(), d["bucket512"])


def test_c3_skewed_join_and_masks_at_full_size():
    dev = torch.device("cuda")
    nb, nd, np_ = 10_000_000, 2_500_000, 400_000_000
    say("C3 generate")
    build, probe, want = make_skewed_join(nb, nd, np_, 3, dev)
    plan = HashJoin(nb, np_, unique=False, device=dev)
    n = plan.prepare(build, probe)
    assert n == want["pairs"]
    op = torch.empty(n, dtype=torch.int64, device=dev)
    ob = torch.empty_like(op)
    plan.write(op, ob)
    torch.cuda.synchronize()
    del plan
    say(f"C3 properties of {n} pairs")
    assert _all_equal_in_chunks(lambda a, b: build[ob[a:b]] == probe[op[a:b]], n)
    assert int(torch.unique(op * nb + ob).numel()) == n   # no pair twice
    m = 2_000_000
    sel = op < m
    got = torch.stack([ob[sel], op[sel]], 1).cpu().numpy()     # (build row, probe row)
    del op, ob
    say("C3 oracle slice")
    want_pairs = oracle.join_pairs_sorted(build.cpu().numpy(), probe[:m].cpu().numpy())
    assert np.array_equal(got[np.lexsort((got[:, 1], got[:, 0]))], want_pairs)
    say("C3 masks")
    present = probe <= nd
    semi = semi_anti_mask(build, probe, anti=False)
    assert int(semi.sum()) == want["present"] and bool((semi == present).all())
    del semi
    anti = semi_anti_mask(build, probe, anti=True)
    assert bool((anti == ~present).all())


@pytest.mark.parametrize("op", ['<', '<>'])
def test_c3_residual_inside_the_join_at_full_size(op):
    """C3 with the residual b_pay OP p_pay (int64 payloads in [0, 1000))
    evaluated inside the join ('<' takes the range form, '<>' the fast form
    with window sweeps): the same pairs as the join followed by the filter
    (count and sorted (build, probe) keys at full size), every kept pair's
    keys equal and condition true, and the oracle's filtered pairs on the
    first 2M probe rows."""
    from pageflow_amd.operators import residual
    dev = torch.device("cuda")
    nb, nd, np_ = 10_000_000, 2_500_000, 400_000_000
    say(f"C3 residual {op}: generate")
    build, probe, want = make_skewed_join(nb, nd, np_, 3, dev)
    g = torch.Generator(device=dev).manual_seed(11)
    bpay = torch.randint(0, 1000, (nb,), generator=g, device=dev)
    ppay = torch.randint(0, 1000, (np_,), generator=g, device=dev)
    cond = [['b_pay#11', op, 'p_pay#12']]
    prog = residual.compile_conditions({11: bpay}, {12: ppay}, cond)
    plan = HashJoin(nb, np_, unique=False, device=dev)
    kept = plan.prepare(build, probe, residual=prog)
    assert plan.candidates == want["pairs"]
    fp = torch.empty(kept, dtype=torch.int64, device=dev)
    fb = torch.empty_like(fp)
    plan.write(fp, fb)
    del plan
    say(f"C3 residual {op}: the join, then the filter")
    plan = HashJoin(nb, np_, unique=False, device=dev)
    n = plan.prepare(build, probe)
    op_ = torch.empty(n, dtype=torch.int64, device=dev)
    ob = torch.empty_like(op_)
    plan.write(op_, ob)
    del plan
    kb, kp = residual.filter_pairs(prog, ob, op_)
    del op_, ob
    assert kb.numel() == kept
    assert torch.equal(torch.sort(fb * np_ + fp).values, torch.sort(kb * np_ + kp).values)
    del kb, kp
    say(f"C3 residual {op}: properties of {kept} pairs")
    cmp = (lambda a, b: a < b) if op == '<' else (lambda a, b: a != b)
    assert _all_equal_in_chunks(lambda a, b: build[fb[a:b]] == probe[fp[a:b]], kept)
    assert _all_equal_in_chunks(lambda a, b: cmp(bpay[fb[a:b]], ppay[fp[a:b]]), kept)
    m = 2_000_000
    sel = fp < m
    got = torch.stack([fb[sel], fp[sel]], 1).cpu().numpy()
    del fb, fp
    say(f"C3 residual {op}: oracle slice")
    bk, pk = build.cpu().numpy(), probe[:m].cpu().numpy()
    lm, rm = oracle.sort_search_pairs(bk, pk)
    wl, wr = oracle.residual_filter({11: bpay.cpu().numpy()}, {12: ppay[:m].cpu().numpy()},
                                    cond, lm, rm)
    want_pairs = np.stack([wl, wr], 1)
    want_pairs = want_pairs[np.lexsort((want_pairs[:, 1], want_pairs[:, 0]))]
    assert np.array_equal(got[np.lexsort((got[:, 1], got[:, 0]))], want_pairs)


def test_c4_batch_of_256_joins_at_full_size():
    say("C4 generate")
    nb, np_, bkeys, pkeys, expect = make_batch()
    assert len(nb) == 256
    bj = BatchJoin(nb, np_)
    st = bj.capture(bkeys, pkeys)
    bj.replay()
    st.synchronize()
    counts = bj.counts.cpu().numpy()
    assert np.array_equal(counts, expect)
    bo = np.concatenate(([0], np.cumsum(nb)))
    po = np.concatenate(([0], np.cumsum(np_)))
    say("C4 key equality of every pair")
    for j in range(256):
        op, ob = bj.pairs(j)
        assert bool((bkeys[bo[j] + ob] == pkeys[po[j] + op]).all()), j
        assert int(torch.unique(op).numel()) == op.numel(), j
    say("C4 oracle joins")
    order = np.argsort(nb)
    for j in [int(order[0]), int(order[64]), int(order[128]), int(order[192]), int(order[-1])]:
        wb, wp = oracle.unique_hash_pairs(
            [bkeys[bo[j]:bo[j + 1]].cpu().numpy().astype(np.int64)],
            [pkeys[po[j]:po[j + 1]].cpu().numpy().astype(np.int64)])
        op, ob = bj.pairs(j)
        srt = torch.argsort(op, stable=True)
        assert np.array_equal(op[srt].cpu().numpy(), wp), j
        assert np.array_equal(ob[srt].cpu().numpy(), wb), j
    bj.close()


def _group_sums(keys, grp, fk, amt, G):
    """SUM and COUNT per group of the fact rows' matched dimension rows, by
    an int64 scatter-add in plain torch (dimension keys are 1..n)."""
    lookup = torch.full((int(keys.max()) + 2,), -1, dtype=torch.int64, device=keys.device)
    lookup[keys.long()] = grp.long()
    g = lookup[fk.long()]
    hit = g >= 0
    s = torch.zeros(G, dtype=torch.int64, device=keys.device).scatter_add_(0, g[hit], amt[hit])
    c = torch.bincount(g[hit], minlength=G)
    return s.cpu().numpy(), c.cpu().numpy()


def test_c5_serving_stream_at_full_size():
    say("C5 generate")
    wl = serving.make_workload(2048, rate=0.0, seed=5)
    eng = serving.ServingEngine(hbm_budget=8 << 30, max_active=40, step_rows=1 << 25)
    for d, (k, g) in wl.dims.items():
        eng.register_dim(d, k, g, wl.G)
    reqs = serving.with_arrivals(wl.requests, rate=20000.0)
    say("C5 serve")
    res = eng.run(reqs)
    eng.check_tables()
    assert eng.stats["max_items"] > 1
    expected = {}
    for req, r in zip(reqs, res):
        key = (req.dim, req.fk.data_ptr())
        if key not in expected:
            k, g = wl.dims[req.dim]
            expected[key] = _group_sums(k, g, req.fk, req.amt, wl.G)
        ws, wc = expected[key]
        assert np.array_equal(r.sum, ws) and np.array_equal(r.cnt, wc), req.rid
    say("C5 oracle and alone-vs-mix sample")
    smallest = sorted(range(len(reqs)), key=lambda i: reqs[i].fk.numel())[:4]
    for i in smallest:
        req = reqs[i]
        k, g = wl.dims[req.dim]
        b, p = oracle.sort_search_pairs(k.cpu().numpy().astype(np.int64),
                                        req.fk.cpu().numpy().astype(np.int64))
        ws, wc = oracle.group_sum_count(g.cpu().numpy()[b], req.amt.cpu().numpy()[p], wl.G)
        assert np.array_equal(res[i].sum, ws) and np.array_equal(res[i].cnt, wc), i
    for i in smallest + [0, 1, 1000, 2047]:
        alone = eng.run_alone(reqs[i])
        assert np.array_equal(alone.sum, res[i].sum) and np.array_equal(alone.cnt, res[i].cnt)
