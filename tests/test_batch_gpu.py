"""Batched joins (pf_batch_*): every join's pairs equal the oracle's
unique-hash pairs of that join, in the batched pipeline, in the per-join
baseline, and on hipGraph replays (several, with eager work between them);
every tunable shape gives the same answer; a join probed by slices writes
its pairs in (slice, probe row) order, with slices held in LDS whole or in
sub-ranges, and in several passes when a table's build rows overflow it."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from pageflow_amd.batch import BatchJoin, make_batch

pytestmark = pytest.mark.gpu


def _check_all(bj, nb, np_, bkeys, pkeys):
    """Each join's pairs against the oracle's (probe-row order); a sliced
    join's pairs come in slice order, so they are compared sorted by probe
    row (a probe row has at most one pair)."""
    counts = bj.counts.cpu().numpy()
    bo = np.concatenate(([0], np.cumsum(nb)))
    po = np.concatenate(([0], np.cumsum(np_)))
    bk = bkeys.cpu().numpy().astype(np.int64)
    pk = pkeys.cpu().numpy().astype(np.int64)
    op = bj.out_probe.cpu().numpy()
    ob = bj.out_build.cpu().numpy()
    for j in range(len(nb)):
        rb, rp = oracle.unique_hash_pairs([bk[bo[j]:bo[j + 1]]], [pk[po[j]:po[j + 1]]])
        assert counts[j] == rb.size, j
        gp, gb = op[po[j]:po[j] + counts[j]], ob[po[j]:po[j] + counts[j]]
        order = np.argsort(gp, kind="stable")
        assert np.array_equal(gp[order], rp), j
        assert np.array_equal(gb[order], rb), j


@pytest.fixture(scope="module")
def small_batch():
    return make_batch(n_joins=24, seed=9, lo=10, hi=6e5)


@pytest.mark.parametrize("per_join", [False, True])
def test_batch_matches_oracle(small_batch, per_join):
    nb, np_, bkeys, pkeys, expect = small_batch
    bj = BatchJoin(nb, np_)
    bj.run(bkeys, pkeys, per_join=per_join)
    torch.cuda.synchronize()
    assert np.array_equal(bj.counts.cpu().numpy(), expect)
    _check_all(bj, nb, np_, bkeys, pkeys)


@pytest.mark.parametrize("tune", [[(4, 4096, 512, 0)] * 3,
                                  [(2, 1024, 256, 0), (4, 65536, 512, 0), (2, 16384, 256, 0)],
                                  [(2, 4096, 256, 0), (2, 4096, 512, 128), (4, 16384, 512, 256)],
                                  [(2, 4096, 256, 64), (2, 1024, 256, 64), (2, 16384, 256, 1024)]])
def test_every_tunable_gives_the_same_pairs(small_batch, tune):
    nb, np_, bkeys, pkeys, expect = small_batch
    bj = BatchJoin(nb, np_, tune)
    bj.run(bkeys, pkeys)
    torch.cuda.synchronize()
    _check_all(bj, nb, np_, bkeys, pkeys)


def test_graph_replays_equal_eager(small_batch):
    nb, np_, bkeys, pkeys, expect = small_batch
    eager = BatchJoin(nb, np_)
    eager.run(bkeys, pkeys)
    torch.cuda.synchronize()
    want = (eager.out_probe.clone(), eager.out_build.clone(), eager.counts.clone())
    bj = BatchJoin(nb, np_)
    st = bj.capture(bkeys, pkeys)
    for rep in range(4):
        bj.out_probe.fill_(-7)
        bj.counts.fill_(-7)
        torch.cuda.synchronize()
        bj.replay()
        st.synchronize()
        n = int(want[2].sum())
        assert torch.equal(bj.counts, want[2]), rep
        for j in range(len(nb)):
            a, b = bj.pairs(j)
            ea, eb = eager.pairs(j)
            assert torch.equal(a, ea) and torch.equal(b, eb), (rep, j)
        # eager work between replays
        eager.run(bkeys, pkeys)
        torch.cuda.synchronize()
    assert n == int(expect.sum())


def test_empty_and_tiny_joins():
    nb = np.array([0, 1, 5, 0, 3])
    np_ = np.array([4, 0, 9, 0, 24])
    g = torch.Generator().manual_seed(1)
    bkeys = torch.randint(-50, 50, (int(nb.sum()),), generator=g, dtype=torch.int64)
    # unique within each join
    bkeys = torch.tensor([7, -3, 4, 9, 11, 12, 5, 6, 1], dtype=torch.int32)[:int(nb.sum())]
    pkeys = torch.randint(-4, 14, (int(np_.sum()),), generator=g).to(torch.int32)
    bj = BatchJoin(nb, np_)
    bj.run(bkeys.cuda(), pkeys.cuda())
    torch.cuda.synchronize()
    _check_all(bj, nb, np_, bkeys, pkeys)


@pytest.mark.parametrize("per_join", [False, True])
@pytest.mark.parametrize("slice_kb,sbits", [(64, 4), (512, 1)])
def test_sliced_join_order(per_join, slice_kb, sbits):
    # one 40K-row join, slices of its 1 MiB table: 16 slices of 64 KiB (each
    # held in LDS whole), or 2 of 512 KiB (each taken in 4 sub-ranges)
    nb, np_ = np.array([40_000]), np.array([320_000])
    g = torch.Generator().manual_seed(3)
    bkeys = (torch.randperm(40_000, generator=g) * 3 - 5).to(torch.int32)
    pkeys = bkeys[torch.randint(0, 40_000, (320_000,), generator=g)].clone()
    pkeys[::7] = 2  # absent (not 1 mod 3)
    tune = [(2, 4096, 256, 0), (2, 4096, 512, slice_kb), (2, 16384, 512, 0)]
    bj = BatchJoin(nb, np_, tune)
    assert bj.info()["sliced_joins"] == 1
    bj.run(bkeys.cuda(), pkeys.cuda(), per_join=per_join)
    torch.cuda.synchronize()
    _check_all(bj, nb, np_, bkeys, pkeys)
    p, _ = bj.pairs(0)
    p = p.cpu().numpy()
    cap = 1 << 17  # pow2 >= 2 * 40000 slots (1 MiB)
    h = oracle.fmix32_np(pkeys.numpy()[p]).astype(np.uint64)  # the batch's slot hash
    sl = (h & np.uint64(cap - 1)) >> np.uint64(17 - sbits)
    key = sl.astype(np.int64) * (1 << 32) + p
    assert np.all(np.diff(key) > 0)  # slices ascending, probe rows ascending in a slice


@pytest.mark.parametrize("per_join", [False, True])
@pytest.mark.parametrize("slice_kb,n_hot,hot_slots", [(64, 9000, 1 << 13), (512, 15000, 1 << 14)])
def test_lds_table_passes(per_join, slice_kb, n_hot, hot_slots):
    # one 40K-row join: 2^17 slots.  n_hot keys home in the table's first
    # hot_slots slots, which is one LDS table (a 64 KiB slice of 2^13 slots,
    # or the first 2^14-slot sub-range of a 512 KiB slice): more build rows
    # than the 7/8 of its slots a pass holds, so it is joined in passes; the
    # hot keys' clusters also wrap around the table's end
    cap, mask = 1 << 17, (1 << 17) - 1
    cand = np.arange(1, 1 << 22, dtype=np.int64)
    slot = (oracle.fmix32_np(cand).astype(np.int64) & mask)  # the batch's slot hash
    hot = cand[slot < hot_slots][:n_hot]
    rest = np.random.default_rng(5).permutation(cand[slot >= hot_slots][:400_000])[:40_000 - n_hot]
    bk = np.random.default_rng(7).permutation(np.concatenate([hot, rest])).astype(np.int32)
    assert hot.size == n_hot and np.unique(bk).size == bk.size == 40_000
    rng = np.random.default_rng(6)
    pk = np.concatenate([bk[rng.integers(0, bk.size, 300_000)], -rng.integers(1, 10**6, 20_000)])
    pk = rng.permutation(pk).astype(np.int32)
    nb, np_ = np.array([bk.size]), np.array([pk.size])
    tune = [(2, 4096, 256, 0), (2, 4096, 512, slice_kb), (2, 16384, 512, 0)]
    bj = BatchJoin(nb, np_, tune)
    bj.run(torch.from_numpy(bk).cuda(), torch.from_numpy(pk).cuda(), per_join=per_join)
    torch.cuda.synchronize()
    assert bj.info()["sliced_joins"] == 1 and bj.info()["lds_overflow"] > 0
    _check_all(bj, nb, np_, torch.from_numpy(bk), torch.from_numpy(pk))


@pytest.mark.parametrize("per_join", [False, True])
def test_nine_bit_slices(per_join):
    # a 1.2M-row join: 2^22 slots (32 MiB) in 64 KiB slices is 512 slices,
    # a 9-bit slice partition of both sides; pairs come in (slice, probe
    # row) order
    nbj, npj = 1_200_000, 2_000_000
    g = torch.Generator().manual_seed(11)
    bkeys = (torch.randperm(nbj, generator=g) * 3 + 1).to(torch.int32)
    pkeys = bkeys[torch.randint(0, nbj, (npj,), generator=g)].clone()
    pkeys[::9] = 3  # absent (not 1 mod 3)
    tune = [(2, 4096, 256, 0), (2, 4096, 512, 0), (2, 16384, 512, 64)]
    bj = BatchJoin(np.array([nbj]), np.array([npj]), tune)
    assert bj.info()["sliced_joins"] == 1
    bj.run(bkeys.cuda(), pkeys.cuda(), per_join=per_join)
    torch.cuda.synchronize()
    _check_all(bj, np.array([nbj]), np.array([npj]), bkeys, pkeys)
    p, _ = bj.pairs(0)
    p = p.cpu().numpy()
    cap = 1 << 22
    h = oracle.fmix32_np(pkeys.numpy()[p]).astype(np.uint64)
    sl = (h & np.uint64(cap - 1)) >> np.uint64(22 - 9)
    key = sl.astype(np.int64) * (1 << 32) + p
    assert np.all(np.diff(key) > 0)


@pytest.mark.parametrize("per_join", [False, True])
def test_mixed_slice_widths(per_join):
    # joins of 9-bit slices, of fewer slice bits and probed whole, in one
    # batch (one segmented partition launch for every slice width)
    nb = np.array([3_000, 1_100_000, 90_000, 1_300_000, 700])
    np_ = nb * 3
    rng = np.random.default_rng(21)
    bk, pk = [], []
    for j in range(len(nb)):
        b = rng.permutation(nb[j]).astype(np.int64) * 5 + 2
        p = b[rng.integers(0, nb[j], np_[j])]
        p[rng.random(np_[j]) < 0.1] = 4  # absent
        bk.append(b)
        pk.append(p)
    bkeys = torch.from_numpy(np.concatenate(bk).astype(np.int32))
    pkeys = torch.from_numpy(np.concatenate(pk).astype(np.int32))
    tune = [(2, 4096, 256, 0), (2, 4096, 512, 64), (2, 16384, 512, 64)]
    bj = BatchJoin(nb, np_, tune)
    assert bj.info()["sliced_joins"] == 3
    bj.run(bkeys.cuda(), pkeys.cuda(), per_join=per_join)
    torch.cuda.synchronize()
    _check_all(bj, nb, np_, bkeys, pkeys)


@pytest.mark.parametrize("per_join", [False, True])
def test_segment_chunk_edges(per_join):
    # sliced joins whose build and probe sides end one row before, on and
    # one row after a segment chunk boundary (64K rows a chunk), and one
    # spanning two chunks and a row
    nb = np.array([65_535, 65_536, 65_537, 131_073])
    np_ = np.array([65_536, 131_071, 131_073, 196_608])
    rng = np.random.default_rng(33)
    bk, pk = [], []
    for j in range(len(nb)):
        b = rng.permutation(nb[j]).astype(np.int64) * 7 - 3
        p = b[rng.integers(0, nb[j], np_[j])]
        p[rng.random(np_[j]) < 0.2] = 5  # absent (not 4 mod 7)
        bk.append(b)
        pk.append(p)
    bkeys = torch.from_numpy(np.concatenate(bk).astype(np.int32))
    pkeys = torch.from_numpy(np.concatenate(pk).astype(np.int32))
    tune = [(2, 4096, 256, 0), (2, 4096, 512, 64), (2, 16384, 512, 64)]
    bj = BatchJoin(nb, np_, tune)
    assert bj.info()["sliced_joins"] == 4
    bj.run(bkeys.cuda(), pkeys.cuda(), per_join=per_join)
    torch.cuda.synchronize()
    _check_all(bj, nb, np_, bkeys, pkeys)
