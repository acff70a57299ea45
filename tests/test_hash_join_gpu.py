"""The partitioned LDS-table join (pf_hash_join_*) against the oracle:
sorted (probe row, build row) pairs bit-exact, payloads gathered by build
row, for unique and multi-match builds, empty and tiny sides, extreme keys,
oversized build buckets (chunked LDS tables), skewed probes, determinism."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from pageflow_amd.join import HashJoin, hash_join

pytestmark = pytest.mark.gpu

I64 = np.iinfo(np.int64)


def _sorted_pairs(probe_idx, build_idx):
    p = probe_idx.cpu().numpy()
    b = build_idx.cpu().numpy()
    pairs = np.stack([b, p], 1) if p.size else np.empty((0, 2), np.int64)
    if pairs.shape[0]:
        pairs = pairs[np.lexsort((pairs[:, 1], pairs[:, 0]))]
    return pairs


def _check(bk, pk, unique, bits=None, pay=True, ordered=True):
    g = torch.Generator().manual_seed(len(bk))
    payload = torch.randint(-(1 << 60), 1 << 60, (len(bk),), generator=g, dtype=torch.int64)
    res = hash_join(torch.as_tensor(bk).cuda(), torch.as_tensor(pk).cuda(),
                    [payload.cuda()] if pay else [], unique=unique, bits=bits, ordered=ordered)
    assert res is not None
    pi, bi, outs = res
    got = _sorted_pairs(pi, bi)
    want = oracle.join_pairs_sorted(np.asarray(bk), np.asarray(pk))
    if unique:
        # a unique build: the oracle's unique-hash pairs are the same multiset
        rb, rp = oracle.unique_hash_pairs([np.asarray(bk)], [np.asarray(pk)])
        w2 = np.stack([rb, rp], 1) if rb.size else np.empty((0, 2), np.int64)
        w2 = w2[np.lexsort((w2[:, 1], w2[:, 0]))] if w2.shape[0] else w2
        assert np.array_equal(w2, want)
    assert np.array_equal(got, want)
    if pay:
        assert np.array_equal(outs[0].cpu().numpy(), payload.numpy()[bi.cpu().numpy()])
    return pi, bi


@pytest.mark.parametrize("ordered", [True, False])
@pytest.mark.parametrize("bits", [0, 4, 8, 12])
def test_unique_random(bits, ordered):
    rng = np.random.default_rng(bits)
    bk = rng.permutation(200_000).astype(np.int64) * 7 - 5
    pk = bk[rng.integers(0, bk.size, 1_000_000)]
    pk[rng.random(pk.size) < 0.1] = 1 << 50
    _check(bk, pk, True, bits, ordered=ordered)


def test_unique_any_order_edges_and_oversized_buckets():
    for bk, pk in (([], [1, 2]), ([1, 2], []), ([1, 2], [3, 4]), ([5], [5, 5, 5])):
        _check(np.array(bk, np.int64), np.array(pk, np.int64), True, pay=False, ordered=False)
    # a build bucket past one LDS table (bits 0: one bucket of 20K rows)
    rng = np.random.default_rng(77)
    bk = rng.permutation(20_000).astype(np.int64)
    pk = rng.integers(-100, 20_100, 300_000)
    _check(bk, pk, True, bits=0, ordered=False)


@pytest.mark.parametrize("bits,nb", [(0, 8000), (6, 150_000), (12, 150_000)])
def test_multi_random(bits, nb):
    # multi-match builds are sorted per bucket in LDS: at most 8192 rows a bucket
    rng = np.random.default_rng(100 + bits)
    bk = rng.integers(0, nb // 3, nb)
    pk = rng.integers(-100, nb // 3 + 10_000, 400_000)
    _check(bk, pk, False, bits)


def test_edges():
    for bk, pk in (([], [1, 2]), ([1, 2], []), ([1, 2], [3, 4]), ([5], [5, 5, 5])):
        for unique in (True, False):
            _check(np.array(bk, np.int64), np.array(pk, np.int64), unique, pay=False)


def test_extreme_keys():
    keys = np.array([I64.min, I64.max, 0, -1, -64, 1], np.int64)
    probe = np.array([I64.min, 5, I64.max, -64, I64.min, 1, 0], np.int64)
    _check(keys, probe, True, bits=2)
    _check(np.concatenate([keys, keys]), probe, False, bits=2)


def test_oversized_build_buckets_take_chunked_tables():
    # 60K unique build rows in one bucket (bits 0): ten LDS chunks
    rng = np.random.default_rng(7)
    bk = rng.permutation(60_000).astype(np.int64)
    pk = rng.integers(-1000, 61_000, 300_000)
    _check(bk, pk, True, bits=0)
    # multi: 6000 rows in one bucket (sorted in LDS), two LDS chunks
    bk = rng.integers(0, 1500, 6000)
    _check(bk, rng.integers(0, 1600, 50_000), False, bits=0)


@pytest.mark.parametrize("bits,nb,nkeys", [(0, 8192, 3000), (0, 8193, 3000), (0, 20_000, 20_000),
                                           (0, 50_000, 3000), (3, 120_000, 40_000)])
def test_multi_sorts_buckets_past_the_lds_capacity_in_hbm(bits, nb, nkeys):
    # buckets of more than 8192 build rows are sorted by (hash, row) in HBM
    # (all 64 hash bits at bits 0), the smaller ones in LDS
    rng = np.random.default_rng(nb + bits)
    bk = rng.integers(0, nkeys, nb)
    pk = rng.integers(-50, nkeys + 50, 200_000)
    _check(bk, pk, False, bits)


def test_multi_many_large_buckets_take_one_global_sort():
    # 64 buckets of ~19K build rows, all past the LDS sort: one sort of the
    # whole side by every hash bit instead of a sort a bucket
    rng = np.random.default_rng(12)
    bk = rng.integers(0, 400_000, 1_200_000)
    pk = rng.integers(-100, 400_100, 300_000)
    _check(bk, pk, False, bits=6)


def test_multi_hot_key_bucket_sorted_in_hbm_is_deterministic():
    rng = np.random.default_rng(8)
    bk = np.concatenate([np.full(12_000, 77, np.int64), rng.integers(0, 30_000, 30_000)])
    rng.shuffle(bk)
    pk = np.where(rng.random(300_000) < 0.0004, 77, rng.integers(0, 31_000, 300_000))
    # 9 bits: 55 hash bits sorted in 7 passes, the last one's output copied back
    _check(bk, pk, False, bits=9)
    b, p = torch.as_tensor(bk).cuda(), torch.as_tensor(pk).cuda()
    x = hash_join(b, p, unique=False, bits=9)
    y = hash_join(b, p, unique=False, bits=9)
    assert torch.equal(x[0], y[0]) and torch.equal(x[1], y[1])


def test_skewed_probe_and_duplicate_heavy_build():
    rng = np.random.default_rng(3)
    bk = np.repeat(np.arange(2000, dtype=np.int64), rng.integers(1, 9, 2000))
    pk = np.where(rng.random(500_000) < 0.3, 17, rng.integers(0, 2500, 500_000))
    _check(bk, pk, False, bits=5)


def test_repeat_is_byte_identical():
    rng = np.random.default_rng(11)
    bk = torch.from_numpy(rng.integers(0, 10_000, 30_000)).cuda()
    pk = torch.from_numpy(rng.integers(0, 12_000, 200_000)).cuda()
    a = hash_join(bk, pk, unique=False, bits=6)
    b = hash_join(bk, pk, unique=False, bits=6)
    assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1])


def test_plan_reuse():
    rng = np.random.default_rng(12)
    bk = torch.from_numpy(rng.permutation(100_000).astype(np.int64)).cuda()
    plan = HashJoin(100_000, 400_000, unique=True)
    for seed in range(2):
        pk = torch.from_numpy(np.random.default_rng(seed).integers(0, 110_000, 400_000)).cuda()
        n = plan.prepare(bk, pk)
        op = torch.empty(n, dtype=torch.int64, device="cuda")
        ob = torch.empty(n, dtype=torch.int64, device="cuda")
        plan.write(op, ob)
        got = _sorted_pairs(op, ob)
        assert np.array_equal(got, oracle.join_pairs_sorted(bk.cpu().numpy(), pk.cpu().numpy()))


@pytest.mark.parametrize("bits", [4, 8])
def test_unique_path_on_repeated_build_keys_answers_each_keys_lowest_row(bits):
    # The unique path's LDS table keeps the lower build row of a repeated hash
    # (a CAS on the u16 slot's 32-bit word, racing with inserts into the
    # word's other slot).  Every bucket here fits one LDS chunk, so each probe
    # row of a present key pairs with exactly that key's first build row.
    rng = np.random.default_rng(21 + bits)
    bk = rng.integers(0, 3000, 20_000).astype(np.int64)
    pk = rng.integers(-50, 3100, 300_000).astype(np.int64)
    pi, bi, _ = hash_join(torch.as_tensor(bk).cuda(), torch.as_tensor(pk).cuda(), [],
                          unique=True, bits=bits)
    keys, first = np.unique(bk, return_index=True)
    pos = np.searchsorted(keys, pk)
    hit = (pos < keys.size) & (keys[np.minimum(pos, keys.size - 1)] == pk)
    want = np.stack([first[pos[hit]], np.nonzero(hit)[0]], 1)
    want = want[np.lexsort((want[:, 1], want[:, 0]))]
    assert np.array_equal(_sorted_pairs(pi, bi), want)


@pytest.mark.parametrize("nb,unique", [(5120, True), (5121, True), (10_240, True),
                                       (4096, False), (4097, False)])
def test_build_bucket_at_the_lds_chunk_capacity(nb, unique):
    # One bucket (bits 0) of exactly an LDS chunk's rows (5120 unique, 4096
    # multi), one row past it (a second chunk of one row) and two full chunks:
    # the u16 table at its fullest, and the chunk boundary.
    rng = np.random.default_rng(nb)
    bk = (rng.permutation(nb).astype(np.int64) * 3 + 1 if unique
          else rng.integers(0, nb // 2, nb).astype(np.int64))
    pk = rng.integers(-10, 3 * nb + 10, 250_000).astype(np.int64)
    _check(bk, pk, unique, bits=0)
