"""The residual program inside the partitioned join (pf_hash_join_residual_*,
``HashJoin.prepare(residual=...)``).

On the pair filter's random condition trees (every column dtype, every
literal kind) the kept pairs equal the oracle's residual filter applied to
the oracle's join pairs, as multisets, for unique and multi-match builds,
one-pass and two-level partitions; in order they equal the unfused path
(the join's pairs, then ``residual.filter_pairs``) except in the range form
(one order or equality between the build column and a row operand, on a
multi-match build), which orders a probe row's build rows by the build
operand; ``candidates`` is the unfused pair count.  The range form is
checked for every relation, both operand orders, int and float columns
(NaN, -0.0, infinities, integer against float), literals, and buckets
past the keyed sort and the LDS sort.  Programs naming more columns than
the join carries (three a side: the rest read at their rows), payloads
gathered for kept pairs, build buckets past the LDS chunk and probe
buckets of several work items, and a column shorter than its side
(raises) are covered."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from pageflow_amd.join import HashJoin, hash_join
from pageflow_amd.operators import residual
from residual_cases import condition, sides

pytestmark = pytest.mark.gpu


def _dev(side):
    return {k: torch.from_numpy(v).cuda() for k, v in side.items()}


def _keys(rng, n, distinct, unique):
    if unique:
        return rng.permutation(n).astype(np.int64) * 7 - 1000
    return rng.integers(0, distinct, n).astype(np.int64) * 3 + 11


def _fused(bk, pk, prog, unique, bits, payloads=()):
    plan = HashJoin(bk.numel(), pk.numel(), unique=unique, bits=bits, device=bk.device)
    kept = plan.prepare(bk, pk, residual=prog)
    op = torch.empty(kept, dtype=torch.int64, device=bk.device)
    ob = torch.empty(kept, dtype=torch.int64, device=bk.device)
    outs = [torch.empty(kept, dtype=torch.int64, device=bk.device) for _ in payloads]
    assert plan.write(op, ob, list(payloads), outs) == kept
    return plan, ob, op, outs


def _sorted_pairs(b, p):
    if len(b) == 0:
        return []
    return sorted(zip(np.asarray(b).tolist(), np.asarray(p).tolist()))


@pytest.mark.parametrize("unique", [True, False])
@pytest.mark.parametrize("bits", [None, 3, 9])
def test_fused_matches_oracle_on_random_conditions(unique, bits):
    rng = np.random.default_rng(101 + (bits or 0) + 7 * unique)
    nb, np_ = 3000, 5000
    left, right = sides(rng, nb, np_)  # odd ids: build columns; even ids: probe columns
    bk = _keys(rng, nb, 700, unique)
    pk = np.where(rng.random(np_) < 0.8, bk[rng.integers(0, nb, np_)], 10 ** 9)
    dbk, dpk = torch.from_numpy(bk).cuda(), torch.from_numpy(pk).cuda()
    dl, dr = _dev(left), _dev(right)
    lm, rm = oracle.sort_search_pairs(bk, pk)
    ran = 0
    for _ in range(30):
        conds = [condition(rng) for _ in range(int(rng.integers(1, 4)))]
        try:
            prog = residual.compile_conditions(dl, dr, conds)
        except NotImplementedError:
            continue
        wl, wr = oracle.residual_filter(left, right, conds, lm, rm)
        plan, ob, op, _ = _fused(dbk, dpk, prog, unique, bits)
        assert plan.candidates == lm.shape[0], conds
        assert _sorted_pairs(ob.cpu(), op.cpu()) == _sorted_pairs(wl, wr), conds
        ran += 1
    assert ran >= 10


def _pair_keys(b, p, n_probe):
    return torch.sort(b * n_probe + p).values


@pytest.mark.parametrize("unique", [True, False])
@pytest.mark.parametrize("bits", [2, 10])
@pytest.mark.parametrize("op", ['<', '<>'])
def test_fused_equals_the_filtered_join_in_order(unique, bits, op):
    rng = np.random.default_rng(5 + bits + unique)
    nb, np_ = 20_000, 150_000
    bk = torch.from_numpy(_keys(rng, nb, 5000, unique)).cuda()
    pk = bk[torch.from_numpy(rng.integers(0, nb, np_)).cuda()]
    pk[::9] = -5
    bv = torch.from_numpy(rng.integers(0, 100, nb)).cuda()
    pv = torch.from_numpy(rng.integers(0, 100, np_).astype(np.int32)).cuda()
    prog = residual.compile_conditions({1: bv}, {2: pv}, [['#1', op, '#2']])
    op_, ob, _ = hash_join(bk, pk, unique=unique, bits=bits)
    kb, kp = residual.filter_pairs(prog, ob, op_)
    plan, fb, fp, _ = _fused(bk, pk, prog, unique, bits)
    assert plan.candidates == ob.numel()
    if unique or op == '<>':
        assert torch.equal(fb, kb) and torch.equal(fp, kp)
    else:  # the range form: a probe row's build rows by the build operand
        assert torch.equal(_pair_keys(fb, fp, np_), _pair_keys(kb, kp, np_))


RANGE_OPS = ['=', '<', '<=', '>', '>=', '<>']


def _range_columns(rng, nb, np_, kind):
    if kind == "int":
        return rng.integers(-5, 6, nb), rng.integers(-5, 6, np_).astype(np.int32)
    if kind == "float":
        b = rng.integers(-3, 4, nb).astype(np.float64) * 0.5
        p = rng.integers(-3, 4, np_).astype(np.float64) * 0.5
        for x, n in ((b, nb), (p, np_)):
            x[rng.random(n) < 0.03] = np.nan
            x[rng.random(n) < 0.03] = -0.0
            x[rng.random(n) < 0.02] = np.inf
            x[rng.random(n) < 0.02] = -np.inf
        return b, p
    # an integer build column against a float probe column
    return rng.integers(-3, 4, nb), rng.integers(-6, 7, np_).astype(np.float64) * 0.5


@pytest.mark.parametrize("kind", ["int", "float", "int_vs_float"])
@pytest.mark.parametrize("build_left", [True, False])
def test_range_form_every_relation(kind, build_left):
    rng = np.random.default_rng(31 + len(kind) + build_left)
    nb, np_ = 8000, 60_000
    bk = _keys(rng, nb, 1500, False)
    pk = np.where(rng.random(np_) < 0.85, bk[rng.integers(0, nb, np_)], -77)
    bv, pv = _range_columns(rng, nb, np_, kind)
    left, right = {1: bv}, {2: pv}
    dbk, dpk = torch.from_numpy(bk).cuda(), torch.from_numpy(pk).cuda()
    dl, dr = _dev(left), _dev(right)
    lm, rm = oracle.sort_search_pairs(bk, pk)
    for op in RANGE_OPS:
        if kind != "int" and op == '=':
            continue  # a float '=' is isclose, which the range form does not take
        cond = [['#1', op, '#2'] if build_left else ['#2', op, '#1']]
        try:
            prog = residual.compile_conditions(dl, dr, cond)
        except NotImplementedError:
            continue
        wl, wr = oracle.residual_filter(left, right, cond, lm, rm)
        plan, ob, op_, _ = _fused(dbk, dpk, prog, False, None)
        assert plan.candidates == lm.shape[0]
        assert _sorted_pairs(ob.cpu(), op_.cpu()) == _sorted_pairs(wl, wr), (kind, cond)
    for lit in ['2', '-1.5']:
        cond = [['#1', '<=', lit]]
        try:
            prog = residual.compile_conditions(dl, dr, cond)
        except NotImplementedError:
            continue
        wl, wr = oracle.residual_filter(left, right, cond, lm, rm)
        _, ob, op_, _ = _fused(dbk, dpk, prog, False, None)
        assert _sorted_pairs(ob.cpu(), op_.cpu()) == _sorted_pairs(wl, wr), (kind, cond)


@pytest.mark.parametrize("unique", [True, False])
def test_three_columns_a_side_and_payloads(unique):
    """Two probe columns ride the partition and the third is read at its
    row; the first build column is staged in LDS, the second read in
    partition order, the third at its row.  Payloads come from kept pairs."""
    rng = np.random.default_rng(77 + unique)
    nb, np_ = 6000, 40_000
    bk = _keys(rng, nb, 1500, unique)
    pk = bk[rng.integers(0, nb, np_)]
    left = {1: rng.integers(0, 9, nb), 3: rng.integers(-3, 3, nb).astype(np.int16),
            5: rng.normal(size=nb)}
    right = {2: rng.integers(0, 9, np_).astype(np.int8), 4: rng.integers(-3, 3, np_),
             6: rng.normal(size=np_)}
    conds = [['#1', '<=', '#2'], [['#3', '<>', '#4'], 'OR', ['#5', '>', '#6']]]
    dl, dr = _dev(left), _dev(right)
    prog = residual.compile_conditions(dl, dr, conds)
    pay = torch.from_numpy(rng.integers(-(1 << 40), 1 << 40, nb)).cuda()
    dbk, dpk = torch.from_numpy(bk).cuda(), torch.from_numpy(pk).cuda()
    plan, ob, op, (po,) = _fused(dbk, dpk, prog, unique, None, [pay])
    lm, rm = oracle.sort_search_pairs(bk, pk)
    wl, wr = oracle.residual_filter(left, right, conds, lm, rm)
    assert _sorted_pairs(ob.cpu(), op.cpu()) == _sorted_pairs(wl, wr)
    assert torch.equal(po, pay[ob])


def test_large_build_buckets_and_many_items():
    """A hot key of 9000 build rows (past the LDS chunk and the LDS sort) and
    probe buckets of several 64K-row work items."""
    rng = np.random.default_rng(3)
    nb, np_ = 30_000, 600_000
    bk = rng.integers(0, 4000, nb).astype(np.int64)
    bk[:9000] = 17
    rng.shuffle(bk)
    pk = rng.integers(0, 4000, np_).astype(np.int64)
    pk[rng.random(np_) < 0.02] = 17
    bv = rng.integers(0, 50, nb)
    pv = rng.integers(0, 50, np_)
    dbk, dpk = torch.from_numpy(bk).cuda(), torch.from_numpy(pk).cuda()
    dbv, dpv = torch.from_numpy(bv).cuda(), torch.from_numpy(pv).cuda()
    prog = residual.compile_conditions({1: dbv}, {2: dpv}, [['#1', '>', '#2']])
    op, ob, _ = hash_join(dbk, dpk, unique=False, bits=2)
    kb, kp = residual.filter_pairs(prog, ob, op)
    for bits in (2, 4):  # 4: buckets of both sizes, narrowed and compared per candidate
        plan, fb, fp, _ = _fused(dbk, dpk, prog, False, bits)
        assert plan.candidates == ob.numel() and plan.n_items > 4
        assert torch.equal(_pair_keys(fb, fp, np_), _pair_keys(kb, kp, np_))
        keep = bv[fb.cpu().numpy()] > pv[fp.cpu().numpy()]
        assert keep.all() and bool((dbk[fb] == dpk[fp]).all())


def test_empty_sides_and_nothing_kept():
    bk = torch.arange(100, device="cuda", dtype=torch.int64)
    v = torch.zeros(100, device="cuda", dtype=torch.int64)
    prog = residual.compile_conditions({1: v}, {2: v}, [['#1', '<', '#2']])
    plan, ob, op, _ = _fused(bk, bk, prog, True, None)
    assert ob.numel() == 0 and plan.candidates == 100
    e = torch.empty(0, device="cuda", dtype=torch.int64)
    prog = residual.compile_conditions({1: e}, {2: v}, [['#1', '<', '#2']])
    plan, ob, op, _ = _fused(e, bk, prog, False, None)
    assert ob.numel() == 0 and plan.candidates == 0


def test_a_column_shorter_than_its_side_raises():
    bk = torch.arange(100, device="cuda", dtype=torch.int64)
    short = torch.zeros(99, device="cuda", dtype=torch.int64)
    full = torch.zeros(100, device="cuda", dtype=torch.int64)
    prog = residual.compile_conditions({1: short}, {2: full}, [['#1', '<', '#2']])
    plan = HashJoin(100, 100, unique=True, device="cuda")
    with pytest.raises(RuntimeError, match="fewer than its side"):
        plan.prepare(bk, bk, residual=prog)
    with pytest.raises(ValueError, match="ordered"):
        HashJoin(100, 100, unique=True, device="cuda", ordered=False).prepare(
            bk, bk, residual=residual.compile_conditions({1: full}, {2: full},
                                                         [['#1', '<', '#2']]))
