"""join_kernel_new's dispatch and tails against dense references and the
known answers of the reference's tests (tests/operators/test_join_core.py,
tests/operators/test_unique_hash_join.py of the reference repository).
Semi/anti outputs are ordered and compared exactly; inner and outer pairs
as multisets."""
from __future__ import annotations

import pytest
import torch

import pageflow_amd.operators.hashjoin as hashjoin
import pageflow_amd.operators.join_core as join_core
from pageflow_amd.operators.hashjoin import (_residual_summary, join_kernel_new,
                                             set_join_core_mode)

pytestmark = pytest.mark.gpu

NULL = -64
I64_MIN = -(1 << 63)
_OPS = {'<>': torch.ne, '<': torch.lt, '<=': torch.le, '>': torch.gt, '>=': torch.ge}


def _cuda(v):
    return torch.tensor(v, dtype=torch.int64, device="cuda")


def _rand(n, lo, hi, seed):
    g = torch.Generator().manual_seed(seed)
    return torch.randint(lo, hi, (n,), generator=g, dtype=torch.int64).cuda()


def _bucket_of(keys, B):
    import oracle
    return torch.from_numpy(oracle.bucket_of(keys.cpu().numpy(), B).astype("int64")).cuda()


@pytest.fixture
def calls(monkeypatch):
    seen = []
    for name in ("expand", "semi_anti", "unique_hash_pairs"):
        real = getattr(join_core, name)

        def counted(*a, _n=name, _r=real, **k):
            seen.append(_n)
            return _r(*a, **k)
        monkeypatch.setattr(join_core, name, counted)
    return seen


@pytest.fixture
def fallbacks(monkeypatch):
    reasons = []

    def record(reason):
        reasons.append(reason)
        return None
    monkeypatch.setattr(hashjoin, "_join_core_fallback", record)
    return reasons


def _tick(keys, B):
    b = _bucket_of(keys, B)
    order = torch.argsort(b, stable=True)
    return order, torch.bincount(b, minlength=B).tolist()


def _join(join_type, lk, rk, lv=None, rv=None, condition=None, B=None):
    lgi = torch.arange(lk.numel(), device="cuda") + 1000
    rgi = torch.arange(rk.numel(), device="cuda") + 5000
    bucket_counts = None
    if B is not None:
        lo, lc = _tick(lk, B)
        ro, rc = _tick(rk, B)
        lk, lgi, rk, rgi = lk[lo], lgi[lo], rk[ro], rgi[ro]
        if lv is not None:
            lv, rv = lv[lo], rv[ro]
        bucket_counts = (lc, rc)
    left, right = {1: lk}, {2: rk}
    if lv is not None:
        left[3], right[4] = lv, rv
    out = join_kernel_new(left, right, 1, 2, [condition] if condition else [], [], True,
                          join_type, lgi, rgi, bucket_counts=bucket_counts,
                          equality_keys=([1], [2]), unique_key_sets=([], []),
                          hash_scratch_limit_bytes=1 << 30)
    return out[0], out[1], (lk, rk, lv, rv, lgi, rgi)


def _reference(join_type, op, data):
    lk, rk, lv, rv, lgi, rgi = data
    hit = lk[:, None] == rk[None, :]
    if op is not None:
        hit &= _OPS[op](lv[:, None], rv[None, :])
    if join_type in ('right-semi', 'right-anti'):
        anyh = hit.any(dim=0) if lk.numel() else torch.zeros(rk.numel(), dtype=torch.bool,
                                                             device="cuda")
        keep = anyh if join_type == 'right-semi' else ~anyh
        return [], rgi[keep].tolist()
    li, ri = torch.nonzero(hit, as_tuple=True)
    pairs = list(zip(lgi[li].tolist(), rgi[ri].tolist()))
    if join_type == 'right-outer':
        matched = hit.any(dim=0) if lk.numel() else torch.zeros(rk.numel(), dtype=torch.bool,
                                                                device="cuda")
        pairs += [(NULL, r) for r in rgi[~matched].tolist()]
    return sorted(pairs)


def _check(join_type, op, got_l, got_r, data):
    ref = _reference(join_type, op, data)
    if join_type in ('right-semi', 'right-anti'):
        assert got_l.numel() == 0
        assert got_r.tolist() == ref[1]
    else:
        assert sorted(zip(got_l.tolist(), got_r.tolist())) == ref


_CASES = [('inner', None, None), ('inner', None, 16), ('inner', '<>', 16), ('inner', '<', None),
          ('right-semi', None, None), ('right-semi', None, 16), ('right-semi', '<>', 16),
          ('right-semi', '>=', 8), ('right-anti', None, 16), ('right-anti', '<>', 16),
          ('right-anti', '<=', None), ('right-outer', None, 16), ('right-outer', None, None)]


@pytest.mark.parametrize("join_type,op,B", _CASES)
@pytest.mark.parametrize("n_left,n_right", [(0, 9), (9, 0), (1, 1), (1500, 900)])
def test_join_kernel_matches_reference(join_type, op, B, n_left, n_right):
    seed = n_left + 3 * n_right
    lk = _rand(n_left, 0, max(n_left // 4, 2), seed)
    rk = _rand(n_right, 0, max(n_left // 2, 3), seed + 1)
    lv = _rand(n_left, 0, 3, seed + 2)
    rv = _rand(n_right, 0, 3, seed + 3)
    cond = ['#3', op, '#4'] if op is not None else None
    got_l, got_r, data = _join(join_type, lk, rk, lv, rv, cond, B)
    _check(join_type, op, got_l, got_r, data)


_Q21 = ['NOT', ['l_suppkey#3', '=', 'l_suppkey#4']]


@pytest.mark.parametrize("join_type", ['right-semi', 'right-anti'])
def test_semi_anti_take_the_hash_set(calls, join_type):
    lk, rk = _rand(500, 0, 100, 41), _rand(400, 0, 200, 42)
    lv, rv = _rand(500, 0, 3, 43), _rand(400, 0, 3, 44)
    got_l, got_r, data = _join(join_type, lk, rk, B=16)
    _check(join_type, None, got_l, got_r, data)
    got_l, got_r, data = _join(join_type, lk, rk, lv, rv, _Q21, B=16)
    _check(join_type, '<>', got_l, got_r, data)
    assert calls == ['semi_anti', 'semi_anti']


@pytest.mark.parametrize("join_type", ['right-semi', 'right-anti'])
@pytest.mark.parametrize("residual", [False, True])
def test_empty_slot_build_key_takes_the_pairs(calls, join_type, residual):
    lk = torch.cat((_rand(300, 0, 60, 71), _cuda([I64_MIN, I64_MIN])))
    rk = torch.cat((_rand(200, 0, 120, 72), _cuda([I64_MIN])))
    lv, rv = _rand(302, 0, 3, 73), _rand(201, 0, 3, 74)
    got_l, got_r, data = _join(join_type, lk, rk, lv, rv, _Q21 if residual else None)
    assert calls == ['semi_anti']  # then the pair path (the LDS-table join)
    _check(join_type, '<>' if residual else None, got_l, got_r, data)


def test_residual_summary_parsing():
    left = {1: _cuda([1]), 3: _cuda([7])}
    right = {2: _cuda([1]), 4: _cuda([8])}

    def op_of(cond):
        found = _residual_summary(cond, left, right)
        return None if found is None else found[2]
    assert op_of(['NOT', ['l_suppkey#3', '=', 'l_suppkey#4']]) == '<>'
    assert op_of([['NOT', [['#3'], '=', ['#4']]]]) == '<>'
    assert op_of(['#3', '<', '#4']) == '<'
    assert op_of(['#4', '<', '#3']) == '>'
    assert op_of(['#4', '>=', '#3']) == '<='
    assert op_of(['NOT', ['#3', '<', '#4']]) == '>='
    assert op_of(['#3', '=', '#4']) is None
    assert op_of(['NOT', ['#3', '<>', '#4']]) is None
    assert op_of(['#3', '<', '5']) is None
    assert op_of(['#1', '<', '#3']) is None


@pytest.fixture
def hash_calls(monkeypatch):
    seen = []
    real = join_core.unique_hash_pairs

    def counted(build_keys, probe_keys):
        seen.append((len(build_keys), build_keys[0].numel()))
        return real(build_keys, probe_keys)
    monkeypatch.setattr(join_core, "unique_hash_pairs", counted)
    return seen


def _run(lk, rk, *, left_unique, right_unique, ls=None, rs=None, scratch=1 << 30, conds=None):
    left, right = {1: lk}, {2: rk}
    le, re_ = [1], [2]
    if ls is not None:
        left[3], right[4] = ls, rs
        le.append(3)
        re_.append(4)
    lgi = torch.arange(lk.numel(), dtype=torch.int64, device="cuda")
    rgi = torch.arange(rk.numel(), dtype=torch.int64, device="cuda")
    return join_kernel_new(left, right, 1, 2, conds or [], [], True, "inner", lgi, rgi,
                           equality_keys=(le, re_), unique_key_sets=(left_unique, right_unique),
                           hash_scratch_limit_bytes=scratch)


def test_orientation_for_both_build_roles(hash_calls):
    r = _run(_cuda([10, 20, 30]), _cuda([30, 10, 10, 99]), left_unique=[[1]], right_unique=[])
    assert r[0].tolist() == [2, 0, 0] and r[1].tolist() == [0, 1, 2]
    r = _run(_cuda([5, 5, 8, 99]), _cuda([5, 8]), left_unique=[], right_unique=[[2]])
    assert r[0].tolist() == [0, 1, 2] and r[1].tolist() == [0, 0, 1]
    assert hash_calls == [(1, 3), (1, 2)]


def test_composite_key_hashes_the_whole_tuple(hash_calls):
    r = _run(_cuda([1, 1, 2, 2]), _cuda([2, 1, 2, 1, 3]), left_unique=[[1, 3]], right_unique=[],
             ls=_cuda([7, 8, 7, 8]), rs=_cuda([8, 8, 9, 7, 7]), conds=[["#3", "=", "#4"]])
    assert r[0].tolist() == [3, 1, 0] and r[1].tolist() == [0, 1, 3]
    assert [k for k, _ in hash_calls] == [2]


def test_guards_fall_back(fallbacks):
    set_join_core_mode("unique_hash")
    large, small = _cuda(list(range(9))), _cuda([4])
    r = _run(large, small, left_unique=[[1]], right_unique=[])
    assert r[0].tolist() == [4] and r[1].tolist() == [0]
    left, right = _cuda([1, 2, 3]), _cuda([3, 1, 9])
    r = _run(left, right, left_unique=[], right_unique=[])
    assert sorted(zip(r[0].tolist(), r[1].tolist())) == [(0, 1), (2, 0)]
    r = _run(left, right, left_unique=[[1]], right_unique=[], scratch=1)
    assert sorted(zip(r[0].tolist(), r[1].tolist())) == [(0, 1), (2, 0)]
    assert fallbacks == ["build_probe_ratio_guard", "selected_keys_not_proven_unique",
                         "scratch_guard"]


def test_true_residual_and_materialized_columns():
    left = {1: _cuda([1, 2, 3]), 5: _cuda([10, 20, 30])}
    right = {2: _cuda([1, 2, 3]), 6: _cuda([15, 15, 35])}
    res = join_kernel_new(left, right, 1, 2, [["#5", ">", "#6"]], [5, 6], True, "inner",
                          _cuda([100, 101, 102]), _cuda([200, 201, 202]),
                          equality_keys=([1], [2]), unique_key_sets=([[1]], []),
                          hash_scratch_limit_bytes=1 << 30)
    assert [t.tolist() for t in res] == [[20], [15], [101], [201]]


def test_right_outer_gathers_the_null_row_for_two_d_payloads():
    left = {1: _cuda([1, 2]), 7: torch.tensor([[1, 2, 3], [4, 5, 6]], dtype=torch.int8,
                                              device="cuda")}
    right = {2: _cuda([2, 9])}
    res = join_kernel_new(left, right, 1, 2, [], [7], False, "right-outer", _cuda([0, 1]),
                          _cuda([0, 1]), equality_keys=([1], [2]), unique_key_sets=([], []))
    assert res[0].tolist() == [[4, 5, 6], [-64, -64, -64]]
