"""``join_kernel_new``: which route a call takes and what it returns.

Results are checked against a brute force over every (left row, right row)
in numpy: semi/anti rows exactly (they come ascending), inner and outer
pairs as sorted multisets (the reference's sort/search core orders a probe
row's matches by an unstable sort).  The route checks follow the known
answers of the reference's tests (tests/operators/test_join_core.py,
tests/operators/test_unique_hash_join.py): orientation, composite keys, the
guard a refusal names, the hashed equality not re-evaluated.  Beyond them:
keys of int32, int16 and float64 (including -0.0 and NaN) and residuals the
device program does not compile (strings, IN, arithmetic), shaped like
TPC-H Q7's and Q19's join conditions.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import pageflow_amd.operators.hashjoin as hashjoin
import pageflow_amd.operators.join_core as join_core
from pageflow_amd.operators.hashjoin import (_residual_summary, join_kernel_new,
                                             set_join_core_mode)

pytestmark = pytest.mark.gpu

NULL = -64
I64_MIN = -(1 << 63)
CMP = {'<>': np.not_equal, '<': np.less, '<=': np.less_equal, '>': np.greater,
       '>=': np.greater_equal}


def _dev(values, dtype=torch.int64):
    return torch.as_tensor(np.asarray(values), dtype=dtype).cuda()


def _ints(n, lo, hi, seed):
    return _dev(np.random.default_rng(seed).integers(lo, hi, n, dtype=np.int64))


class Spy:
    """Records the calls of module functions by name, passing them through."""

    def __init__(self, monkeypatch, module, names, record=lambda name, a, k: name):
        self.calls = []
        for name in names:
            real = getattr(module, name)

            def wrapper(*a, _name=name, _real=real, **k):
                self.calls.append(record(_name, a, k))
                return _real(*a, **k)
            monkeypatch.setattr(module, name, wrapper)


@pytest.fixture
def core_calls(monkeypatch):
    return Spy(monkeypatch, join_core, ("expand", "semi_anti", "unique_hash_pairs")).calls


@pytest.fixture
def hash_builds(monkeypatch):
    """(key columns, build rows) of each unique-hash call."""
    return Spy(monkeypatch, join_core, ("unique_hash_pairs",),
               lambda name, a, k: (len(a[0]), a[0][0].numel())).calls


@pytest.fixture
def residual_trees(monkeypatch):
    """The condition tree of each residual evaluation."""
    return Spy(monkeypatch, hashjoin, ("evaluate_join",), lambda name, a, k: a[3]).calls


@pytest.fixture
def refusals(monkeypatch):
    seen = []

    def note(reason):
        seen.append(reason)
        return None
    monkeypatch.setattr(hashjoin, "_join_core_fallback", note)
    return seen


# -------------------------------------------------------- brute force --

def _brute(join_type, lk, rk, keep_pair=None):
    """Expected result rows: for inner/outer a sorted list of (left, right)
    (left -64 for an unmatched right row of an outer join), for semi/anti
    the ascending right rows.  ``keep_pair(l, r)``: the residual."""
    lk, rk = np.asarray(lk), np.asarray(rk)
    hit = lk[:, None] == rk[None, :]
    if keep_pair is not None and hit.any():
        li, ri = np.nonzero(hit)
        hit[li, ri] = [keep_pair(int(a), int(b)) for a, b in zip(li, ri)]
    matched = hit.any(axis=0) if lk.size else np.zeros(rk.size, bool)
    if join_type == 'right-semi':
        return np.nonzero(matched)[0].tolist()
    if join_type == 'right-anti':
        return np.nonzero(~matched)[0].tolist()
    pairs = [(int(a), int(b)) for a, b in zip(*np.nonzero(hit))]
    if join_type == 'right-outer':
        pairs += [(NULL, int(r)) for r in np.nonzero(~matched)[0]]
    return sorted(pairs)


def _result_rows(join_type, out):
    """The join's returned global indexes as _brute's shape (the global
    index of row i is i + 1000 on the left and i + 5000 on the right)."""
    left = [x - 1000 if x != NULL else NULL for x in out[-2].tolist()]
    right = [x - 5000 for x in out[-1].tolist()]
    if join_type in ('right-semi', 'right-anti'):
        assert not left
        return right
    return sorted(zip(left, right))


def _call(join_type, left, right, conditions=(), bucket_counts=None, unique=([], []),
          lk=1, rk=2, **kw):
    lgi = torch.arange(left[lk].shape[0], device="cuda") + 1000
    rgi = torch.arange(right[rk].shape[0], device="cuda") + 5000
    return join_kernel_new(left, right, lk, rk, list(conditions), [], True, join_type, lgi,
                           rgi, bucket_counts=bucket_counts, equality_keys=([lk], [rk]),
                           unique_key_sets=unique, hash_scratch_limit_bytes=1 << 30, **kw)


def _bucketed(keys, B):
    """A tick's rows in bucket order and its bucket counts."""
    import oracle
    b = oracle.bucket_of(keys.cpu().numpy(), B)
    order = np.argsort(b, kind="stable")
    return torch.from_numpy(order).cuda(), np.bincount(b, minlength=B).tolist()


# --------------------------------------------------- every route, dense --

ROUTE_CASES = [('inner', None, None), ('inner', None, 16), ('inner', '<>', 16),
               ('inner', '<', None), ('right-semi', None, None), ('right-semi', None, 16),
               ('right-semi', '<>', 16), ('right-semi', '>=', 8), ('right-anti', None, 16),
               ('right-anti', '<>', 16), ('right-anti', '<=', None), ('right-outer', None, 16),
               ('right-outer', None, None)]


@pytest.mark.parametrize("join_type,op,B", ROUTE_CASES)
@pytest.mark.parametrize("n_left,n_right", [(0, 9), (9, 0), (1, 1), (1500, 900)])
def test_every_join_type_against_the_brute_force(join_type, op, B, n_left, n_right):
    seed = 11 * n_left + n_right
    lk = _ints(n_left, 0, max(n_left // 4, 2), seed)
    rk = _ints(n_right, 0, max(n_left // 2, 3), seed + 1)
    lv, rv = _ints(n_left, 0, 3, seed + 2), _ints(n_right, 0, 3, seed + 3)
    counts = None
    if B is not None:
        lo, lc = _bucketed(lk, B)
        ro, rc = _bucketed(rk, B)
        lk, lv, rk, rv = lk[lo], lv[lo], rk[ro], rv[ro]
        counts = (lc, rc)
    cond = [['#3', op, '#4']] if op else []
    out = _call(join_type, {1: lk, 3: lv}, {2: rk, 4: rv}, cond, counts)
    lvn, rvn = lv.cpu().numpy(), rv.cpu().numpy()
    keep = (lambda a, b: bool(CMP[op](lvn[a], rvn[b]))) if op else None
    assert _result_rows(join_type, out) == _brute(join_type, lk.cpu(), rk.cpu(), keep)


Q21 = ['NOT', ['l_suppkey#3', '=', 'l_suppkey#4']]


@pytest.mark.parametrize("join_type", ['right-semi', 'right-anti'])
def test_semi_and_anti_take_the_hash_set(core_calls, join_type):
    lk, rk = _ints(500, 0, 100, 41), _ints(400, 0, 200, 42)
    lv, rv = _ints(500, 0, 3, 43), _ints(400, 0, 3, 44)
    out = _call(join_type, {1: lk}, {2: rk})
    assert _result_rows(join_type, out) == _brute(join_type, lk.cpu(), rk.cpu())
    out = _call(join_type, {1: lk, 3: lv}, {2: rk, 4: rv}, [Q21])
    lvn, rvn = lv.cpu().numpy(), rv.cpu().numpy()
    assert _result_rows(join_type, out) == _brute(join_type, lk.cpu(), rk.cpu(),
                                                  lambda a, b: lvn[a] != rvn[b])
    assert core_calls == ['semi_anti', 'semi_anti']


@pytest.mark.parametrize("join_type", ['right-semi', 'right-anti'])
@pytest.mark.parametrize("with_residual", [False, True])
def test_an_int64_min_build_key_moves_to_the_pairs(core_calls, join_type, with_residual):
    lk = torch.cat((_ints(300, 0, 60, 71), _dev([I64_MIN, I64_MIN])))
    rk = torch.cat((_ints(200, 0, 120, 72), _dev([I64_MIN])))
    lv, rv = _ints(302, 0, 3, 73), _ints(201, 0, 3, 74)
    out = _call(join_type, {1: lk, 3: lv}, {2: rk, 4: rv}, [Q21] if with_residual else [])
    assert core_calls == ['semi_anti']      # refused (None), then the pair route
    lvn, rvn = lv.cpu().numpy(), rv.cpu().numpy()
    keep = (lambda a, b: lvn[a] != rvn[b]) if with_residual else None
    assert _result_rows(join_type, out) == _brute(join_type, lk.cpu(), rk.cpu(), keep)


def test_the_summary_reads_negations_and_swaps():
    left = {1: _dev([1]), 3: _dev([7])}
    right = {2: _dev([1]), 4: _dev([8])}

    def op(cond):
        s = _residual_summary(cond, left, right)
        return s and s[2]
    assert op(['NOT', ['l_suppkey#3', '=', 'l_suppkey#4']]) == '<>'
    assert op([['NOT', [['#3'], '=', ['#4']]]]) == '<>'
    assert op(['NOT', ['NOT', ['#3', '<', '#4']]]) == '<'
    assert op(['#3', '<', '#4']) == '<'
    assert op(['#4', '<', '#3']) == '>'
    assert op(['#4', '>=', '#3']) == '<='
    assert op(['NOT', ['#3', '<', '#4']]) == '>='
    assert op(['NOT', ['#4', '>', '#3']]) == '>='
    assert not op(['#3', '=', '#4'])
    assert not op(['NOT', ['#3', '<>', '#4']])
    assert not op(['#3', '<', '5'])
    assert not op(['#1', '<', '#3'])


# ------------------------------------------------------- unique route --

def _unique_call(lk, rk, *, left_unique, right_unique, second=None, limit=1 << 30,
                 conds=()):
    left, right = {1: lk}, {2: rk}
    eq = ([1], [2])
    if second is not None:
        left[3], right[4] = second
        eq = ([1, 3], [2, 4])
    return join_kernel_new(left, right, 1, 2, list(conds), [], True, "inner",
                           torch.arange(lk.numel(), device="cuda"),
                           torch.arange(rk.numel(), device="cuda"), equality_keys=eq,
                           unique_key_sets=(left_unique, right_unique),
                           hash_scratch_limit_bytes=limit)


def test_both_build_orientations_return_left_right_pairs(hash_builds):
    out = _unique_call(_dev([10, 20, 30]), _dev([30, 10, 10, 99]), left_unique=[[1]],
                       right_unique=[])
    assert out[0].tolist() == [2, 0, 0] and out[1].tolist() == [0, 1, 2]
    out = _unique_call(_dev([5, 5, 8, 99]), _dev([5, 8]), left_unique=[], right_unique=[[2]])
    assert out[0].tolist() == [0, 1, 2] and out[1].tolist() == [0, 0, 1]
    assert hash_builds == [(1, 3), (1, 2)]


def test_both_sides_unique_builds_the_smaller(hash_builds):
    out = _unique_call(_dev([5, 8, 9, 10]), _dev([5, 8]), left_unique=[[1]],
                       right_unique=[[2]])
    assert out[0].tolist() == [0, 1] and out[1].tolist() == [0, 1]
    assert hash_builds == [(1, 2)]


def test_a_composite_key_is_hashed_whole_and_not_re_evaluated(hash_builds, residual_trees):
    out = _unique_call(_dev([1, 1, 2, 2]), _dev([2, 1, 2, 1, 3]), left_unique=[[1, 3]],
                       right_unique=[], second=(_dev([7, 8, 7, 8]), _dev([8, 8, 9, 7, 7])),
                       conds=[["#3", "=", "#4"]])
    assert out[0].tolist() == [3, 1, 0] and out[1].tolist() == [0, 1, 3]
    assert [k for k, _ in hash_builds] == [2]
    assert residual_trees == []


def test_each_refusal_names_its_guard(refusals):
    set_join_core_mode("unique_hash")
    out = _unique_call(_dev(list(range(9))), _dev([4]), left_unique=[[1]], right_unique=[])
    assert out[0].tolist() == [4] and out[1].tolist() == [0]
    out = _unique_call(_dev([4]), _dev(list(range(9))), left_unique=[], right_unique=[[2]])
    assert out[0].tolist() == [0] and out[1].tolist() == [4]
    a, b = _dev([1, 2, 3]), _dev([3, 1, 9])
    out = _unique_call(a, b, left_unique=[], right_unique=[])
    assert sorted(zip(out[0].tolist(), out[1].tolist())) == [(0, 1), (2, 0)]
    out = _unique_call(a, b, left_unique=[[1]], right_unique=[], limit=1)
    assert sorted(zip(out[0].tolist(), out[1].tolist())) == [(0, 1), (2, 0)]
    assert refusals == ["build_probe_ratio_guard"] * 2 + ["selected_keys_not_proven_unique",
                                                         "scratch_guard"]


def test_the_scratch_guard_uses_the_reference_estimate(refusals):
    # 8 B a build row + 20 B a probe row: 100 build and 100 probe rows are
    # 2800 B, inside a 2800-B limit and outside a 2799-B one
    assert join_core.unique_hash_peak_bytes(100, 100) == 2800
    keys = _dev(np.arange(100))
    _unique_call(keys, keys, left_unique=[[1]], right_unique=[], limit=2800)
    _unique_call(keys, keys, left_unique=[[1]], right_unique=[], limit=2799)
    assert refusals == ["scratch_guard"]


def test_a_true_residual_is_evaluated_once_and_columns_gathered(residual_trees):
    left = {1: _dev([1, 2, 3]), 5: _dev([10, 20, 30])}
    right = {2: _dev([1, 2, 3]), 6: _dev([15, 15, 35])}
    out = join_kernel_new(left, right, 1, 2, [["#5", ">", "#6"]], [5, 6], True, "inner",
                          _dev([100, 101, 102]), _dev([200, 201, 202]),
                          equality_keys=([1], [2]), unique_key_sets=([[1]], []),
                          hash_scratch_limit_bytes=1 << 30)
    assert [t.tolist() for t in out] == [[20], [15], [101], [201]]
    assert residual_trees == [["#5", ">", "#6"]]


def test_right_outer_fills_two_d_payloads_with_null_rows():
    left = {1: _dev([1, 2]), 7: _dev([[1, 2, 3], [4, 5, 6]], torch.int8)}
    out = join_kernel_new(left, {2: _dev([2, 9])}, 1, 2, [], [7], False, "right-outer",
                          _dev([0, 1]), _dev([0, 1]), equality_keys=([1], [2]),
                          unique_key_sets=([], []))
    assert out[0].tolist() == [[4, 5, 6], [-64, -64, -64]]


def test_right_outer_null_rows_of_a_float_payload_are_minus_64():
    left = {1: _dev([1, 2]), 7: _dev([0.5, 2.5], torch.float64)}
    out = join_kernel_new(left, {2: _dev([2, 9])}, 1, 2, [], [7], False, "right-outer",
                          _dev([0, 1]), _dev([0, 1]), equality_keys=([1], [2]),
                          unique_key_sets=([], []))
    assert out[0].tolist() == [2.5, -64.0]


# ------------------------------------------------ keys that are not int64 --

@pytest.mark.parametrize("dtype", [torch.int32, torch.int16])
@pytest.mark.parametrize("join_type", ['inner', 'right-semi', 'right-anti', 'right-outer'])
def test_narrow_integer_keys(dtype, join_type, core_calls):
    lk = _ints(3000, -400, 400, 5).to(dtype)
    rk = _ints(2000, -500, 500, 6).to(dtype)
    out = _call(join_type, {1: lk}, {2: rk}, unique=([[1]], []))
    assert _result_rows(join_type, out) == _brute(join_type, lk.cpu(), rk.cpu())
    assert 'unique_hash_pairs' not in core_calls and 'semi_anti' not in core_calls


@pytest.mark.parametrize("join_type", ['inner', 'right-semi', 'right-anti', 'right-outer'])
def test_float64_keys_with_signed_zeros_and_nan(join_type):
    rng = np.random.default_rng(9)
    lk = rng.integers(0, 50, 700).astype(np.float64) * 0.25
    rk = rng.integers(0, 60, 900).astype(np.float64) * 0.25
    lk[:5] = [0.0, -0.0, np.nan, np.inf, -np.inf]
    rk[:6] = [-0.0, 0.0, np.nan, np.inf, 1e-300, -np.inf]
    out = _call(join_type, {1: _dev(lk, torch.float64)}, {2: _dev(rk, torch.float64)})
    assert _result_rows(join_type, out) == _brute(join_type, lk, rk)


def test_int32_against_float64_keys_compare_as_float64():
    lk = np.array([1, 2, 3, -7, 1 << 30], np.int32)
    rk = np.array([1.0, 2.5, -7.0, 3.0, float(1 << 30), 1.0], np.float64)
    out = _call('inner', {1: _dev(lk, torch.int32)}, {2: _dev(rk, torch.float64)})
    assert _result_rows('inner', out) == _brute('inner', lk.astype(np.float64), rk)


# ------------------------------------- residuals the program cannot compile --

def _text_column(words, width):
    out = np.zeros((len(words), width), np.int8)
    for i, w in enumerate(words):
        out[i, :len(w)] = [ord(c) for c in w]
    return out


def _text(row) -> str:
    return "".join(chr(c) for c in row if c)


NATIONS = ["FRANCE", "GERMANY", "PERU", "CHINA", "KENYA"]


def test_q7_shaped_string_condition(residual_trees):
    rng = np.random.default_rng(17)
    n_left, n_right = 600, 800
    lk, rk = rng.integers(0, 40, n_left), rng.integers(0, 40, n_right)
    n1 = _text_column([NATIONS[i] for i in rng.integers(0, 5, n_left)], 25)
    n2 = _text_column([NATIONS[i] for i in rng.integers(0, 5, n_right)], 25)
    cond = [[['n_name#41', '=', 'FRANCE'], 'AND', ['n_name#45', '=', 'GERMANY']], 'OR',
            [['n_name#41', '=', 'GERMANY'], 'AND', ['n_name#45', '=', 'FRANCE']]]
    out = _call('inner', {1: _dev(lk), 41: _dev(n1, torch.int8)},
                {2: _dev(rk), 45: _dev(n2, torch.int8)}, [cond])

    def keep(a, b):
        x, y = _text(n1[a]), _text(n2[b])
        return (x, y) in (("FRANCE", "GERMANY"), ("GERMANY", "FRANCE"))
    assert _result_rows('inner', out) == _brute('inner', lk, rk, keep)
    assert residual_trees == [cond]


def test_q19_shaped_in_list_and_literals_mixed_with_program_terms():
    rng = np.random.default_rng(19)
    n_left, n_right = 900, 700
    lk, rk = rng.integers(0, 50, n_left), rng.integers(0, 50, n_right)
    brands = ["Brand#12", "Brand#23", "Brand#34"]
    conts = ["SM CASE", "SM BOX", "MED BAG", "LG PKG", "SM PKG"]
    brand = _text_column([brands[i] for i in rng.integers(0, 3, n_right)], 10)
    cont = _text_column([conts[i] for i in rng.integers(0, 5, n_right)], 10)
    qty = rng.integers(1, 40, n_left)
    size = rng.integers(1, 20, n_right).astype(np.int32)
    cond = [[[['p_brand#20', '=', 'Brand#12'], 'AND',
              ['p_container#21', 'IN', ['SM CASE', 'SM BOX', 'SM PACK', 'SM PKG']]], 'AND',
             [['l_quantity#5', '>=', '1'], 'AND', ['l_quantity#5', '<=', '11']]], 'AND',
            [['p_size#22', '<=', '5'], 'AND', ['p_size#22', '>=', '1']]]
    left = {1: _dev(lk), 5: _dev(qty)}
    right = {2: _dev(rk), 20: _dev(brand, torch.int8), 21: _dev(cont, torch.int8),
             22: _dev(size, torch.int32)}
    out = _call('inner', left, right, [cond])

    def keep(a, b):
        return (_text(brand[b]) == "Brand#12"
                and _text(cont[b]) in ("SM CASE", "SM BOX", "SM PACK", "SM PKG")
                and 1 <= qty[a] <= 11 and 1 <= size[b] <= 5)
    assert _result_rows('inner', out) == _brute('inner', lk, rk, keep)


def test_arithmetic_and_numeric_in_residuals():
    rng = np.random.default_rng(23)
    lk, rk = rng.integers(0, 30, 500), rng.integers(0, 30, 600)
    a, b = rng.integers(0, 100, 500), rng.integers(0, 100, 600)
    conds = [[['#3', '+', '#4'], '>', '90'], ['#4', 'IN', ['1', '5', '7', '42']]]
    out = _call('right-semi', {1: _dev(lk), 3: _dev(a)}, {2: _dev(rk), 4: _dev(b)}, conds)
    want = _brute('right-semi', lk, rk,
                  lambda x, y: a[x] + b[y] > 90 and b[y] in (1, 5, 7, 42))
    assert _result_rows('right-semi', out) == want
