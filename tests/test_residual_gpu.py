"""The residual filter on pairs (pf_pair_filter_*, operators.residual,
hashjoin.evaluate_join / _filter_pairs): masks and kept pairs equal the
oracle's on random condition trees over every column dtype and literal
kind, at pair counts around the 4096-pair tile; the isclose boundary in
float64; pair order kept; at 2^26 pairs the kept and dropped pairs of P and
NOT P partition the pairs and filtering twice changes nothing; pair rows
outside their column raise; programs the reference cannot evaluate, or the
library cannot take, are refused."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

import oracle
from pageflow_amd import _lib
from pageflow_amd.operators import residual
from pageflow_amd.operators.hashjoin import _filter_pairs, evaluate_join
from residual_cases import condition, sides, torch_mask

pytestmark = pytest.mark.gpu


def _dev(side):
    return {k: torch.from_numpy(v).cuda() for k, v in side.items()}


def _oracle_or_none(fn):
    try:
        return fn()
    except (TypeError, RuntimeError, OverflowError, ValueError, NotImplementedError,
            AttributeError):
        return None


@pytest.mark.parametrize("n", [0, 1, 4095, 4096, 4097, 100_003])
def test_filter_matches_oracle_on_random_conditions(n):
    rng = np.random.default_rng(n + 17)
    left, right = sides(rng, 5000, 3000)
    dl, dr = _dev(left), _dev(right)
    lm = rng.integers(0, 5000, n)
    rm = rng.integers(0, 3000, n)
    tl, tr = torch.from_numpy(lm).cuda(), torch.from_numpy(rm).cuda()
    ran = 0
    for _ in range(40):
        conds = [condition(rng) for _ in range(int(rng.integers(1, 4)))]
        want = _oracle_or_none(lambda: oracle.residual_filter(left, right, conds, lm, rm))
        try:
            prog = residual.compile_conditions(dl, dr, conds)
        except NotImplementedError:
            continue
        assert want is not None, conds
        gl, gr = residual.filter_pairs(prog, tl, tr)
        assert np.array_equal(gl.cpu().numpy(), want[0]), conds
        assert np.array_equal(gr.cpu().numpy(), want[1]), conds
        ran += 1
    assert ran >= 15


@pytest.mark.parametrize("seed", range(3))
def test_evaluate_join_mask_matches_oracle(seed):
    """A tree the device program compiles must equal the oracle; one it
    does not (a float against an integer, an integer column against a float
    literal, ...) is evaluated with torch on the GPU and must equal torch's
    own evaluation on the CPU, raising exactly where torch raises."""
    rng = np.random.default_rng(seed)
    left, right = sides(rng, 2000, 2000)
    dl, dr = _dev(left), _dev(right)
    n = 70_001
    lm, rm = rng.integers(0, 2000, n), rng.integers(0, 2000, n)
    tl, tr = torch.from_numpy(lm).cuda(), torch.from_numpy(rm).cuda()
    on_program = on_torch = 0
    for _ in range(40):
        tree = condition(rng)
        try:
            residual.compile_conditions(dl, dr, [tree])
            compiled = True
        except NotImplementedError:
            compiled = False
        if compiled:
            want = _oracle_or_none(lambda: oracle.residual_mask(left, right, tree, lm, rm))
            assert want is not None, tree
            on_program += 1
        else:
            want = _oracle_or_none(lambda: torch_mask(left, right, tree, lm, rm))
            on_torch += 1
        if want is None:
            with pytest.raises((TypeError, RuntimeError, ValueError)):
                evaluate_join("cuda", dl, dr, tree, tl, tr)
            continue
        got = evaluate_join("cuda", dl, dr, tree, tl, tr)
        assert got.dtype == torch.bool
        assert np.array_equal(got.cpu().numpy(), want), tree
    assert on_program >= 15 and on_torch >= 1


@pytest.mark.parametrize("op", ['=', '<>', '<', '<=', '>', '>='])
@pytest.mark.parametrize("a,b", [(1, 2), (3, 10), (7, 8), (9, 4), (11, 12), (11, 2), (1, 12),
                                 (5, 6)])
def test_every_relation_and_column_pair(op, a, b):
    rng = np.random.default_rng([ord(op[0]), len(op), a, b])
    left, right = sides(rng, 1000, 1000)
    n = 50_000
    lm, rm = rng.integers(0, 1000, n), rng.integers(0, 1000, n)
    cond = [f"x#{a}", op, f"y#{b}"]
    want = _oracle_or_none(lambda: oracle.residual_filter(left, right, [cond], lm, rm))
    tl, tr = torch.from_numpy(lm).cuda(), torch.from_numpy(rm).cuda()
    if want is None:  # float = integer column: the reference's isclose refuses it
        with pytest.raises(NotImplementedError):
            residual.compile_conditions(_dev(left), _dev(right), [cond])
        return
    gl, gr = _filter_pairs(_dev(left), _dev(right), [cond], tl, tr)
    assert np.array_equal(gl.cpu().numpy(), want[0])
    assert np.array_equal(gr.cpu().numpy(), want[1])


def test_isclose_boundary_in_float64():
    rng = np.random.default_rng(9)
    b = rng.standard_normal(4000) * 10.0 ** rng.integers(-8, 8, 4000)
    with np.errstate(invalid="ignore", over="ignore"):
        tol = 1e-5 + np.abs(1e-5 * b)
        a = np.concatenate([b + tol, b - tol, np.nextafter(b + tol, np.inf),
                            np.nextafter(b - tol, -np.inf), b * (1 + 2e-5),
                            [np.inf, -np.inf, np.nan, 0.0, 1e-5, 1.1e-5]])
    bb = np.concatenate([b, b, b, b, b, [np.inf, np.inf, np.nan, -0.0, 0.0, 0.0]])
    left, right = {1: a}, {2: bb}
    idx = np.arange(a.shape[0])
    t = torch.from_numpy(idx).cuda()
    for cond in (['#1', '=', '#2'], ['#2', '=', '#1'], ['#1', '<>', '#2'],
                 ['NOT', ['#1', '=', '#2']]):
        got = evaluate_join("cuda", _dev(left), _dev(right), cond, t, t).cpu().numpy()
        assert np.array_equal(got, oracle.residual_mask(left, right, cond, idx, idx)), cond
    # against a literal: isclose(a, full(c)); <= / >= shift by 1e-5
    for c in ('0.5', '-1.5', '2.0'):
        x = float(c)
        vals = np.array([x, x + 1e-5, x + 1.000001e-5, x - 1e-5, x - 2e-5, np.nextafter(x, 9)])
        L = {1: vals}
        k = np.arange(vals.shape[0])
        tk = torch.from_numpy(k).cuda()
        for op in ('=', '<=', '>=', '<', '>', '<>'):
            cond = ['#1', op, c]
            got = evaluate_join("cuda", _dev(L), {}, cond, tk, tk).cpu().numpy()
            assert np.array_equal(got, oracle.residual_mask(L, {}, cond, k, k)), cond


def test_reference_known_answers_through_the_join():
    from pageflow_amd.operators.hashjoin import join_kernel_new

    def cuda(v):
        return torch.tensor(v, dtype=torch.int64, device="cuda")
    # tests/operators/test_unique_hash_join.py:251-272
    left = {1: cuda([1, 2, 3]), 5: cuda([10, 20, 30])}
    right = {2: cuda([1, 2, 3]), 6: cuda([15, 15, 35])}
    res = join_kernel_new(left, right, 1, 2, [["#5", ">", "#6"]], [5, 6], True, "inner",
                          cuda([100, 101, 102]), cuda([200, 201, 202]),
                          equality_keys=([1], [2]), unique_key_sets=([[1]], []),
                          hash_scratch_limit_bytes=1 << 30)
    assert [t.tolist() for t in res] == [[20], [15], [101], [201]]
    # the same through the multi-match path (no uniqueness proof), Q21's form
    res = join_kernel_new({1: cuda([1, 1, 2]), 3: cuda([7, 8, 7])},
                          {2: cuda([1, 2, 1]), 4: cuda([7, 7, 9])}, 1, 2,
                          [['NOT', ['l_suppkey#3', '=', 'l_suppkey#4']]], [3, 4], True, "inner",
                          cuda([0, 1, 2]), cuda([0, 1, 2]), equality_keys=([1], [2]),
                          unique_key_sets=([], []))
    got = sorted(zip(*[t.tolist() for t in res]))
    # pairs with equal keys: (0,0) (1,0) (0,2) (1,2) (2,1); Q21 keeps unequal suppkeys
    want = sorted([(8, 7, 1, 0), (7, 9, 0, 2), (8, 9, 1, 2)])
    assert got == want


def test_pair_order_is_kept_and_large_counts_partition():
    n = 1 << 26
    g = torch.Generator(device="cuda").manual_seed(4)
    lcol = torch.randint(-1000, 1000, (1 << 20,), generator=g, device="cuda")
    rcol = torch.randint(-1000, 1000, (1 << 22,), generator=g, device="cuda")
    lm = torch.randint(0, 1 << 20, (n,), generator=g, device="cuda")
    rm = torch.arange(n, device="cuda") % (1 << 22)
    L, R = {1: lcol}, {2: rcol}
    P = ['#1', '<', '#2']
    keep_l, keep_r = _filter_pairs(L, R, [P], lm, rm)
    drop_l, drop_r = _filter_pairs(L, R, [['NOT', P]], lm, rm)
    assert keep_l.numel() + drop_l.numel() == n
    assert bool((lcol[keep_l] < rcol[keep_r]).all())
    assert bool((lcol[drop_l] >= rcol[drop_r]).all())
    # idempotent: filtering the kept pairs again keeps them all
    again_l, again_r = _filter_pairs(L, R, [P], keep_l, keep_r)
    assert torch.equal(again_l, keep_l) and torch.equal(again_r, keep_r)
    # pair order: with pair positions as the left rows, the kept rows ascend
    pos = torch.arange(n, device="cuda")
    pcol = torch.randint(0, 2, (n,), generator=g, device="cuda")
    kl, kr = _filter_pairs({1: pcol}, R, [['#1', '=', '1']], pos, rm)
    assert kl.numel() == int(pcol.sum())
    assert bool((kl[1:] > kl[:-1]).all())
    assert torch.equal(kr, rm[kl])


def test_rows_outside_their_column_raise():
    L = {1: torch.arange(10, device="cuda")}
    R = {2: torch.arange(5, device="cuda")}
    lm = torch.tensor([0, 3, 10, 2], device="cuda")
    rm = torch.tensor([0, 1, 2, -1], device="cuda")
    with pytest.raises(RuntimeError, match="outside their column"):
        _filter_pairs(L, R, [['#1', '<', '#2']], lm, rm)


def test_refused_programs():
    L = {1: torch.arange(10, device="cuda"), 3: torch.zeros(10, 4, dtype=torch.int8,
                                                            device="cuda"),
         5: torch.zeros(10, dtype=torch.float32, device="cuda"),
         7: torch.zeros(10, dtype=torch.float64, device="cuda")}
    R = {2: torch.arange(10, device="cuda")}
    for cond in (['#1', '<', '1.5'],          # float32 in torch
                 ['#7', '=', '3'],            # isclose of a float and an int
                 ['#7', '=', '#2'],
                 ['3', '<', '#1'],            # a literal on the left
                 ['#3', '=', 'ABC'],          # strings
                 ['#5', '<', '#1'],           # float32 column
                 [['#1', '+', '#2'], '<', '3'],
                 ['#1', 'IN', ['1', '2']]):
        with pytest.raises(NotImplementedError):
            residual.compile_conditions(L, R, [cond])
    with pytest.raises(ValueError, match="Unsupported operation"):
        residual.compile_conditions(L, R, [['#1', '~', '#2']])
    cols = {c: torch.arange(10, device="cuda") for c in range(1, 40, 2)}
    many = [[f'#{c}', '<', '#2'] for c in range(1, 33, 2)]
    residual.compile_conditions(cols, R, many[:15])
    with pytest.raises(NotImplementedError, match="columns"):
        residual.compile_conditions(cols, R, many + [['#33', '<', '#2']])


def _raw_count(ins, types=(0,), sides_=(0,), rows=(4,), n=0):
    lib = _lib.load()
    col = torch.arange(4, device="cuda")
    host = torch.zeros(2, dtype=torch.int64, pin_memory=True)
    ws = _lib.workspace(lib.pf_pair_filter_workspace_bytes(n), col.device)
    flat = [w for t in ins for w in t]
    return lib.pf_pair_filter_count(
        _lib.ptr(col), _lib.ptr(col), n, _lib.void_array([col.data_ptr()] * len(types)),
        ctypes.cast(_lib.i32_array(types), ctypes.c_void_p),
        ctypes.cast(_lib.i32_array(sides_), ctypes.c_void_p),
        ctypes.cast(_lib.i64_array(rows), ctypes.c_void_p), len(types),
        ctypes.cast(_lib.i64_array(flat), ctypes.c_void_p), len(ins), None,
        ctypes.c_void_p(host.data_ptr()), _lib.ptr(ws), ws.numel(), _lib.stream_of(col.device))


def _last_error() -> str:
    return _lib.load().pf_last_error().decode(errors="replace")


def test_c_abi_validates_programs():
    R = residual
    cmp_ = R.CMP | (R.REL['<'] << 8) | R.A_COL | R.B_COL
    assert _raw_count([(cmp_, 0, 0)]) == 0
    bad = {
        "opcode": [(9, 0, 0)],
        "fewer than two": [(cmp_, 0, 0), (R.AND, 0, 0)],
        "not one": [(cmp_, 0, 0), (cmp_, 0, 0)],
        "slot": [(cmp_, 0, 3)],
        "relation": [(R.CMP | (R.REL_CLOSE << 8) | R.A_COL | R.B_COL, 0, 0)],
        "NOT on no result": [(R.NOT, 0, 0)],
    }
    for what, ins in bad.items():
        assert _raw_count(ins) != 0, what
        assert what in _last_error(), (what, _last_error())
    assert _raw_count([(cmp_, 0, 0)], types=(5,)) != 0      # float64 in an int64 compare
    assert "float64" in _last_error()
    deep = [(R.TRUE, 0, 0)] * 33 + [(R.AND, 0, 0)] * 31
    assert _raw_count(deep) != 0 and "deeper" in _last_error()
    ok = [(R.TRUE, 0, 0)] * 32 + [(R.AND, 0, 0)] * 31
    assert _raw_count(ok) == 0


def test_sixteen_columns_and_a_deep_program():
    # the column limit: 16 slots (8 a side, every dtype), and 35 program steps
    rng = np.random.default_rng(21)
    dts = [np.int64, np.int32, np.int16, np.int8, np.uint8, np.float64, np.int64, np.int32]
    left = {2 * j + 1: rng.integers(0, 4, 900).astype(dt) for j, dt in enumerate(dts)}
    right = {2 * j + 2: rng.integers(0, 4, 700).astype(dt) for j, dt in enumerate(dts)}
    right[12] = right[12] + 0.5
    ops = ['=', '<>', '<', '<=', '>', '>=']
    conds = []
    for j in range(8):
        a, b = 2 * j + 1, 2 * ((j + 3) % 8) + 2
        op = ops[j % 6]
        if dts[j] == np.float64 and op == '=':
            op = '<>'
        inner = [f'l#{a}', op, f'r#{b}']
        conds.append(['NOT', [inner, 'OR', [f'l#{a}', '>=', '2']]] if j % 2 else
                     [inner, 'AND', ['isnotnull', f'r#{b}']])
    n = 123_457
    lm, rm = rng.integers(0, 900, n), rng.integers(0, 700, n)
    prog = residual.compile_conditions(_dev(left), _dev(right), conds)
    assert len(prog.cols) == 16 and len(prog.ins) <= residual.MAX_INS
    gl, gr = residual.filter_pairs(prog, torch.from_numpy(lm).cuda(), torch.from_numpy(rm).cuda())
    wl, wr = oracle.residual_filter(left, right, conds, lm, rm)
    assert gl.numel() > 0
    assert np.array_equal(gl.cpu().numpy(), wl) and np.array_equal(gr.cpu().numpy(), wr)
