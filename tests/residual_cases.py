"""Cases for the residual-condition tests: columns of every dtype the pair
evaluator reads, random condition trees in the reference's parsed form
(string tokens, single-item wrappers, ``[name, '#', id]`` leaves, AS), and
the conditions evaluated with torch on the CPU, step by step as the
reference's evaluator calls torch (operators/operator_utils.py:235-478:
index, squeeze, the comparison ops with torch's scalar casts and
promotions, torch.isclose(atol=1e-5) for a float ``=``, the 1e-5 shift of a
float literal under ``<=`` / ``>=``)."""
from __future__ import annotations

import datetime

import numpy as np
import torch

EPS = 1e-5
_TORCH_OPS = {'<=': torch.le, '>=': torch.ge, '<': torch.lt, '>': torch.gt, '=': torch.eq,
              '<>': torch.ne}
# column id -> dtype; odd ids on the left, even on the right
DTYPES = {1: np.int64, 3: np.int32, 5: np.int16, 7: np.int8, 9: np.uint8, 11: np.float64,
          2: np.int64, 4: np.int32, 6: np.int16, 8: np.int8, 10: np.uint8, 12: np.float64}
INT_IDS = [c for c, d in DTYPES.items() if d != np.float64]
FLOAT_IDS = [11, 12]
DATE = '1995-03-15'  # 1899 days after 1990-01-01


def column(rng, dtype, n):
    if dtype == np.float64:
        base = rng.integers(-3, 4, n).astype(np.float64)
        # values near each other: equal, within and just outside the tolerance
        jitter = rng.choice([0.0, 5e-6, -5e-6, 3e-5, 1e-3, 0.5], n)
        x = base + jitter
        x[rng.random(n) < 0.01] = np.inf
        x[rng.random(n) < 0.01] = -np.inf
        x[rng.random(n) < 0.01] = np.nan
        return x
    info = np.iinfo(dtype)
    x = rng.integers(max(info.min, -4), min(info.max, 4) + 1, n).astype(dtype)
    ext = rng.random(n) < 0.02
    x[ext] = rng.choice(np.array([info.min, info.max], dtype=dtype), int(ext.sum()))
    if dtype in (np.int32, np.int64):
        near = rng.random(n) < 0.05
        x[near] = 1899 + rng.integers(-1, 2, int(near.sum()))
    if dtype == np.int8:
        ch = rng.random(n) < 0.05
        x[ch] = 65
    return x


def sides(rng, n_left_rows, n_right_rows):
    left = {c: column(rng, d, n_left_rows) for c, d in DTYPES.items() if c % 2}
    right = {c: column(rng, d, n_right_rows) for c, d in DTYPES.items() if c % 2 == 0}
    return left, right


def _ref(rng, cid):
    form = rng.integers(0, 4)
    if form == 0:
        return f"c#{cid}"
    if form == 1:
        return f"#{cid}L"
    if form == 2:
        return [f"c", '#', str(cid)]
    return [[f"col_{cid}#{cid}"]]


def _literal(rng, dtype):
    k = rng.integers(0, 8)
    if k <= 2:
        return str(int(rng.integers(-3, 4)))
    if k == 3:
        return str(int(rng.choice([300, -1, 2 ** 40 + 2, -(2 ** 33) + 1, 128, 65535 + 3])))
    if k == 4:
        return DATE
    if k == 5:
        return 'A'
    return str(float(rng.choice([0.5, -1.5, 2.0, 1e-6, 3.00001])))


def comparison(rng, allow_unsupported=False):
    op = str(rng.choice(['=', '<>', '<', '<=', '>', '>=']))
    a = int(rng.choice(list(DTYPES)))
    if rng.random() < 0.5:
        b = int(rng.choice(list(DTYPES)))
        rhs = _ref(rng, b)
    else:
        rhs = _literal(rng, DTYPES[a])
    lhs = _ref(rng, a)
    if allow_unsupported and rng.random() < 0.1:
        lhs, rhs = _literal(rng, None), lhs
    tree = [lhs, op, rhs]
    return [tree] if rng.random() < 0.2 else tree


def condition(rng, depth=0, allow_unsupported=False):
    r = rng.random()
    if depth >= 3 or r < 0.45:
        return comparison(rng, allow_unsupported)
    if r < 0.6:
        return ['NOT', condition(rng, depth + 1, allow_unsupported)]
    if r < 0.65:
        return ['isnotnull', _ref(rng, int(rng.choice(list(DTYPES))))]
    if r < 0.7:
        return [condition(rng, depth + 1, allow_unsupported), 'AS', 'x']
    return [condition(rng, depth + 1, allow_unsupported), str(rng.choice(['AND', 'OR'])),
            condition(rng, depth + 1, allow_unsupported)]


def _norm(t):
    while isinstance(t, list) and len(t) == 1:
        t = t[0]
    return t


def _cid(tok):
    import re
    if not isinstance(tok, str) or tok.startswith('Brand#'):
        return None
    head = tok.split(' as', 1)[0] if ' as' in tok else tok
    m = re.search(r'#(\d+)L?$', head)
    return int(m.group(1)) if m else None


def torch_mask(left, right, tree, lm, rm):
    """The condition on the pairs, computed with torch on the CPU."""
    lt = torch.from_numpy(np.asarray(lm, dtype=np.int64))
    rt = torch.from_numpy(np.asarray(rm, dtype=np.int64))
    n = lt.numel()

    def leaf(cid):
        if cid in left:
            return torch.from_numpy(left[cid])[lt].squeeze()
        return torch.from_numpy(right[cid])[rt].squeeze()

    def ev(t):
        t = _norm(t)
        if isinstance(t, str):
            if len(t) == 10 and t.count('-') == 2 and t.replace('-', '').isdigit():
                y, m, d = (int(x) for x in t.split('-'))
                return (datetime.date(y, m, d) - datetime.date(1990, 1, 1)).days
            if t.replace('.', '').replace('-', '').replace('E', '').isdigit():
                return float(t) if '.' in t else int(t)
            c = _cid(t)
            return leaf(c) if c is not None else t
        if len(t) == 3 and t[1] == '#':
            return leaf(int(t[-1]))
        if len(t) == 2:
            if t[0] == 'NOT':
                return ~ev(t[1])
            if t[0] == 'isnotnull':
                return torch.ones(n, dtype=torch.bool)
            raise ValueError(t[0])
        op = t[-2]
        if op == 'AS':
            return ev(t[0])
        if op in ('AND', 'OR'):
            a, b = ev(t[0]), ev(t[-1])
            return torch.logical_and(a, b) if op == 'AND' else torch.logical_or(a, b)
        lhs, rhs = ev(t[0]), ev(t[-1])
        if isinstance(rhs, float):
            if op == '<=':
                return torch.le(lhs, rhs + EPS)
            if op == '>=':
                return torch.ge(lhs, rhs - EPS)
        if isinstance(rhs, str):
            return _TORCH_OPS[op](lhs, ord(rhs))
        if op == '=' and lhs.dtype == torch.float64:
            if isinstance(rhs, float):
                rhs = torch.full((lhs.shape[0],), rhs, dtype=lhs.dtype)
            return torch.isclose(lhs, rhs, atol=EPS)
        return _TORCH_OPS[op](lhs, rhs)

    out = ev(tree)
    return torch.broadcast_to(out, (n,)).numpy().copy()
