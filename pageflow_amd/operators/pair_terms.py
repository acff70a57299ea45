"""Residual join conditions the device pair program does not compile,
evaluated on the pairs with torch tensor operations on the GPU.

The pair program (``operators.residual``) covers comparisons of numeric
columns and literals under AND/OR/NOT.  A join's residual can also hold the
other forms of the reference's filter evaluator
(operators/operator_utils.py:235-478, filter mode): string columns (int8
``[rows, width]``, zero padded) compared with multi-character literals,
``IN`` lists, ``LIKE '%a%b%'``, ``StartsWith``/``EndsWith``/``Contains``,
``substring``, arithmetic, ``CASE`` and ``cast(... as double)``.  Spark
plans put such terms in join conditions (TPC-H Q7's ``n_name = 'FRANCE' OR
...``, Q19's ``p_container IN (...)``).  This module evaluates them with
the same rules, on the operand columns gathered at the pair rows by the
library's gather kernel; the result is one bool per pair.

It is the slow form: every operand is materialised per pair.  The join
dispatch uses it only for a condition the program rejects.
"""
from __future__ import annotations

import datetime
from typing import Dict

import torch

from .. import ops
from .residual import EPS, column_ref, normalize_expr_tree

_EPOCH = datetime.date(1990, 1, 1)  # conversion.py:13
STRING_DTYPE = torch.int8           # constants.py:3

_BINARY = {
    '+': torch.add, '-': torch.sub, '*': torch.mul, '/': torch.div,
    '<=': torch.le, '>=': torch.ge, '<': torch.lt, '>': torch.gt,
    '=': torch.eq, '<>': torch.ne, 'AND': torch.logical_and, 'OR': torch.logical_or,
}


def _is_date(s: str) -> bool:
    return len(s) == 10 and s.count('-') == 2 and s.replace('-', '').isdigit()


def _is_number(s: str) -> bool:
    return s.replace('.', '').replace('-', '').replace('E', '').isdigit()


def _days(s: str) -> int:
    y, m, d = (int(x) for x in s.split('-'))
    return (datetime.date(y, m, d) - _EPOCH).days


def _codes(text: str, width: int, device) -> torch.Tensor:
    """A string literal as the column encoding: code points, zero padded
    to ``width`` (conversion.py:32-33)."""
    codes = [ord(c) for c in text] + [0] * (width - len(text))
    return torch.tensor(codes, dtype=STRING_DTYPE, device=device)


def _string_equal(col: torch.Tensor, text: str) -> torch.Tensor:
    col = col if col.dim() == 2 else col.unsqueeze(1)   # a one-character column
    if len(text) > col.shape[1]:
        # longer than the column's width: no value can equal it
        return torch.zeros(col.shape[0], dtype=torch.bool, device=col.device)
    return (col == _codes(text, col.shape[1], col.device)).all(dim=1)


def _contains(col: torch.Tensor, text: str) -> torch.Tensor:
    """Rows whose string holds ``text`` at some position."""
    n, width = col.shape
    pat = _codes(text, len(text), col.device)
    hit = torch.zeros(n, dtype=torch.bool, device=col.device)
    for at in range(width - len(text) + 1):
        hit |= (col[:, at:at + len(text)] == pat).all(dim=1)
    return hit


def _like_two_words(col: torch.Tensor, first: str, second: str) -> torch.Tensor:
    """``LIKE '%first%second%'``: ``first`` ends at or before where
    ``second`` starts."""
    n, width = col.shape
    p1 = _codes(first, len(first), col.device)
    p2 = _codes(second, len(second), col.device)
    seen_first = torch.zeros(n, dtype=torch.bool, device=col.device)
    hit = torch.zeros(n, dtype=torch.bool, device=col.device)
    for at in range(width - len(first) - len(second) + 1):
        seen_first |= (col[:, at:at + len(first)] == p1).all(dim=1)
        second_here = (col[:, at + len(first):at + len(first) + len(second)] == p2).all(dim=1)
        hit |= seen_first & second_here
    return hit


class PairTerms:
    """Evaluates condition trees on the pairs ``(left_rows[i],
    right_rows[i])``.  A column id the left side holds is read on the left
    rows, any other on the right rows (operators/hashjoin.py:235-242)."""

    def __init__(self, tensors_left: Dict, tensors_right: Dict, left_rows: torch.Tensor,
                 right_rows: torch.Tensor):
        self.left, self.right = tensors_left, tensors_right
        self.left_rows, self.right_rows = left_rows, right_rows
        self.n = int(left_rows.numel())
        self.device = left_rows.device
        self._gathered: Dict[int, torch.Tensor] = {}

    def column(self, cid: int) -> torch.Tensor:
        if cid not in self._gathered:
            # a pair row outside its column raises, as indexing does
            if cid in self.left:
                t = ops.gather_rows(self.left[cid], self.left_rows, check=True)
            else:
                t = ops.gather_rows(self.right[cid], self.right_rows, check=True)
            if t.dim() == 2 and t.shape[1] == 1:
                t = t.reshape(-1)
            self._gathered[cid] = t
        return self._gathered[cid]

    def _leaf(self, token: str):
        if token == "Subquery":
            raise NotImplementedError("a subquery result in a join condition")
        if _is_date(token):
            return _days(token)
        if _is_number(token):
            return float(token) if "." in token else int(token)
        cid = column_ref(token)
        if cid is not None:
            return self.column(cid)
        return token

    def value(self, tree):
        tree = normalize_expr_tree(tree)
        if isinstance(tree, str):
            return self._leaf(tree)
        if isinstance(tree, (int, float)) and not isinstance(tree, bool):
            return tree
        if len(tree) == 3 and tree[1] == '#':
            return self.column(int(tree[-1]))
        if len(tree) == 8 and tree[0] == 'CASE':
            return torch.where(self.value(tree[2]), self.value(tree[4]), self.value(tree[6]))
        if len(tree) == 2:
            return self._function(tree[0], tree[1])
        op = tree[-2]
        if op == 'IN':
            return self._in(tree[0], tree[-1])
        if op == 'LIKE':
            pattern = tree[-1]
            if pattern.count('%') != 3:
                raise NotImplementedError(f"LIKE pattern {pattern!r}: only '%a%b%' is read")
            _, first, second, _ = pattern.split('%')
            return _like_two_words(self.value(tree[0]), first, second)
        if op == 'AS':
            return self.value(tree[0])
        if op not in _BINARY:
            raise ValueError(f"Unsupported operation: {op}")
        return self._binary(op, self.value(tree[0]), self.value(tree[-1]))

    def _function(self, fn, arg):
        if fn in ('StartsWith', 'EndsWith', 'Contains'):
            col, text = self.column(column_ref(arg[0])), arg[1]
            if fn == 'StartsWith':
                return (col[:, :len(text)] == _codes(text, len(text), col.device)).all(dim=1)
            return _contains(col, text)
        if fn == 'NOT':
            return ~self.value(arg)
        if fn == 'isnotnull':
            return torch.ones(self.n, dtype=torch.bool, device=self.device)
        if fn == 'cast':
            literal, _, cast_type = arg[0].split(' ')
            if cast_type != 'double':
                raise ValueError(f"Unsupported cast type: {cast_type}")
            v = self.value(literal)
            return v.to(torch.float64) if torch.is_tensor(v) else float(v)
        if fn == 'substring':
            base, pos, width = arg
            start = int(pos) - 1
            return self.value(base)[:, start:start + int(width)]
        raise ValueError(f"Unsupported function: {fn}")

    def _in(self, lhs_tree, names):
        lhs = self.value(lhs_tree)
        hit = torch.zeros(self.n, dtype=torch.bool, device=self.device)
        if lhs.dtype == STRING_DTYPE:
            for name in names:
                hit |= _string_equal(lhs, name)
            return hit
        cast = int if lhs.dtype in (torch.int64, torch.int32) else float
        for name in names:
            hit |= lhs == cast(name)
        return hit

    def _binary(self, op, lhs, rhs):
        if isinstance(rhs, float) and op in ('<=', '>='):
            return torch.le(lhs, rhs + EPS) if op == '<=' else torch.ge(lhs, rhs - EPS)
        if isinstance(rhs, str):
            if len(rhs) == 1:
                return _BINARY[op](lhs, ord(rhs))
            if op != '=':
                raise NotImplementedError(f"string {op} against {rhs!r}")
            return _string_equal(lhs, rhs)
        if op == '=' and torch.is_tensor(lhs) and lhs.dtype == torch.float64:
            # a float literal becomes a tensor; an integer one stays a scalar,
            # which torch.isclose refuses, as it does in the reference
            if isinstance(rhs, float):
                rhs = torch.full((lhs.shape[0],), rhs, dtype=lhs.dtype, device=lhs.device)
            return torch.isclose(lhs, rhs, atol=EPS)
        return _BINARY[op](lhs, rhs)

    def mask(self, tree) -> torch.Tensor:
        """The condition's value on every pair as a bool mask."""
        v = self.value(tree)
        if not torch.is_tensor(v) or v.dtype != torch.bool or v.shape != (self.n,):
            raise ValueError(f"join condition {tree!r} is not a per-pair predicate")
        return v
