"""Residual join conditions on pairs, compiled to one device program.

The reference evaluates a join's residual conditions on its (left row,
right row) pairs with its general expression evaluator: per condition, per
column operand, a gather of the column at the pair rows, a torch comparison,
then ``match_mask &=``, a ``nonzero`` and two index selects
(operators/hashjoin.py:235-242 ``evaluate_join``, :264-290
``_apply_fast_path_conditions``, :428-439 in the sort/search core; the
evaluation rules are operators/operator_utils.py:235-478 ``_evaluate_core``
in filter mode).  Here the whole condition list becomes one postfix program
(``pf_pair_filter_*``) that a kernel evaluates for every pair, gathering the
operands itself, and the kept pairs are compacted in pair order.

The forms compiled, with the reference's semantics:

* column references (``name#id``, ``#id``, ``[name, '#', id]``): read on
  the left rows when the left side holds the id, else on the right rows;
  1-D columns (or ``(n, 1)``) of int64, int32, int16, int8, uint8, float64;
* literals: integers, dates (days since 1990-01-01, conversion.py:24-27),
  floats, one-character strings (their code point);
* ``=``, ``<>``, ``<``, ``<=``, ``>``, ``>=`` (operator_utils.py:438-478):
  integers compare in int64, an integer literal first wrapped to the
  column's dtype as torch casts it; a float64 column makes the comparison
  float64; ``=`` with a float64 left operand is ``torch.isclose(a, b,
  atol=1e-5)`` (rtol 1e-5); ``<=`` / ``>=`` against a float literal compare
  with the literal plus / minus 1e-5;
* ``AND``, ``OR``, ``NOT``, ``isnotnull`` (true), ``AS`` (its operand).

Forms the reference evaluates some other way, or rejects (a literal on the
left, an integer column against a float literal, which torch compares in
float32, float ``=`` against an integer, strings longer than one character,
arithmetic, IN, LIKE, CASE), raise NotImplementedError; an operator the
reference does not know raises ValueError, as it does.
"""
from __future__ import annotations

import ctypes
import datetime
import re
import struct
from typing import Dict, List, Optional, Sequence, Tuple

import torch

from .. import _lib

EPS = 1e-5  # constants.py:8
NULL = -64

# pageflow_join.h
CMP, AND, OR, NOT, TRUE = 1, 2, 3, 4, 5
REL = {'=': 0, '<>': 1, '<': 2, '<=': 3, '>': 4, '>=': 5}
REL_CLOSE = 6
F64, A_COL, B_COL = 1 << 16, 1 << 17, 1 << 18
COL_TYPE = {torch.int64: 0, torch.int32: 1, torch.int16: 2, torch.int8: 3, torch.uint8: 4,
            torch.float64: 5}
MAX_COLS, MAX_INS = 16, 64
_EPOCH = datetime.date(1990, 1, 1)  # conversion.py:13
_COLUMN_REF = re.compile(r'#(\d+)L?$')
_LOGIC = {'AND': AND, 'OR': OR}
_KNOWN_OPS = {'+', '-', '*', '/', '<=', '>=', '<', '>', '=', '<>', 'AND', 'OR'}


def normalize_expr_tree(tree):
    """Unwraps single-item list wrappers (operator_utils.py:22-26)."""
    while isinstance(tree, list) and len(tree) == 1:
        tree = tree[0]
    return tree


def column_ref(token) -> Optional[int]:
    """The column id a token references, or None (parsing/expressions.py:
    17-32): ``name#id`` or ``name#idL``, a cast's ``name#id as double``;
    ``Brand#NN`` is a literal."""
    if not isinstance(token, str) or token.startswith('Brand#'):
        return None
    head = token.split(' as', 1)[0] if ' as' in token else token
    m = _COLUMN_REF.search(head)
    return int(m.group(1)) if m else None


def _is_date(s: str) -> bool:
    return len(s) == 10 and s.count('-') == 2 and s.replace('-', '').isdigit()


def _is_float(s: str) -> bool:
    return s.replace('.', '').replace('-', '').replace('E', '').isdigit()


def _wrap(v: int, dtype: torch.dtype) -> int:
    """v cast to an integer dtype as torch casts a scalar operand (two's
    complement wrap-around)."""
    bits = torch.iinfo(dtype).bits
    v &= (1 << bits) - 1
    if dtype != torch.uint8 and v >> (bits - 1):
        v -= 1 << bits
    return v


def _f64_bits(x: float) -> int:
    return struct.unpack('<q', struct.pack('<d', float(x)))[0]


class Program:
    """A compiled condition list: the column slots and the instructions."""

    def __init__(self, left: Dict, right: Dict):
        self.left, self.right = left, right
        self.cols: List[torch.Tensor] = []
        self.types: List[int] = []
        self.sides: List[int] = []
        self.slot_of: Dict[Tuple[int, int], int] = {}
        self.ins: List[Tuple[int, int, int]] = []

    # -- operands: ('col', slot, dtype) | ('int', v) | ('float', v) | ('str', s)
    def _column(self, cid: int):
        side = 0 if cid in self.left else 1
        t = (self.left if side == 0 else self.right)[cid]
        key = (side, cid)
        if key not in self.slot_of:
            _lib.require_cuda(t, f"residual column #{cid}")
            if t.dim() == 2 and t.shape[1] == 1:
                t = t.reshape(-1)
            if t.dim() != 1:
                raise NotImplementedError(f"residual column #{cid} of shape {tuple(t.shape)}: "
                                          "the pair evaluator compares 1-D columns")
            if t.dtype not in COL_TYPE:
                raise NotImplementedError(f"residual column #{cid} of {t.dtype}: the pair "
                                          "evaluator reads int64/32/16/8, uint8 and float64")
            if len(self.cols) == MAX_COLS:
                raise NotImplementedError(f"more than {MAX_COLS} residual columns")
            self.slot_of[key] = len(self.cols)
            self.cols.append(t.contiguous())
            self.types.append(COL_TYPE[t.dtype])
            self.sides.append(side)
        s = self.slot_of[key]
        return ('col', s, self.cols[s].dtype)

    def _operand(self, tok):
        tok = normalize_expr_tree(tok)
        if isinstance(tok, list) and len(tok) == 3 and tok[1] == '#':
            return self._column(int(tok[-1]))
        if isinstance(tok, list) and len(tok) == 3 and tok[1] == 'AS':
            return self._operand(tok[0])
        if isinstance(tok, bool):
            raise NotImplementedError(f"residual operand {tok!r}")
        if isinstance(tok, int):
            return ('int', tok)
        if isinstance(tok, float):
            return ('float', tok)
        if isinstance(tok, str):
            if _is_date(tok):
                y, m, d = (int(x) for x in tok.split('-'))
                return ('int', (datetime.date(y, m, d) - _EPOCH).days)
            if _is_float(tok):
                return ('float', float(tok)) if '.' in tok else ('int', int(tok))
            cid = column_ref(tok)
            if cid is not None:
                return self._column(cid)
            return ('str', tok)
        raise NotImplementedError(f"residual operand {tok!r} is outside the pair evaluator")

    def _emit(self, w0: int, a: int = 0, b: int = 0) -> None:
        if len(self.ins) == MAX_INS:
            raise NotImplementedError(f"a residual program of more than {MAX_INS} steps")
        self.ins.append((w0, a, b))

    def _cmp(self, tree) -> None:
        """One comparison, in the reference's order of rules
        (operator_utils.py:458-478): a float literal under <= / >= is
        shifted by 1e-5; a one-character string compares as its code point;
        a float64 left operand under = is isclose; otherwise the torch
        comparison."""
        op = tree[1]
        a, b = self._operand(tree[0]), self._operand(tree[2])
        if a[0] != 'col':
            raise NotImplementedError(f"residual {tree!r}: the reference compares a column "
                                      "on the left of a literal")
        a_f = a[2] == torch.float64
        rel = REL[op]
        if b[0] == 'float' and op in ('<=', '>='):
            if not a_f:
                raise NotImplementedError(f"residual {tree!r}: an integer column against a "
                                          "float literal compares in float32")
            c = b[1] + EPS if op == '<=' else b[1] - EPS
            self._emit(CMP | (rel << 8) | F64 | A_COL, a[1], _f64_bits(c))
            return
        if b[0] == 'str':
            if len(b[1]) != 1:
                raise NotImplementedError(f"residual {tree!r}: string comparison")
            self._literal_cmp(tree, rel, a, ord(b[1]))
            return
        if op == '=' and a_f:
            if b[0] == 'float':
                self._emit(CMP | (REL_CLOSE << 8) | F64 | A_COL, a[1], _f64_bits(b[1]))
            elif b[0] == 'col' and b[2] == torch.float64:
                self._emit(CMP | (REL_CLOSE << 8) | F64 | A_COL | B_COL, a[1], b[1])
            else:
                raise NotImplementedError(f"residual {tree!r}: isclose of a float column and "
                                          "an integer")
            return
        if b[0] == 'float':
            if not a_f:
                raise NotImplementedError(f"residual {tree!r}: an integer column against a "
                                          "float literal compares in float32")
            self._emit(CMP | (rel << 8) | F64 | A_COL, a[1], _f64_bits(b[1]))
            return
        if b[0] == 'int':
            self._literal_cmp(tree, rel, a, b[1])
            return
        b_f = b[2] == torch.float64
        self._emit(CMP | (rel << 8) | (F64 if a_f or b_f else 0) | A_COL | B_COL, a[1], b[1])

    def _literal_cmp(self, tree, rel: int, a, v: int) -> None:
        """A column against an integer: float64 for a float64 column, else
        the integer cast to the column's dtype (torch wraps it)."""
        if not -(1 << 63) <= v < (1 << 63):
            raise NotImplementedError(f"residual {tree!r}: literal outside int64")
        if a[2] == torch.float64:
            self._emit(CMP | (rel << 8) | F64 | A_COL, a[1], _f64_bits(float(v)))
        else:
            self._emit(CMP | (rel << 8) | A_COL, a[1], _wrap(v, a[2]))

    def add(self, tree) -> None:
        tree = normalize_expr_tree(tree)
        if isinstance(tree, list) and len(tree) == 2:
            if tree[0] == 'NOT':
                self.add(tree[1])
                self._emit(NOT)
                return
            if tree[0] == 'isnotnull':
                self._emit(TRUE)
                return
            raise NotImplementedError(f"residual function {tree[0]!r} is outside the pair "
                                      "evaluator")
        if isinstance(tree, list) and len(tree) >= 3:
            op = tree[-2]
            if op == 'AS':
                self.add(tree[0])
                return
            if op in _LOGIC:
                self.add(tree[0])
                self.add(tree[-1])
                self._emit(_LOGIC[op])
                return
            if op in REL and len(tree) == 3:
                self._cmp(tree)
                return
            if op in _KNOWN_OPS or op in ('IN', 'LIKE') or tree[0] == 'CASE':
                raise NotImplementedError(f"residual {tree!r} is outside the pair evaluator")
            raise ValueError(f"Unsupported operation: {op}")
        raise NotImplementedError(f"residual condition {tree!r} is outside the pair evaluator")


def compile_conditions(tensors_left: Dict, tensors_right: Dict,
                       conditions: Sequence) -> Program:
    """One program that is true where every condition holds."""
    if not conditions:
        raise ValueError("compile_conditions: no condition")
    p = Program(tensors_left, tensors_right)
    for k, cond in enumerate(conditions):
        p.add(cond)
        if k:
            p._emit(AND)
    return p


def _count(prog: Program, left_m: torch.Tensor, right_m: torch.Tensor,
           mask: Optional[torch.Tensor]):
    _lib.require_cuda(left_m, "the left pair rows")
    _lib.require_cuda(right_m, "the right pair rows")
    if left_m.dtype != torch.int64 or right_m.dtype != torch.int64:
        raise ValueError("residual: int64 pair rows")
    if left_m.shape != right_m.shape or left_m.dim() != 1:
        raise ValueError("residual: the pair rows are two 1-D tensors of one length")
    n = left_m.numel()
    dev = left_m.device
    if right_m.device != dev:
        raise ValueError("residual: the pair rows are on two devices")
    for t in prog.cols:
        if t.device != dev:
            raise ValueError(f"residual: a column on {t.device}, the pairs on {dev}")
    lib = _lib.load()
    ws_bytes = lib.pf_pair_filter_workspace_bytes(n)
    ws = _lib.workspace(ws_bytes, dev)
    host = torch.empty(2, dtype=torch.int64, pin_memory=True)  # per call: thread-safe
    flat = [w for ins in prog.ins for w in ins]
    with torch.cuda.device(dev):
        _lib.check(lib.pf_pair_filter_count(
            _lib.ptr(left_m), _lib.ptr(right_m), n,
            _lib.void_array([t.data_ptr() for t in prog.cols]),
            ctypes.cast(_lib.i32_array(prog.types), ctypes.c_void_p),
            ctypes.cast(_lib.i32_array(prog.sides), ctypes.c_void_p),
            ctypes.cast(_lib.i64_array([t.numel() for t in prog.cols]), ctypes.c_void_p),
            len(prog.cols), ctypes.cast(_lib.i64_array(flat), ctypes.c_void_p), len(prog.ins),
            _lib.ptr(mask) if mask is not None else None, ctypes.c_void_p(host.data_ptr()),
            _lib.ptr(ws), ws_bytes, _lib.stream_of(dev)))
    kept, bad = int(host[0]), int(host[1])
    if bad:
        raise RuntimeError(f"residual: {bad} pair rows lie outside their column")
    return kept, ws


def filter_pairs(prog: Program, left_m: torch.Tensor, right_m: torch.Tensor
                 ) -> Tuple[torch.Tensor, torch.Tensor]:
    """The pairs where the program holds, in pair order."""
    left_m, right_m = left_m.contiguous(), right_m.contiguous()
    n = left_m.numel()
    if n == 0:
        return left_m, right_m
    kept, ws = _count(prog, left_m, right_m, None)
    dev = left_m.device
    lo = torch.empty(kept, dtype=torch.int64, device=dev)
    ro = torch.empty(kept, dtype=torch.int64, device=dev)
    if kept:
        with torch.cuda.device(dev):
            _lib.check(_lib.load().pf_pair_filter_write(
                _lib.ptr(left_m), _lib.ptr(right_m), n, _lib.ptr(lo), _lib.ptr(ro),
                _lib.ptr(ws), _lib.stream_of(dev)))
    return lo, ro


def evaluate_pairs(prog: Program, left_m: torch.Tensor, right_m: torch.Tensor
                   ) -> torch.Tensor:
    """The program's value on every pair, as a bool mask."""
    left_m, right_m = left_m.contiguous(), right_m.contiguous()
    mask = torch.empty(left_m.numel(), dtype=torch.bool, device=left_m.device)
    if left_m.numel():
        _count(prog, left_m, right_m, mask)
    return mask


__all__ = ["compile_conditions", "filter_pairs", "evaluate_pairs", "column_ref",
           "normalize_expr_tree", "Program"]
