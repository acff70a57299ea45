"""Whole-side equi-joins on the partitioned LDS-table kernels
(``pf_hash_join_*``, include/pageflow_join.h): the large-join tier and the
multi-match join that stands in for the reference's sort/search core
(operators/hashjoin.py:380-441) on non-unique builds.

``HashJoin`` keeps its workspace between calls (a serving loop or a
benchmark reuses it); ``hash_join`` is the one-shot form.  Either takes a
residual program (``operators.residual.compile_conditions`` over the build
columns as the left side and the probe columns as the right), evaluated on
every candidate pair inside the count and write passes
(``pf_hash_join_residual_*``): only the kept pairs are counted and written.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence, Tuple

import torch

from . import _lib

CAP_UNIQUE = 6144   # build rows an LDS table holds (unique builds)
CAP_MULTI = 4096    # multi: build rows an LDS chunk; buckets sorted in LDS up to 8192 rows
MAX_BITS = 12


def default_bits(n_build: int, unique: bool) -> int:
    """Bucket bits so the mean build bucket sits well inside one LDS table
    (about 2/3 of it for a unique build, 60% of a multi chunk)."""
    target = CAP_UNIQUE * 2 // 3 if unique else CAP_MULTI * 3 // 5
    bits = 0
    while bits < MAX_BITS and (n_build >> bits) > target:
        bits += 1
    return bits


def _check_keys(t: torch.Tensor, what: str) -> None:
    _lib.require_cuda(t, what)
    if t.dtype != torch.int64 or t.dim() != 1 or not t.is_contiguous():
        raise ValueError(f"{what} must be a contiguous 1-D int64 tensor")
    if t.numel() >= (1 << 31):
        raise ValueError(f"{what}: a side takes fewer than 2^31 rows")


class HashJoin:
    """One join's plan: sizes, bucket bits and the workspace.

    ``ordered=False`` (unique builds only) skips the count pass: ``prepare``
    returns None, the outputs of ``write`` must hold every probe row (a
    unique build gives a probe row at most one pair), pairs come grouped by
    the probe kernel's rounds in no fixed order, and ``write`` returns the
    pair count after one wait."""

    def __init__(self, n_build: int, n_probe: int, unique: bool = True,
                 bits: Optional[int] = None, device=None, ordered: bool = True):
        self.nb, self.np = int(n_build), int(n_probe)
        self.unique = bool(unique)
        self.ordered = bool(ordered) or not self.unique
        # pf_hash_join_*'s multi: 0 unique, 1 multi-match, 2 unique in any order
        self.mode = 1 if not self.unique else (0 if self.ordered else 2)
        self.bits = default_bits(self.nb, self.unique) if bits is None else int(bits)
        if not 0 <= self.bits <= MAX_BITS:
            raise ValueError(f"bits must be in [0, {MAX_BITS}]")
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        if self.device.index is None:
            self.device = torch.device(self.device.type, torch.cuda.current_device())
        lib = _lib.load()
        self.ws_bytes = int(lib.pf_hash_join_workspace_bytes(self.nb, self.np, self.bits,
                                                             self.mode))
        self.ws = _lib.workspace(self.ws_bytes, self.device)
        self.host = torch.empty(4, dtype=torch.int64, pin_memory=True)
        self.n_items = 0
        self.pairs = 0
        self.candidates = 0
        self.residual = None
        self.prepared = False

    def _program_args(self, prog):
        flat = [w for ins in prog.ins for w in ins]
        return (_lib.void_array([t.data_ptr() for t in prog.cols]),
                ctypes.cast(_lib.i32_array(prog.types), ctypes.c_void_p),
                ctypes.cast(_lib.i32_array(prog.sides), ctypes.c_void_p),
                ctypes.cast(_lib.i64_array([t.numel() for t in prog.cols]), ctypes.c_void_p),
                len(prog.cols), ctypes.cast(_lib.i64_array(flat), ctypes.c_void_p),
                len(prog.ins))

    def _residual_workspace(self, prog) -> None:
        if not self.ordered:
            raise ValueError("HashJoin: a residual program needs an ordered plan")
        for t in prog.cols:
            if t.device != self.device:
                raise ValueError(f"HashJoin: a residual column on {t.device}, the join on "
                                 f"{self.device}")
        n_build = sum(1 for sd in prog.sides if sd == 0)
        n_probe = len(prog.sides) - n_build
        need = int(_lib.load().pf_hash_join_residual_workspace_bytes(
            self.nb, self.np, self.bits, self.mode, n_probe, n_build))
        if need > self.ws.numel():
            self.ws = _lib.workspace(need, self.device)
        self.ws_bytes = max(self.ws_bytes, need)

    def prepare(self, build_keys: torch.Tensor, probe_keys: torch.Tensor,
                residual=None) -> Optional[int]:
        """Partitions both sides and, when ordered, counts the pairs (two
        host waits; three when a multi-match build bucket is larger than the
        LDS sort takes and is sorted in HBM; one when not ordered).  Returns
        the pair count, or None when not ordered.  With a ``residual``
        program (ordered plans) the count is of the pairs it keeps, and
        ``candidates`` holds the pairs before it."""
        _check_keys(build_keys, "build keys")
        _check_keys(probe_keys, "probe keys")
        if build_keys.numel() != self.nb or probe_keys.numel() != self.np:
            raise ValueError("HashJoin: sides differ from the plan's sizes")
        lib = _lib.load()
        self.residual = None
        if residual is not None:
            self._residual_workspace(residual)
            with torch.cuda.device(self.device):
                _lib.check(lib.pf_hash_join_residual_prepare(
                    _lib.ptr(build_keys), self.nb, _lib.ptr(probe_keys), self.np, self.bits,
                    self.mode, *self._program_args(residual),
                    ctypes.c_void_p(self.host.data_ptr()), _lib.ptr(self.ws), self.ws_bytes,
                    _lib.stream_of(self.device)))
            pairs, _, self.n_items, self.candidates = (int(v) for v in self.host.tolist())
            self.residual = residual
            self.prepared = True
            self.pairs = pairs
            return self.pairs
        with torch.cuda.device(self.device):
            _lib.check(lib.pf_hash_join_prepare(
                _lib.ptr(build_keys), self.nb, _lib.ptr(probe_keys), self.np, self.bits,
                self.mode, ctypes.c_void_p(self.host.data_ptr()), _lib.ptr(self.ws),
                self.ws_bytes, _lib.stream_of(self.device)))
        pairs, _, self.n_items = (int(v) for v in self.host.tolist()[:3])
        self.prepared = True
        if not self.ordered:
            return None
        self.pairs = self.candidates = pairs
        return self.pairs

    def write(self, out_probe: torch.Tensor, out_build: torch.Tensor,
              payloads: Sequence[torch.Tensor] = (),
              pay_out: Sequence[torch.Tensor] = ()) -> int:
        """Writes the prepared pairs (and gathered build payloads) and
        returns their count; no wait when ordered, one otherwise."""
        if not self.prepared:
            raise RuntimeError("HashJoin: write before prepare")
        if self.ordered and self.pairs < 0:
            raise RuntimeError("HashJoin: the prepared join took no LDS path")
        need = self.pairs if self.ordered else self.np
        if out_probe.numel() < need or out_build.numel() < need:
            raise ValueError(f"HashJoin: outputs of {need} rows are needed")
        if len(payloads) != len(pay_out) or len(payloads) > 2:
            raise ValueError("HashJoin: at most two payload columns, one output each")
        for t in payloads:
            if t.dtype != torch.int64 or t.numel() != self.nb or not t.is_contiguous():
                raise ValueError("HashJoin: payloads are contiguous int64 build columns")
        for t in pay_out:
            if t.dtype != torch.int64 or t.numel() < need:
                raise ValueError(f"HashJoin: payload outputs are int64 of {need} rows")
        lib = _lib.load()
        if self.residual is not None:
            with torch.cuda.device(self.device):
                _lib.check(lib.pf_hash_join_residual_write(
                    self.nb, self.np, self.bits, self.mode, self.n_items,
                    _lib.void_array([t.data_ptr() for t in payloads]), len(payloads),
                    _lib.ptr(out_probe), _lib.ptr(out_build),
                    _lib.void_array([t.data_ptr() for t in pay_out]),
                    *self._program_args(self.residual), _lib.ptr(self.ws), self.ws_bytes,
                    _lib.stream_of(self.device)))
            return self.pairs
        with torch.cuda.device(self.device):
            _lib.check(lib.pf_hash_join_write(
                self.nb, self.np, self.bits, self.mode, self.n_items,
                _lib.void_array([t.data_ptr() for t in payloads]), len(payloads),
                _lib.ptr(out_probe), _lib.ptr(out_build),
                _lib.void_array([t.data_ptr() for t in pay_out]),
                ctypes.c_void_p(self.host.data_ptr()), _lib.ptr(self.ws), self.ws_bytes,
                _lib.stream_of(self.device)))
        if not self.ordered:
            self.pairs = int(self.host[0])
        return self.pairs


def hash_join(build_keys: torch.Tensor, probe_keys: torch.Tensor,
              payloads: Sequence[torch.Tensor] = (), unique: bool = True,
              bits: Optional[int] = None, ordered: bool = True, residual=None
              ) -> Tuple[torch.Tensor, torch.Tensor, List[torch.Tensor]]:
    """(probe rows, build rows, gathered build payloads) of an inner
    equi-join of int64 keys, in the deterministic partition order (or, for
    a unique build with ``ordered=False``, in no fixed order); with a
    ``residual`` program only the pairs it keeps."""
    plan = HashJoin(build_keys.numel(), probe_keys.numel(), unique, bits, build_keys.device,
                    ordered=ordered or residual is not None)
    n = plan.prepare(build_keys, probe_keys, residual)
    dev = build_keys.device
    rows = n if n is not None else probe_keys.numel()
    out_p = torch.empty(rows, dtype=torch.int64, device=dev)
    out_b = torch.empty(rows, dtype=torch.int64, device=dev)
    outs = [torch.empty(rows, dtype=torch.int64, device=dev) for _ in payloads]
    n = plan.write(out_p, out_b, list(payloads), outs)
    return out_p[:n], out_b[:n], [t[:n] for t in outs]
