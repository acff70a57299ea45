"""The reference's extension module ``TQPlib.tqpkernels`` over the C ABI.

The reference binds four functions through pybind11
(operators/kernels/tqpkernels.cu:22-46).  This module has the same four,
with the same arguments, outputs and ``host_out`` protocol, calling the
gfx950 library through ctypes:

* ``partition_sort(cols, comp, key_idx, mask, n_valid, goff, B,
  select_first, host_out)`` (partition_sort.cu:610-613);
* ``join_expand(sorted_ids, start, len, host_out)`` (join_core.cu:263-264);
* ``join_unique_hash(build_keys, probe_keys, host_out)``
  (join_core.cu:301-303);
* ``join_semi_anti(build_key, probe_key, build_val, probe_val, op, anti,
  host_out)`` (join_core.cu:358-361).

Outputs are new tensors at the extension's upper bounds on the key's
device; the caller narrows them by what the call leaves in ``host_out``.
Tensor checks raise RuntimeError (the extension's TORCH_CHECK); the work
runs on the current stream of the key's device; each call waits once.  The
operator modules (``pageflow_amd.operators``) call these exactly as the
reference's operators call the pybind module, so a reference checkout can
also import this module in place of ``TQPlib.tqpkernels``.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence

import torch

from . import _lib
from .layout import row_bytes, rows_per_unit


def _hp(host: torch.Tensor):
    if host.device.type != "cpu" or not host.is_pinned() or host.dtype != torch.int64:
        raise RuntimeError("host_out must be a pinned int64 CPU tensor")
    return ctypes.c_void_p(host.data_ptr())


def _i64_col(t: torch.Tensor, what: str, device=None) -> None:
    if not t.is_cuda or t.dtype != torch.int64 or t.dim() != 1 or not t.is_contiguous():
        raise RuntimeError(f"tqpkernels: {what} must be a contiguous 1-D int64 CUDA tensor")
    if device is not None and t.device != device:
        raise RuntimeError(f"tqpkernels: {what} is on {t.device}, expected {device}")


def partition_sort(cols: Sequence[torch.Tensor], comp: Sequence[int], key_idx: int,
                   mask: Optional[torch.Tensor], n_valid: int, goff: int, B: int,
                   select_first: bool, host_out: torch.Tensor) -> List[torch.Tensor]:
    """The tick's columns, then its gidx (int64), each partitioned into the
    padded layout and sized n_valid + B (u - 1) rows (u = lcm(row bytes,
    512) / row bytes).  host_out[:B] the bucket counts, then each row-width
    class's padded total, and host_out[B + 4] the kept rows a mask prologue
    found (-1: none ran)."""
    if not cols:
        raise RuntimeError("partition_sort: no columns")
    key = cols[key_idx]
    device = key.device
    for t in cols:
        if not t.is_cuda or t.device != device or not t.is_contiguous():
            raise RuntimeError("partition_sort: columns must be contiguous on the key's device")
    if mask is not None and (not mask.is_cuda or mask.device != device):
        raise RuntimeError("partition_sort: the mask must be on the key's device")
    n_pre = mask.shape[0] if mask is not None else key.shape[0]
    lib = _lib.load()
    outs = []
    for t in cols:
        shape = list(t.shape)
        shape[0] = n_valid + B * (rows_per_unit(row_bytes(t)) - 1)
        outs.append(torch.empty(shape, dtype=t.dtype, device=device))
    gidx = torch.empty(n_valid + B * (rows_per_unit(8) - 1), dtype=torch.int64, device=device)
    ws_bytes = lib.pf_partition_sort_workspace_bytes(n_pre, n_valid, B, int(mask is not None),
                                                     int(bool(select_first)),
                                                     int(any(comp)))
    ws = _lib.workspace(ws_bytes, device)
    rbs = _lib.i32_array([row_bytes(t) for t in cols])
    cps = _lib.i32_array([int(c) for c in comp])
    with torch.cuda.device(device):
        _lib.check(lib.pf_partition_sort(
            _lib.void_array([t.data_ptr() for t in cols]),
            _lib.void_array([o.data_ptr() for o in outs]),
            ctypes.cast(rbs, ctypes.c_void_p), ctypes.cast(cps, ctypes.c_void_p),
            len(cols), key_idx, _lib.ptr(mask), n_pre, n_valid, goff, B,
            int(bool(select_first)), _lib.ptr(gidx), _hp(host_out), _lib.ptr(ws), ws_bytes,
            _lib.stream_of(device)))
    outs.append(gidx)
    return outs


def join_expand(sorted_ids: torch.Tensor, start: torch.Tensor, len: torch.Tensor,
                host_out: torch.Tensor) -> List[torch.Tensor]:
    """[build rows, probe rows]: (sorted_ids[start[i] + j], i) for every
    probe row i and j < len[i], probe-row then run order; host_out[0] the
    pair count.  The third parameter is named ``len`` as in the reference's
    binding (tqpkernels.cu:29-34), so keyword callers keep working."""
    runs = len  # the run lengths; the builtin is not used below
    device = sorted_ids.device
    for t, what in ((sorted_ids, "sorted_ids"), (start, "start"), (runs, "len")):
        _i64_col(t, what, device)
    if start.numel() != runs.numel():
        raise RuntimeError("join_expand: start and len differ in length")
    m = runs.numel()
    if m == 0:
        host_out[0] = 0
        return [torch.empty(0, dtype=torch.int64, device=device) for _ in range(2)]
    lib = _lib.load()
    ws_bytes = lib.pf_join_expand_workspace_bytes(m)
    ws = _lib.workspace(ws_bytes, device)
    st = _lib.stream_of(device)
    with torch.cuda.device(device):
        _lib.check(lib.pf_join_expand_count(_lib.ptr(runs), m, _hp(host_out), _lib.ptr(ws),
                                            ws_bytes, st))
        n_flat = int(host_out[0])
        cs = torch.empty(n_flat, dtype=torch.int64, device=device)
        cr = torch.empty(n_flat, dtype=torch.int64, device=device)
        if n_flat > 0:
            _lib.check(lib.pf_join_expand_write(_lib.ptr(sorted_ids), _lib.ptr(start), m,
                                                n_flat, _lib.ptr(cs), _lib.ptr(cr),
                                                _lib.ptr(ws), st))
    return [cs, cr]


def join_unique_hash(build_keys: Sequence[torch.Tensor], probe_keys: Sequence[torch.Tensor],
                     host_out: torch.Tensor) -> List[torch.Tensor]:
    """[build rows, probe rows] of the inner join on a unique build tuple of
    1 or 2 int64 keys, each sized for every probe row; the first
    host_out[0] hold the matches in probe-row order."""
    if len(build_keys) not in (1, 2) or len(build_keys) != len(probe_keys):
        raise RuntimeError("join_unique_hash: one or two aligned build/probe keys")
    device = build_keys[0].device
    for t in list(build_keys) + list(probe_keys):
        _i64_col(t, "a key", device)
    nb, np_ = build_keys[0].numel(), probe_keys[0].numel()
    if any(t.numel() != nb for t in build_keys) or any(t.numel() != np_ for t in probe_keys):
        raise RuntimeError("join_unique_hash: key lengths disagree")
    build_idx = torch.empty(np_, dtype=torch.int64, device=device)
    probe_idx = torch.empty(np_, dtype=torch.int64, device=device)
    if nb == 0 or np_ == 0:
        host_out[0] = 0
        return [build_idx, probe_idx]
    nkeys = len(build_keys)
    lib = _lib.load()
    ws_bytes = lib.pf_join_unique_hash_workspace_bytes(nb, np_, nkeys)
    ws = _lib.workspace(ws_bytes, device)
    b1 = build_keys[1] if nkeys == 2 else build_keys[0]
    p1 = probe_keys[1] if nkeys == 2 else probe_keys[0]
    with torch.cuda.device(device):
        _lib.check(lib.pf_join_unique_hash(
            _lib.ptr(build_keys[0]), _lib.ptr(b1), nb, _lib.ptr(probe_keys[0]), _lib.ptr(p1),
            np_, nkeys, _lib.ptr(build_idx), _lib.ptr(probe_idx), _hp(host_out),
            _lib.ptr(ws), ws_bytes, _lib.stream_of(device)))
    return [build_idx, probe_idx]


def join_semi_anti(build_key: torch.Tensor, probe_key: torch.Tensor,
                   build_val: Optional[torch.Tensor], probe_val: Optional[torch.Tensor],
                   op: int, anti: bool, host_out: torch.Tensor) -> torch.Tensor:
    """The kept probe rows, ascending, sized for every probe row:
    host_out[0] of them are valid, host_out[1] = 1 when a build key equals
    int64 min (then nothing is)."""
    device = build_key.device
    _i64_col(build_key, "the build key", device)
    _i64_col(probe_key, "the probe key", device)
    nb, np_ = build_key.numel(), probe_key.numel()
    if op:
        if build_val is None or probe_val is None:
            raise RuntimeError("join_semi_anti: a summary needs both value columns")
        _i64_col(build_val, "the build value", device)
        _i64_col(probe_val, "the probe value", device)
        if build_val.numel() != nb or probe_val.numel() != np_:
            raise RuntimeError("join_semi_anti: a value column differs in length from its key")
    out = torch.empty(np_, dtype=torch.int64, device=device)
    if nb == 0 or np_ == 0:
        n = np_ if (anti and nb == 0) else 0
        if n:
            torch.arange(np_, out=out)
        host_out[0] = n
        host_out[1] = 0
        return out
    lib = _lib.load()
    ws_bytes = lib.pf_join_semi_anti_workspace_bytes(nb, np_, op)
    ws = _lib.workspace(ws_bytes, device)
    with torch.cuda.device(device):
        _lib.check(lib.pf_join_semi_anti(
            _lib.ptr(build_key), nb, _lib.ptr(probe_key), np_,
            _lib.ptr(build_val if op else None), _lib.ptr(probe_val if op else None), op,
            int(bool(anti)), _lib.ptr(out), _hp(host_out), _lib.ptr(ws), ws_bytes,
            _lib.stream_of(device)))
    return out


__all__ = ["partition_sort", "join_expand", "join_unique_hash", "join_semi_anti"]
