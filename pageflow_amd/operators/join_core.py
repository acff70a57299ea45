"""The join core's kernels over the gfx950 library
(include/pageflow_join.h), called once per core call by
``pageflow_amd.operators.hashjoin``.

Drop-in for the reference's ``operators/join_core.py`` (:54-143):

- ``unique_hash_pairs``: an inner join on a build tuple (1-2 int64 keys)
  proven unique, one partner per probe row;
- ``semi_anti``: a semi or anti join on a hash set of the build keys, with an
  optional per-key min/max summary answering one residual comparison
  between a build column and a probe column;
- ``expand``: each probe row's run of equal keys in a sorted build as pairs.

Each call waits for the GPU once.  The hash paths' outputs are allocated for
every probe row and narrowed here by the count the call reads back;
``expand`` waits for the pair count, which sizes its outputs.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence, Tuple

import torch

from .. import _lib, tqpkernels

# Rows a call takes on a side (operators/join_core.py:30).
MAX_ROWS = (1 << 31) - (1 << 16)

# The residual operators a summary answers, oriented build.v OP probe.v.
SUMMARY_OPS = {'<>': 1, '<': 2, '<=': 3, '>': 4, '>=': 5}

_host: Optional[torch.Tensor] = None


def _host_buffer() -> torch.Tensor:
    """The process's pinned read-back buffer (a call has read it before it
    returns, so one buffer serves every call of a thread at a time)."""
    global _host
    if _host is None:
        _host = torch.empty(2, dtype=torch.int64, pin_memory=True)
    return _host


def _empty(device) -> torch.Tensor:
    return torch.empty(0, dtype=torch.int64, device=device)


def _hp(host: torch.Tensor):
    return ctypes.c_void_p(host.data_ptr())


def expand(sorted_ids: torch.Tensor, start: torch.Tensor,
           length: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor]:
    """The sort/search probe's pairs.  ``sorted_ids``: the build rows in key
    order; ``start[i]``, ``length[i]``: probe row i's run in it (0: no
    match).  Returns (build rows, probe rows) in probe-row order and, within
    a probe row, run order."""
    return tuple(tqpkernels.join_expand(sorted_ids, start, length, _host_buffer()))


def unique_hash_peak_bytes(n_build: int, n_probe: int) -> int:
    """The device memory of a ``unique_hash_pairs`` call: the table (an
    int64 key and an int32 row a slot, two slots a build row), the partner
    and the two outputs, sized for every probe row."""
    return 2 * n_build * 12 + n_probe * (4 + 8 + 8)


def _validate_keys(build_keys: Sequence[torch.Tensor],
                   probe_keys: Sequence[torch.Tensor]) -> None:
    if len(build_keys) not in (1, 2) or len(build_keys) != len(probe_keys):
        raise ValueError("unique_hash_pairs expects one or two aligned build/probe keys")
    device = build_keys[0].device
    for side, keys in (("build", build_keys), ("probe", probe_keys)):
        rows = keys[0].numel()
        for key in keys:
            if not key.is_cuda or key.dtype != torch.int64 or key.dim() != 1:
                raise ValueError(
                    f"{side} keys must be 1-D CUDA int64 tensors; got device={key.device}, "
                    f"dtype={key.dtype}, shape={tuple(key.shape)}")
            if not key.is_contiguous():
                raise ValueError(f"{side} keys must be contiguous")
            if key.numel() != rows:
                raise ValueError(f"{side} key lengths disagree")
            if key.device != device:
                raise ValueError("all build/probe keys must be on one CUDA device")
        if rows > MAX_ROWS:
            raise ValueError(f"unique hash core takes at most {MAX_ROWS} rows a side")


def unique_hash_pairs(build_keys: Sequence[torch.Tensor],
                      probe_keys: Sequence[torch.Tensor]) -> Tuple[torch.Tensor, torch.Tensor]:
    """``(build rows, probe rows)`` of an inner join whose build tuple is
    unique, in probe-row order; a probe row takes the lowest build row with
    its tuple."""
    _validate_keys(build_keys, probe_keys)
    device = build_keys[0].device
    nb, np_ = build_keys[0].numel(), probe_keys[0].numel()
    if nb == 0 or np_ == 0:
        return _empty(device), _empty(device)
    host = _host_buffer()
    build_idx, probe_idx = tqpkernels.join_unique_hash(list(build_keys), list(probe_keys), host)
    n = int(host[0])
    return build_idx[:n], probe_idx[:n]


def _check_column(t: torch.Tensor, what: str, like: torch.Tensor) -> None:
    if not (t.is_cuda and t.device == like.device and t.dtype == torch.int64
            and t.dim() == 1 and t.is_contiguous()):
        raise RuntimeError(f"join_core: {what} must be a contiguous 1-D int64 tensor on the "
                           f"key's device")


def semi_anti(build_key: torch.Tensor, probe_key: torch.Tensor, anti: bool,
              summary: Optional[Tuple[torch.Tensor, torch.Tensor, str]] = None
              ) -> Optional[torch.Tensor]:
    """The probe rows, ascending, that have a matching build key (semi) or
    none (anti); with ``summary = (build column, probe column, op)`` a match
    must also pass ``build column op probe column`` (``SUMMARY_OPS``).
    None when a build key equals the table's empty-slot value (int64 min):
    the caller then takes the pair path."""
    device = build_key.device
    n_build, n_probe = build_key.numel(), probe_key.numel()
    if n_probe == 0:
        return _empty(device)
    if n_build == 0:
        return (torch.arange(n_probe, dtype=torch.int64, device=device)
                if anti else _empty(device))
    _check_column(build_key, "the build key", build_key)
    _check_column(probe_key, "the probe key", build_key)
    if n_probe > MAX_ROWS:
        raise RuntimeError(f"join_core: the probe key has {n_probe} rows; a call takes at "
                           f"most {MAX_ROWS}")
    build_val = probe_val = None
    op = 0
    if summary is not None:
        build_val, probe_val, op_name = summary
        op = SUMMARY_OPS[op_name]
        _check_column(build_val, "the build value", build_key)
        _check_column(probe_val, "the probe value", build_key)
        if build_val.numel() != n_build or probe_val.numel() != n_probe:
            raise RuntimeError("join_core: a value column differs in length from its key")
    host = _host_buffer()
    out = tqpkernels.join_semi_anti(build_key, probe_key, build_val, probe_val, op, anti, host)
    n, bad = host.tolist()
    if bad:
        return None
    return out[:n]


def semi_anti_mask(build_key: torch.Tensor, probe_key: torch.Tensor, anti: bool,
                   summary: Optional[Tuple[torch.Tensor, torch.Tensor, str]] = None
                   ) -> Optional[torch.Tensor]:
    """``semi_anti`` as a bool row mask over the probe rows (True: kept).
    None when a build key equals int64 min."""
    device = probe_key.device
    n_build, n_probe = build_key.numel(), probe_key.numel()
    if n_build == 0 or n_probe == 0:
        return torch.full((n_probe,), bool(anti), dtype=torch.bool, device=device)
    _check_column(build_key, "the build key", build_key)
    _check_column(probe_key, "the probe key", build_key)
    build_val = probe_val = None
    op = 0
    if summary is not None:
        build_val, probe_val, op_name = summary
        op = SUMMARY_OPS[op_name]
        _check_column(build_val, "the build value", build_key)
        _check_column(probe_val, "the probe value", build_key)
    lib = _lib.load()
    host = _host_buffer()
    ws_bytes = lib.pf_semi_anti_mask_workspace_bytes(n_build, n_probe, op)
    ws = _lib.workspace(ws_bytes, device)
    mask = torch.empty(n_probe, dtype=torch.bool, device=device)
    st = _lib.stream_of(device)
    with torch.cuda.device(device):
        _lib.check(lib.pf_semi_anti_mask(
            _lib.ptr(build_key), n_build, _lib.ptr(probe_key), n_probe, _lib.ptr(build_val),
            _lib.ptr(probe_val), op, int(bool(anti)), _lib.ptr(mask), _hp(host),
            _lib.ptr(ws), ws_bytes, st))
        torch.cuda.current_stream(device).synchronize()
    if int(host[0]):
        return None
    return mask
