"""Per-call join kernels for one core tick, on the gfx950 library.

Stands in for the reference's ``operators/join_core.py`` (its public
functions at :54-143 keep their names, arguments, return conventions and
exception types; the bodies here are this package's own):

``unique_hash_pairs(build_keys, probe_keys)``
    inner join of a build tuple (one or two int64 columns) that the planner
    proved unique; every probe row meets at most one build row.
``semi_anti(build_key, probe_key, anti, summary=None)``
    the probe rows that do (semi) or do not (anti) meet a build key, a match
    optionally also passing one comparison between a build value and a probe
    value, answered from a per-key min/max summary.
``expand(sorted_ids, start, length)``
    every probe row's run of a key-sorted build as (build row, probe row).

A call's device work ends in one host wait for the count it needs (the
result rows, or the pair count that sizes ``expand``'s outputs), which the
library leaves in a pinned buffer; outputs come back narrowed to it.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence, Tuple

import torch

from .. import _lib, tqpkernels

# Rows a side may have (the reference's operators/join_core.py:30): the
# kernels index rows in int32 tiles just below 2^31.
MAX_ROWS = (1 << 31) - (1 << 16)

# Comparisons a per-key summary answers, written build.value OP probe.value.
SUMMARY_OPS = {'<>': 1, '<': 2, '<=': 3, '>': 4, '>=': 5}


class _Readback:
    """The pinned int64 words the library reports counts in.  One set serves
    the whole process: every call has consumed them before it returns, so
    calls from one thread never overlap on them (calls from several threads
    must not run at once, as with the reference's module-level buffer)."""

    def __init__(self, words: int):
        self.words = words
        self._buf: Optional[torch.Tensor] = None

    def buffer(self) -> torch.Tensor:
        if self._buf is None:
            self._buf = torch.zeros(self.words, dtype=torch.int64, pin_memory=True)
        return self._buf

    def address(self) -> ctypes.c_void_p:
        return ctypes.c_void_p(self.buffer().data_ptr())


_READBACK = _Readback(2)


def _host_buffer() -> torch.Tensor:
    return _READBACK.buffer()


def _no_rows(device) -> torch.Tensor:
    return torch.empty(0, dtype=torch.int64, device=device)


def _vector_fault(t: torch.Tensor) -> Optional[str]:
    """Why ``t`` cannot be a key or value column of these kernels, or None:
    they read 1-D contiguous int64 device vectors."""
    if not t.is_cuda:
        return f"is on {t.device}, not a GPU"
    if t.dtype != torch.int64:
        return f"is {t.dtype}, not int64"
    if t.dim() != 1:
        return f"has shape {tuple(t.shape)}, not one dimension"
    if not t.is_contiguous():
        return "is not contiguous"
    return None


# ------------------------------------------------------------- expand --

def expand(sorted_ids: torch.Tensor, start: torch.Tensor,
           length: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor]:
    """Pairs of the sort/search probe.  ``sorted_ids`` lists the build rows
    in key order; probe row i's matches are ``sorted_ids[start[i] :
    start[i] + length[i]]`` (length 0: none).  Returns (build rows, probe
    rows) ordered by probe row, then by position in the run."""
    cs, cr = tqpkernels.join_expand(sorted_ids, start, length, _host_buffer())
    return cs, cr


# -------------------------------------------------------- unique hash --

def unique_hash_peak_bytes(n_build: int, n_probe: int) -> int:
    """The device memory a ``unique_hash_pairs`` call is budgeted at by the
    join dispatch's scratch guard: an int32 table of two slots a build row,
    an int32 partner and two int64 outputs a probe row.  The figure is the
    reference's (operators/join_core.py:64-68), so the guard takes the same
    decision on the same call.  This library's table also keeps each slot's
    int64 key beside its row (one probe load a step instead of two dependent
    ones); its exact workspace is ``pf_join_unique_hash_workspace_bytes``."""
    return 8 * n_build + 20 * n_probe


def _key_tuple_check(build_keys: Sequence[torch.Tensor],
                     probe_keys: Sequence[torch.Tensor]) -> Tuple[int, int]:
    """(build rows, probe rows) of a well-formed key tuple; ValueError
    naming the first problem otherwise."""
    if len(build_keys) != len(probe_keys) or len(build_keys) not in (1, 2):
        raise ValueError(f"unique_hash_pairs expects one or two aligned build/probe keys; "
                         f"got {len(build_keys)} build and {len(probe_keys)} probe keys")
    device = build_keys[0].device
    rows = []
    for side, keys in (("build", build_keys), ("probe", probe_keys)):
        n = keys[0].numel()
        for i, k in enumerate(keys):
            fault = _vector_fault(k)
            if fault:
                raise ValueError(f"{side} key {i} {fault}: keys must be 1-D contiguous CUDA "
                                 f"int64 tensors")
            if k.device != device:
                raise ValueError(f"{side} key {i} is on {k.device}, the first build key on "
                                 f"{device}: all keys must share one CUDA device")
            if k.numel() != n:
                raise ValueError(f"{side} keys have {n} and {k.numel()} rows")
        if n > MAX_ROWS:
            raise ValueError(f"{side} side has {n} rows; the unique hash core takes at most "
                             f"{MAX_ROWS} a side")
        rows.append(n)
    return rows[0], rows[1]


def unique_hash_pairs(build_keys: Sequence[torch.Tensor],
                      probe_keys: Sequence[torch.Tensor]) -> Tuple[torch.Tensor, torch.Tensor]:
    """(build rows, probe rows) of the inner join, ordered by probe row.  The
    build tuple is unique by the planner's proof; should a tuple repeat, a
    probe row takes its lowest build row."""
    n_build, n_probe = _key_tuple_check(build_keys, probe_keys)
    device = build_keys[0].device
    if n_build == 0 or n_probe == 0:
        return _no_rows(device), _no_rows(device)
    host = _host_buffer()
    build_rows, probe_rows = tqpkernels.join_unique_hash(list(build_keys), list(probe_keys),
                                                         host)
    matched = int(host[0])
    return build_rows[:matched], probe_rows[:matched]


# -------------------------------------------------------- semi / anti --

def _summary_columns(summary, build_key: torch.Tensor, n_build: int, n_probe: int):
    """(build values, probe values, op code) of a summary, checked as the
    extension checks its tensors (RuntimeError); (None, None, 0) without
    one."""
    if summary is None:
        return None, None, 0
    build_val, probe_val, op_name = summary
    if op_name not in SUMMARY_OPS:
        raise KeyError(op_name)
    for what, col, n in (("build value", build_val, n_build), ("probe value", probe_val,
                                                                n_probe)):
        fault = _vector_fault(col)
        if fault:
            raise RuntimeError(f"join_core: the {what} column {fault}")
        if col.device != build_key.device:
            raise RuntimeError(f"join_core: the {what} column is on {col.device}, the keys on "
                               f"{build_key.device}")
        if col.numel() != n:
            raise RuntimeError(f"join_core: the {what} column has {col.numel()} rows, its key "
                               f"{n}")
    return build_val, probe_val, SUMMARY_OPS[op_name]


def _set_keys_check(build_key: torch.Tensor, probe_key: torch.Tensor) -> None:
    for what, k in (("build key", build_key), ("probe key", probe_key)):
        fault = _vector_fault(k)
        if fault:
            raise RuntimeError(f"join_core: the {what} {fault}")
    if probe_key.device != build_key.device:
        raise RuntimeError("join_core: the build and probe keys are on two devices")
    if probe_key.numel() > MAX_ROWS:
        raise RuntimeError(f"join_core: the probe key has {probe_key.numel()} rows; a call "
                           f"takes at most {MAX_ROWS}")


def semi_anti(build_key: torch.Tensor, probe_key: torch.Tensor, anti: bool,
              summary: Optional[Tuple[torch.Tensor, torch.Tensor, str]] = None
              ) -> Optional[torch.Tensor]:
    """Ascending probe rows that meet a build key (semi) or meet none
    (anti).  With ``summary = (build values, probe values, op)`` a meeting
    counts only if ``build value op probe value`` holds for some build row of
    the key.  Returns None when a build key is int64 min, the set's empty
    marker; the caller then takes the pair path."""
    device = build_key.device
    n_build, n_probe = build_key.numel(), probe_key.numel()
    if n_probe == 0:
        return _no_rows(device)
    if n_build == 0:
        # nothing to meet: anti keeps every probe row, semi none
        return torch.arange(n_probe, dtype=torch.int64, device=device) if anti \
            else _no_rows(device)
    _set_keys_check(build_key, probe_key)
    build_val, probe_val, op = _summary_columns(summary, build_key, n_build, n_probe)
    host = _host_buffer()
    rows = tqpkernels.join_semi_anti(build_key, probe_key, build_val, probe_val, op, anti, host)
    kept, empty_marker_met = (int(v) for v in host.tolist())
    if empty_marker_met:
        return None
    return rows[:kept]


def semi_anti_mask(build_key: torch.Tensor, probe_key: torch.Tensor, anti: bool,
                   summary: Optional[Tuple[torch.Tensor, torch.Tensor, str]] = None
                   ) -> Optional[torch.Tensor]:
    """``semi_anti``'s answer as a bool mask over the probe rows (True:
    kept), with the same None convention.  No counterpart in the reference:
    the mask form skips the compaction when the caller wants a mask."""
    device = probe_key.device
    n_build, n_probe = build_key.numel(), probe_key.numel()
    if n_build == 0 or n_probe == 0:
        return torch.full((n_probe,), bool(anti), dtype=torch.bool, device=device)
    _set_keys_check(build_key, probe_key)
    build_val, probe_val, op = _summary_columns(summary, build_key, n_build, n_probe)
    lib = _lib.load()
    ws_bytes = lib.pf_semi_anti_mask_workspace_bytes(n_build, n_probe, op)
    ws = _lib.workspace(ws_bytes, device)
    mask = torch.empty(n_probe, dtype=torch.bool, device=device)
    with torch.cuda.device(device):
        _lib.check(lib.pf_semi_anti_mask(
            _lib.ptr(build_key), n_build, _lib.ptr(probe_key), n_probe, _lib.ptr(build_val),
            _lib.ptr(probe_val), op, int(bool(anti)), _lib.ptr(mask), _READBACK.address(),
            _lib.ptr(ws), ws_bytes, _lib.stream_of(device)))
        torch.cuda.current_stream(device).synchronize()
    return None if int(_host_buffer()[0]) else mask
