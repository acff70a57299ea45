"""The reference's zero-copy reader extension (``TQPlib.zero_copy_reader``,
IO/zero_copy_reader.cu:224-318) over the C ABI.

Three functions with the extension's arguments, outputs and checks:

* ``device_alias(host)`` (zero_copy_reader.cu:224-233): the device address
  of a pinned host tensor's first byte, as an int for ``seg_ptrs``;
* ``zero_copy_mask(seg_ptrs, seg_rows, mask, result, row_start, row_stop)``
  (:235-265): the rows a bool mask keeps, into ``result``;
* ``zero_copy_rearrange(seg_ptrs, seg_rows, rearrange, row_start, row_stop,
  dtype, width)`` (:267-292): the gather of the join's late
  materialisation, ``result[r]`` = row ``row_start + rearrange[r]``, or -64
  in every element where ``rearrange[r]`` is -64 (an outer join's null row).

A column in pinned host memory is a list of segments: ``seg_ptrs`` (int64,
on the GPU) holds each segment's device address, ``seg_rows`` (int64, on
the GPU, one more entry) each segment's first row in the column's row space
and the end of the last.  The GPU reads the host rows itself.  Where the
extension traps on a row outside every segment, these raise RuntimeError
(after waiting for the call's count of such rows); argument errors raise
RuntimeError, as TORCH_CHECK does.
"""
from __future__ import annotations

import ctypes
from typing import Sequence, Tuple

import torch

from . import _lib

_ELEM = {torch.uint8: (1, 0), torch.int8: (1, 0), torch.int16: (2, 0), torch.int32: (4, 0),
         torch.int64: (8, 0), torch.float32: (4, 1), torch.float64: (8, 1)}


def device_alias(host: torch.Tensor) -> int:
    if host.device.type != "cpu":
        raise RuntimeError("device_alias needs a host tensor")
    if not host.is_pinned():
        raise RuntimeError("device_alias needs pinned memory")
    lib = _lib.load()
    dev = ctypes.c_void_p()
    _lib.check(lib.pf_device_alias(ctypes.c_void_p(host.data_ptr()), ctypes.byref(dev)))
    return int(dev.value or 0)


def _check_segments(seg_ptrs: torch.Tensor, seg_rows: torch.Tensor, row_start: int,
                    row_stop: int) -> None:
    # the extension's check_segments (zero_copy_reader.cu:122-138)
    if not (seg_ptrs.is_cuda and seg_rows.is_cuda):
        raise RuntimeError("segment tensors must be on the GPU")
    if seg_ptrs.dtype != torch.int64 or seg_rows.dtype != torch.int64:
        raise RuntimeError("segment tensors must be int64")
    if seg_ptrs.dim() != 1 or not seg_ptrs.is_contiguous() or not seg_rows.is_contiguous():
        raise RuntimeError("segment tensors must be contiguous and 1-D")
    if seg_ptrs.numel() == 0:
        raise RuntimeError("no segment to read")
    if seg_rows.numel() != seg_ptrs.numel() + 1:
        raise RuntimeError(f"seg_rows needs one entry per segment plus the end "
                           f"({seg_rows.numel()} for {seg_ptrs.numel()} segments)")
    if not 0 <= row_start <= row_stop:
        raise RuntimeError(f"rows [{row_start}, {row_stop}) are not a range")


def _elem(dtype) -> Tuple[int, int]:
    if dtype not in _ELEM:
        raise RuntimeError(f"Unsupported dtype {dtype} for a zero-copy read")
    return _ELEM[dtype]


def _raise_bad(bad: torch.Tensor, what: str) -> None:
    n = int(bad.item())
    if n:
        raise RuntimeError(f"{what}: {n} rows lie outside every segment")


def zero_copy_mask(seg_ptrs: torch.Tensor, seg_rows: torch.Tensor, mask: torch.Tensor,
                   result: torch.Tensor, row_start: int, row_stop: int) -> None:
    _check_segments(seg_ptrs, seg_rows, row_start, row_stop)
    if not mask.is_cuda:
        raise RuntimeError("mask must be on GPU")
    if not result.is_cuda:
        raise RuntimeError("result must be on GPU")
    if not mask.is_contiguous():
        raise RuntimeError("mask must be contiguous")
    if not result.is_contiguous():
        raise RuntimeError("result must be contiguous")
    if mask.dtype != torch.bool:
        raise RuntimeError("mask must be bool")
    if mask.dim() != 1:
        raise RuntimeError("mask must be 1-D")
    if mask.numel() != row_stop - row_start:
        raise RuntimeError(f"mask length ({mask.numel()}) != row_stop - row_start "
                           f"({row_stop - row_start})")
    eb, fl = _elem(result.dtype)
    width = result.shape[1] if result.dim() > 1 else 1
    n_out = result.shape[0] if result.dim() > 0 else 1
    dev = result.device
    lib = _lib.load()
    n = row_stop - row_start
    ws_bytes = lib.pf_zero_copy_mask_workspace_bytes(n)
    ws = _lib.workspace(ws_bytes, dev)
    counts = torch.zeros(2, dtype=torch.int64, device=dev)  # kept rows, rows outside
    with torch.cuda.device(dev):
        _lib.check(lib.pf_zero_copy_mask(
            _lib.ptr(seg_ptrs), _lib.ptr(seg_rows), seg_ptrs.numel(), row_start, row_stop,
            _lib.ptr(mask), eb, fl, width, _lib.ptr(result), n_out,
            ctypes.c_void_p(counts.data_ptr()), ctypes.c_void_p(counts.data_ptr() + 8),
            _lib.ptr(ws), ws_bytes, _lib.stream_of(dev)))
    kept, bad = (int(v) for v in counts.tolist())
    if kept != n_out:
        raise RuntimeError(f"mask selected count ({kept}) vs result rows ({n_out}) mismatch")
    if bad:
        raise RuntimeError(f"zero_copy_mask: {bad} rows lie outside every segment")


def zero_copy_rearrange(seg_ptrs: torch.Tensor, seg_rows: torch.Tensor,
                        rearrange: torch.Tensor, row_start: int, row_stop: int, dtype,
                        width: int) -> torch.Tensor:
    _check_segments(seg_ptrs, seg_rows, row_start, row_stop)
    if not rearrange.is_cuda:
        raise RuntimeError("rearrange must be on GPU")
    if not rearrange.is_contiguous():
        raise RuntimeError("rearrange must be contiguous")
    if rearrange.dtype != torch.int64:
        raise RuntimeError("rearrange must be int64")
    if width < 1:
        raise RuntimeError("width must be positive")
    eb, fl = _elem(dtype)
    dev = rearrange.device
    n = rearrange.numel()
    result = torch.empty((n, width), dtype=dtype, device=dev)
    if n == 0:
        return result
    lib = _lib.load()
    ws_bytes = lib.pf_zero_copy_rearrange_workspace_bytes(n)
    ws = _lib.workspace(ws_bytes, dev)
    bad = torch.zeros(1, dtype=torch.int64, device=dev)
    with torch.cuda.device(dev):
        _lib.check(lib.pf_zero_copy_rearrange(
            _lib.ptr(seg_ptrs), _lib.ptr(seg_rows), seg_ptrs.numel(), row_start, row_stop,
            _lib.ptr(rearrange), n, eb, fl, width, _lib.ptr(result), _lib.ptr(bad),
            _lib.ptr(ws), ws_bytes, _lib.stream_of(dev)))
    _raise_bad(bad, "zero_copy_rearrange")
    return result


def segments(host_parts: Sequence[torch.Tensor], first_row: int = 0,
             device=None) -> Tuple[torch.Tensor, torch.Tensor]:
    """(seg_ptrs, seg_rows) of a column held as pinned host tensors, one
    segment each, back to back from row ``first_row`` of the row space."""
    device = torch.device(device) if device is not None else torch.device("cuda")
    ptrs, rows = [], [first_row]
    for t in host_parts:
        ptrs.append(device_alias(t))
        rows.append(rows[-1] + t.shape[0])
    return (torch.tensor(ptrs, dtype=torch.int64, device=device),
            torch.tensor(rows, dtype=torch.int64, device=device))


__all__ = ["device_alias", "zero_copy_mask", "zero_copy_rearrange", "segments"]
