"""Pass 1 of the reference's partitioned aggregation over the gfx950 kernels.

Drop-in for three helpers of the reference's ``operators/aggregate.py``:

* ``_mix64(h)`` (aggregate.py:95-104): mix64 of an int64 tensor
  (``pf_mix64``);
* ``_bucket_ids(h, B)`` (aggregate.py:107-117): each grouping hash's bucket
  among B (a power of two), the top log2(B) bits of its mix64, as int32
  (``pf_bucket_ids``); a group's bucket is the same in every tick;
* ``_partition_outputs(outs, bucket, counts, counts_t)``
  (aggregate.py:120-152): one pass-1 tick's outputs as
  ``PartitionedTensor``s in the sort pipe's padded layout, bucket b's
  ``counts[b]`` rows from a whole ``lcm(row bytes, 512)``-byte unit, rows in
  their order within a bucket (``pf_partition_by_bucket``: one stable
  partition kernel for up to six outputs, where the reference sorts the
  buckets and ``index_copy_``s every output).

``counts`` (host) and ``counts_t`` (device) are the callers' bucket counts;
the kernel counts the buckets itself and the call raises RuntimeError when
they disagree with ``counts`` or a bucket id lies outside [0, B).
"""
from __future__ import annotations

import ctypes
from math import prod
from typing import List, Optional, Sequence

import torch

from .. import _lib
from ..layout import PartitionedTensor, rows_per_unit
from .partitions import JOIN_B_MAX as _B_MAX

_MAX_COLS = 6
_MAX_WIDTHS = 4

def _host_buffer() -> torch.Tensor:
    """A call's pinned read-back buffer (counts, padded totals, -1,
    out-of-range bucket ids), from torch's caching host allocator, so
    concurrent calls do not share it."""
    return torch.empty(_B_MAX + _MAX_WIDTHS + 2, dtype=torch.int64, pin_memory=True)


def _mix64(h: torch.Tensor) -> torch.Tensor:
    _lib.require_cuda(h, "the hashes")
    if h.dtype != torch.int64:
        raise ValueError("_mix64: int64 hashes")
    h = h.contiguous()
    out = torch.empty_like(h)
    lib = _lib.load()
    with torch.cuda.device(h.device):
        _lib.check(lib.pf_mix64(_lib.ptr(h), h.numel(), _lib.ptr(out), _lib.stream_of(h.device)))
    return out


def _bucket_ids(h: torch.Tensor, B: int) -> torch.Tensor:
    _lib.require_cuda(h, "the hashes")
    if h.dtype != torch.int64 or h.dim() != 1:
        raise ValueError("_bucket_ids: a 1-D int64 hash")
    if B < 1 or B & (B - 1):
        raise ValueError(f"_bucket_ids: B={B} is not a power of two")
    out = torch.empty(h.shape[0], dtype=torch.int32, device=h.device)
    lib = _lib.load()
    with torch.cuda.device(h.device):
        _lib.check(lib.pf_bucket_ids(_lib.ptr(h.contiguous()), h.shape[0], B, _lib.ptr(out),
                                     _lib.stream_of(h.device)))
    return out


def _row_bytes(t: torch.Tensor) -> int:
    return t.element_size() * prod(t.shape[1:])


def _partition_outputs(outs: Sequence[torch.Tensor], bucket: torch.Tensor,
                       counts: Sequence[int], counts_t: torch.Tensor
                       ) -> List[PartitionedTensor]:
    B = len(counts)
    if B < 1 or B > _B_MAX or B & (B - 1):
        raise ValueError(f"_partition_outputs: B={B} buckets, a power of two <= {_B_MAX}")
    _lib.require_cuda(bucket, "the bucket ids")
    if bucket.dtype != torch.int32 or bucket.dim() != 1:
        raise ValueError("_partition_outputs: the bucket ids are 1-D int32")
    n = bucket.shape[0]
    for t in outs:
        _lib.require_cuda(t, "a pass-1 output")
        if t.dim() < 1 or t.shape[0] != n:
            raise ValueError(f"_partition_outputs: an output of {tuple(t.shape)} rows for "
                             f"{n} bucket ids")
    counts = [int(c) for c in counts]
    dev = bucket.device
    lib = _lib.load()
    host = _host_buffer()
    bucket = bucket.contiguous()
    result: List[Optional[PartitionedTensor]] = [None] * len(outs)
    # groups of at most six outputs and four row widths, a kernel call each
    order = list(range(len(outs)))
    while order:
        group, widths = [], []
        for i in list(order):
            rb = _row_bytes(outs[i])
            if len(group) == _MAX_COLS or (rb not in widths and len(widths) == _MAX_WIDTHS):
                continue
            group.append(i)
            if rb not in widths:
                widths.append(rb)
            order.remove(i)
        src = [outs[i].contiguous() for i in group]
        dst = [torch.empty((n + B * (rows_per_unit(_row_bytes(t)) - 1),) + tuple(t.shape[1:]),
                           dtype=t.dtype, device=dev) for t in src]
        rbs = _lib.i32_array([_row_bytes(t) for t in src])
        ws_bytes = lib.pf_partition_by_bucket_workspace_bytes(n, B)
        ws = _lib.workspace(ws_bytes, dev)
        with torch.cuda.device(dev):
            _lib.check(lib.pf_partition_by_bucket(
                _lib.void_array([t.data_ptr() for t in src]),
                _lib.void_array([t.data_ptr() for t in dst]),
                ctypes.cast(rbs, ctypes.c_void_p), len(src), _lib.ptr(bucket), n, B,
                ctypes.c_void_p(host.data_ptr()), _lib.ptr(ws), ws_bytes,
                _lib.stream_of(dev)))
        got = host[:B].tolist()
        totals = host[B:B + len(widths)].tolist()
        bad = int(host[B + _MAX_WIDTHS + 1])
        if bad:
            raise RuntimeError(f"_partition_outputs: {bad} bucket ids outside [0, {B})")
        if got != counts:
            raise RuntimeError("_partition_outputs: the bucket counts given disagree with "
                               "the bucket ids")
        for i, t in zip(group, dst):
            total = totals[widths.index(_row_bytes(outs[i]))]
            result[i] = PartitionedTensor(t[:total], counts)
    return result  # type: ignore[return-value]


__all__ = ["_mix64", "_bucket_ids", "_partition_outputs"]
