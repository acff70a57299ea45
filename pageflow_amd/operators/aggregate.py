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

And the grouping of ``KeyedAggregate.pass1`` (aggregate.py:1511-1563) for
int64 keys, with SUM and COUNT of one int64 column
(operator_utils.py:342-399): ``keyed_pass1`` groups a chunk's rows by the
exact tuple of one or two key columns on the device (``pf_group_by_keys``)
and returns pass 1's ``[hash] + keys + values`` and buckets.  The hash of a
group is the reference's linear grouping hash (``_make_key_coeffs``,
``_hash_key_columns``: sum of key x coefficient, wrapping int64), and the
groups come in its ascending order, as ``torch.unique`` orders them there.
Keys of other types (strings, narrow integers packed into words) and other
aggregate functions are outside this mirror.
"""
from __future__ import annotations

import ctypes
import os
from math import prod
from typing import List, Optional, Sequence, Tuple

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


# The seed of the grouping hash's coefficients (the reference's
# _COEFF_BASE_SEED, aggregate.py:526, overridable by the same variable).
_COEFF_BASE_SEED = int(os.environ.get('TQP_HASH_SEED', '2029')) & 0x7FFF_FFFF_FFFF_FFFF


def _make_key_coeffs(key_tensors: Sequence[torch.Tensor], device) -> List[torch.Tensor]:
    """One coefficient vector per key column, one coefficient per int64 word
    of its row, drawn in [2, 10^7) from a generator reseeded to the fixed
    seed at every call, so they depend only on the seed and the key shapes
    (the reference's aggregate.py:529-548).  An int64 key of one element
    takes one coefficient."""
    g = torch.Generator(device=device)
    g.manual_seed(_COEFF_BASE_SEED)
    out = []
    for col in key_tensors:
        width = col.shape[1] if col.dim() == 2 else 1
        words = (width * col.element_size() + 7) // 8
        out.append(torch.randint(2, 10_000_000, (words,), dtype=torch.int64, device=device,
                                 generator=g))
    return out


def _key_vector(t: torch.Tensor, what: str) -> torch.Tensor:
    _lib.require_cuda(t, what)
    if t.dim() == 2 and t.shape[1] == 1:
        t = t.reshape(-1)
    if t.dtype != torch.int64 or t.dim() != 1:
        raise ValueError(f"keyed_pass1: {what} must be a 1-D int64 column (or (n, 1))")
    return t.contiguous()


def group_by_keys(keys: Sequence[torch.Tensor], values: Optional[torch.Tensor],
                  coeffs: Sequence[int], B: int = 1):
    """Groups of the rows by the exact tuple of ``keys`` (one or two int64
    columns): (hash, key columns, SUM, COUNT, bucket ids, bucket counts),
    groups in ascending hash order.  SUM/COUNT are None without ``values``;
    bucket ids None at B = 1.  Raises RuntimeError when two groups share a
    hash (the reference would have merged them)."""
    if len(keys) not in (1, 2) or len(coeffs) != len(keys):
        raise ValueError("keyed_pass1: one or two key columns, one coefficient each")
    ks = [_key_vector(k, f"key {i}") for i, k in enumerate(keys)]
    n = ks[0].numel()
    dev = ks[0].device
    if any(k.numel() != n or k.device != dev for k in ks):
        raise ValueError("keyed_pass1: key columns of one length on one device")
    if values is not None:
        values = _key_vector(values, "the value column")
        if values.numel() != n or values.device != dev:
            raise ValueError("keyed_pass1: the value column must match the keys")
    lib = _lib.load()
    ws_bytes = lib.pf_group_by_keys_workspace_bytes(n, len(ks), B)
    ws = _lib.workspace(ws_bytes, dev)
    m = max(n, 1)
    hash_out = torch.empty(m, dtype=torch.int64, device=dev)
    key_out = [torch.empty(m, dtype=torch.int64, device=dev) for _ in ks]
    sums = torch.empty(m, dtype=torch.int64, device=dev) if values is not None else None
    cnts = torch.empty(m, dtype=torch.int64, device=dev) if values is not None else None
    bucket = torch.empty(m, dtype=torch.int32, device=dev) if B > 1 else None
    host = torch.zeros(2 + B, dtype=torch.int64, pin_memory=True)
    c = (ctypes.c_int64 * 2)(*[int(x) for x in coeffs], *([0] * (2 - len(coeffs))))
    with torch.cuda.device(dev):
        _lib.check(lib.pf_group_by_keys(
            _lib.ptr(ks[0]), _lib.ptr(ks[-1]), len(ks), n, ctypes.cast(c, ctypes.c_void_p),
            _lib.ptr(values), B, _lib.ptr(hash_out), _lib.ptr(key_out[0]),
            _lib.ptr(key_out[-1]), _lib.ptr(sums), _lib.ptr(cnts), _lib.ptr(bucket),
            ctypes.c_void_p(host.data_ptr()), _lib.ptr(ws), ws_bytes, _lib.stream_of(dev)))
    words = host.tolist()
    G, repeats, counts = words[0], words[1], words[2:]
    if repeats:
        raise RuntimeError(f"keyed_pass1: {repeats} groups share a grouping hash with another; "
                           f"the coefficients do not separate these keys")
    cut = (lambda t: t[:G] if t is not None else None)
    return (hash_out[:G], [cut(k) for k in key_out], cut(sums), cut(cnts), cut(bucket),
            counts)


def keyed_pass1(columns, key_ids: Sequence[int], value_id: Optional[int], B: int,
                coeffs: Optional[Sequence[int]] = None):
    """One chunk's partials as ``KeyedAggregate.pass1`` returns them
    (aggregate.py:1511-1563), for int64 GROUP BY keys and the results
    SUM(value) and COUNT(value): ``[hash, key..., sum, count]`` (without a
    value column, ``[hash, key...]``) and the buckets, None at B = 1 or with
    no group, else ``(bucket ids int32, counts on the host, counts on the
    GPU)``.  ``coeffs``: the grouping hash's coefficients, by default the
    reference's (``_make_key_coeffs`` on the key columns)."""
    keys = [columns[k] for k in key_ids]
    if coeffs is None:
        coeffs = [int(c[0]) for c in _make_key_coeffs(keys, keys[0].device)]
    values = columns[value_id] if value_id is not None else None
    h, key_out, sums, cnts, bucket, counts = group_by_keys(keys, values, coeffs, B)
    outs = [h] + key_out + ([sums, cnts] if values is not None else [])
    if B == 1 or h.numel() == 0:
        return outs, None
    return outs, (bucket, counts, torch.tensor(counts, dtype=torch.int64, device=h.device))


__all__ = ["_mix64", "_bucket_ids", "_partition_outputs", "_make_key_coeffs", "group_by_keys",
           "keyed_pass1"]
