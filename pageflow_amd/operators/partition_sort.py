"""The join sort pipe's hash partition over the gfx950 kernel
(``pf_partition_sort``, include/pageflow_join.h).

Drop-in for the reference's ``operators/partition_sort.py`` (``partition``,
operators/partition_sort.py:68-143): one call puts a tick's transfer set into
the padded per-bucket layout, each non-empty bucket starting on a whole
``lcm(row bytes, 512)``-byte unit.  A row's bucket is the top log2(B) bits of
mix64(key) (bucket 0 for every row at B = 1); rows keep their source order in
a bucket; the global index (``goff`` + the row's pre-mask position) is
computed on the device.

Under a mask a column arrives raw (the mask's rows) or compacted (the kept
rows).  Below a selectivity threshold the kernel first collects the kept
rows' pre-mask ids (select-first): ``SELECT_FIRST_RAW_BELOW`` when a column
is raw, ``SELECT_FIRST_BELOW`` when every column is compacted; above it the
tiles walk the pre-mask rows with the mask as a predicate.
"""
from __future__ import annotations

from typing import List, Optional

import torch

from .. import _lib, tqpkernels
from ..layout import PartitionedTensor, row_bytes
from .partitions import JOIN_B_MAX as _B_MAX

# Kept rows per mask row below which select-first runs (every column
# compacted, then with a raw column).  The reference's crossovers
# (partition_sort.py:39-46); tests move them to force a mode.
SELECT_FIRST_BELOW = 0.5
SELECT_FIRST_RAW_BELOW = 0.3

_MAX_COLS = 6
_MAX_WIDTHS = 4          # row widths, the gidx's 8 B included
_MAX_ROWS = (1 << 31) - 1

_host: Optional[torch.Tensor] = None


def _host_buffer() -> torch.Tensor:
    """The process's pinned read-back buffer: the counts, the padded
    totals, then the kept rows a prologue found."""
    global _host
    if _host is None:
        _host = torch.empty(_B_MAX + _MAX_WIDTHS + 1, dtype=torch.int64, pin_memory=True)
    return _host


def partition(cols: List[torch.Tensor], key_pos: int, mask: Optional[torch.Tensor],
              goff: int, B: int, n_valid: Optional[int] = None) -> List[PartitionedTensor]:
    """``cols``: the tick's transfer set in output order, the int64 join key
    at ``key_pos``; with ``mask`` (the pre-mask mask) each column compacted
    or raw, and ``n_valid`` the mask's kept rows (required when a column is
    raw); ``goff``: the tick's first global row.  Returns one
    ``PartitionedTensor`` per column, then the gidx's, in the padded layout.
    Raises ValueError for a tick outside the kernel's limits and
    RuntimeError when the mask contradicts the counts given."""
    key = cols[key_pos]
    n_pre = mask.shape[0] if mask is not None else key.shape[0]
    widths: List[int] = []
    for t in cols:
        rb = row_bytes(t)
        if rb not in widths:
            widths.append(rb)
    if 8 not in widths:
        widths.append(8)
    if (not 1 <= B <= _B_MAX or B & (B - 1) or len(cols) > _MAX_COLS
            or len(widths) > _MAX_WIDTHS or key.dtype != torch.int64
            or key.dim() != 1 or n_pre > _MAX_ROWS
            or any(t.dim() not in (1, 2) for t in cols)
            or (mask is not None and (mask.dim() != 1 or mask.element_size() != 1))):
        raise ValueError(
            f"partition_sort: the tick is outside the kernel's limits (B={B}, "
            f"{len(cols)} columns of widths {widths}, a {key.dtype} key, {n_pre} rows; "
            f"at most B={_B_MAX}, {_MAX_COLS} columns, {_MAX_WIDTHS} widths with the "
            f"gidx's 8 B, a 1-D int64 key, {_MAX_ROWS} rows, 1-D or 2-D columns, a 1-D "
            f"1-byte mask)")
    select_first = False
    if mask is None:
        comp = [0] * len(cols)
        n_valid = n_pre
    else:
        comp = [int(t.shape[0] != n_pre) for t in cols]
        lengths = {t.shape[0] for t, c in zip(cols, comp) if c}
        if len(lengths) > 1:
            raise ValueError(f"partition_sort: compacted columns of {sorted(lengths)} rows")
        if lengths:
            rows = lengths.pop()
            if n_valid is not None and n_valid != rows:
                raise RuntimeError(f"partition_sort: compacted columns hold {rows} rows, "
                                   f"the mask keeps {n_valid}")
            n_valid = rows
        elif n_valid is None:
            raise ValueError("partition_sort: raw columns under a mask need the mask's "
                             "kept rows")
        mask = mask.contiguous()
        below = SELECT_FIRST_BELOW if all(comp) else SELECT_FIRST_RAW_BELOW
        select_first = n_valid < below * n_pre
    for t in cols:
        _lib.require_cuda(t, "a partition column")
    if mask is not None:
        _lib.require_cuda(mask, "the mask")
    host = _host_buffer()
    outs = tqpkernels.partition_sort([t.contiguous() for t in cols], comp, key_pos, mask,
                                     n_valid, goff, B, select_first, host)
    counts = host[:B].tolist()
    totals = host[B:B + len(widths)].tolist()
    kept = sum(counts)
    found = int(host[B + _MAX_WIDTHS])
    if found < 0:
        found = kept
    if found != n_valid or kept != n_valid:
        raise RuntimeError(f"partition_sort: the mask keeps {found} rows and {kept} were "
                           f"partitioned; the caller gave {n_valid}")
    return [PartitionedTensor(t[:totals[widths.index(row_bytes(t))]], counts) for t in outs]
