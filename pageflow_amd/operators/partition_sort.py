"""Hash partition of one sort-pipe tick into the padded bucket layout, on the
gfx950 kernels behind ``pf_partition_sort`` (include/pageflow_join.h).

Stands in for ``partition`` of the reference's operators/partition_sort.py
(:68-143): same signature, outputs, exception types and the message
fragments its tests match; the body is this package's own.

What a call produces, per column and then for the global row index (gidx,
int64): the tick's kept rows grouped by bucket, each bucket's rows in their
source order, each non-empty bucket starting on a whole unit of
``lcm(row bytes, 512)`` bytes (the unwritten padding before it belongs to
no row).  A row's bucket is the top log2(B) bits of mix64 of its int64 key
(every row in bucket 0 at B = 1); its gidx is ``goff`` plus its position
before the mask.

Under a mask each column is either raw (one row per mask row) or already
compacted (one row per kept row).  The kernel either walks every mask row
with the mask as a predicate, or first lists the kept rows (select-first).
Select-first pays below a kept fraction: ``SELECT_FIRST_BELOW`` when every
column is compacted, ``SELECT_FIRST_RAW_BELOW`` when one is raw (the values
the reference measured, operators/partition_sort.py:39-46; tests move them
to force a mode).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional

import torch

from .. import _lib, tqpkernels
from ..layout import PartitionedTensor, row_bytes
from .partitions import JOIN_B_MAX as _B_MAX

SELECT_FIRST_BELOW = 0.5
SELECT_FIRST_RAW_BELOW = 0.3

# What one kernel call takes.
_MAX_COLS = 6
_MAX_WIDTHS = 4          # distinct row widths, counting the gidx's 8 B
_MAX_ROWS = (1 << 31) - 1

# host_out: B counts, the padded total of each width class, then the kept
# rows a mask prologue counted (-1 when none ran).
_host: Optional[torch.Tensor] = None


def _host_buffer() -> torch.Tensor:
    global _host
    if _host is None:
        _host = torch.zeros(_B_MAX + _MAX_WIDTHS + 1, dtype=torch.int64, pin_memory=True)
    return _host


@dataclass
class _Tick:
    """One call's shape: the pre-mask rows, the row-width classes in the
    kernel's numbering (the columns' widths by first appearance, then the
    gidx's 8 B unless already present)."""
    n_pre: int
    widths: List[int]


def _describe(cols: List[torch.Tensor], key_pos: int, mask: Optional[torch.Tensor]) -> _Tick:
    widths: List[int] = []
    for t in cols:
        w = row_bytes(t)
        if w not in widths:
            widths.append(w)
    if 8 not in widths:
        widths.append(8)
    n_pre = int(mask.shape[0]) if mask is not None else int(cols[key_pos].shape[0])
    return _Tick(n_pre, widths)


def _limit_breaches(cols, key: torch.Tensor, mask, tick: _Tick, B: int) -> List[str]:
    """Each limit of the kernel the call breaks, in words."""
    out = []
    if not (1 <= B <= _B_MAX) or (B & (B - 1)):
        out.append(f"B={B} is not a power of two in [1, {_B_MAX}]")
    if len(cols) > _MAX_COLS:
        out.append(f"{len(cols)} columns, at most {_MAX_COLS}")
    if len(tick.widths) > _MAX_WIDTHS:
        out.append(f"row widths {tick.widths}, at most {_MAX_WIDTHS} with the gidx's 8 B")
    if key.dtype != torch.int64 or key.dim() != 1:
        out.append(f"a {key.dtype} key of {key.dim()} dimensions, not a 1-D int64 key")
    if tick.n_pre > _MAX_ROWS:
        out.append(f"{tick.n_pre} rows, at most {_MAX_ROWS}")
    bad_dims = [t.dim() for t in cols if t.dim() not in (1, 2)]
    if bad_dims:
        out.append(f"columns of {bad_dims} dimensions, only 1-D or 2-D")
    if mask is not None and (mask.dim() != 1 or mask.element_size() != 1):
        out.append("a mask that is not one byte a row in one dimension")
    return out


def _mask_plan(cols: List[torch.Tensor], mask: Optional[torch.Tensor], n_pre: int,
               n_valid: Optional[int]):
    """(per-column compacted flags, the kept rows, select-first?)."""
    if mask is None:
        return [0] * len(cols), n_pre, False
    comp = [int(t.shape[0] != n_pre) for t in cols]
    compacted_rows = sorted({int(t.shape[0]) for t, c in zip(cols, comp) if c})
    if len(compacted_rows) > 1:
        raise ValueError(f"partition_sort: the compacted columns disagree on their row count "
                         f"({compacted_rows})")
    if compacted_rows:
        held = compacted_rows[0]
        if n_valid is not None and n_valid != held:
            raise RuntimeError(f"partition_sort: compacted columns hold {held} rows, but the "
                               f"caller says the mask keeps {n_valid}")
        n_valid = held
    elif n_valid is None:
        raise ValueError("partition_sort: with only raw columns, n_valid is required: the "
                         "kernel would need the mask's kept rows to size its outputs")
    threshold = SELECT_FIRST_BELOW if all(comp) else SELECT_FIRST_RAW_BELOW
    return comp, n_valid, n_valid < threshold * n_pre


def partition(cols: List[torch.Tensor], key_pos: int, mask: Optional[torch.Tensor],
              goff: int, B: int, n_valid: Optional[int] = None) -> List[PartitionedTensor]:
    """``cols``: the tick's transfer set, the int64 join key at ``key_pos``.
    ``mask``: the pre-mask row mask (one byte a row) or None; under it each
    column is raw or compacted, and ``n_valid`` (the mask's kept rows) must
    be given when no column is compacted.  ``goff``: the tick's first
    global row.  Returns a ``PartitionedTensor`` per column and one for the
    gidx, in the padded layout.  ValueError: the tick is outside the
    kernel's limits or under-specified; RuntimeError: the mask and the
    counts given disagree."""
    key = cols[key_pos]
    tick = _describe(cols, key_pos, mask)
    breaches = _limit_breaches(cols, key, mask, tick, B)
    if breaches:
        raise ValueError("partition_sort: the tick is outside the kernel's limits: " +
                         "; ".join(breaches))
    comp, n_valid, select_first = _mask_plan(cols, mask, tick.n_pre, n_valid)
    for t in cols:
        _lib.require_cuda(t, "a partition column")
    if mask is not None:
        _lib.require_cuda(mask, "the mask")
        mask = mask.contiguous()
    host = _host_buffer()
    outs = tqpkernels.partition_sort([t.contiguous() for t in cols], comp, key_pos, mask,
                                     n_valid, goff, B, select_first, host)
    words = host.tolist()
    counts = words[:B]
    padded_total = dict(zip(tick.widths, words[B:B + len(tick.widths)]))
    partitioned = sum(counts)
    prologue = words[B + _MAX_WIDTHS]
    seen = partitioned if prologue < 0 else prologue
    if seen != n_valid or partitioned != n_valid:
        # the kernel stayed inside the n_valid-sized outputs; the call fails
        raise RuntimeError(f"partition_sort: the mask keeps {seen} rows ({partitioned} "
                           f"partitioned), not the {n_valid} the caller gave")
    return [PartitionedTensor(t[:padded_total[row_bytes(t)]], counts) for t in outs]
