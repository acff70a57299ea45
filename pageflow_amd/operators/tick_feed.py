"""The feed of a partitioned join: how the reference's streaming engine cuts
one join into core calls (operators/hashjoin.py ``tqp_join``, :581-633 and
:856-976), for sides already resident in device memory.

1. B, the same for both sides, is the smallest power of two whose predicted
   largest bucket pair fits the partition budget (``partition_count`` over
   each side's rows, bytes a row and distinct keys; a side proven unique on
   its key has one row a key, and the other side of such a join is taken to
   draw its keys from it), as ``join_partition_count`` (:581-633) derives it.
2. Each side is partitioned in sort ticks of ``tick_rows`` rows by
   ``partition`` (the padded 512-byte-unit bucket layout, gidx = the tick's
   first global row + the row), as ``sort_kernel`` (:726-740) calls it.
3. The buckets of every tick are collected bucket-major (bucket 0 of every
   tick, then bucket 1, ...), the sort pipe's partition-major order.
4. Whole bucket pairs are packed into core calls up to half the core budget
   (``pack_core_calls``, :567-578 and :872-884): matches never cross a
   bucket, so a call's rows are a run of buckets of both sides.
5. Every call with rows runs ``join_kernel_new`` on its buckets' rows with
   their per-bucket counts (``bucket_counts``) and the two gidx columns, as
   ``join_wrapper`` (:962-988) does; the calls' outputs are concatenated.

The data movement of steps 3-5 (the reference's pages, codecs and host
transfers) is out of scope: here a call's rows are copied out of the padded
buckets on the device.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple

import torch

from ..layout import row_bytes
from .hashjoin import join_kernel_new
from .partition_sort import partition
from .partitions import JOIN_B_MAX, pack_core_calls, partition_count

LEFT, RIGHT = 0, 1
PARTITION_BUDGET_B = 512 << 20     # constants.py:76, the default partition budget


@dataclass
class SideFeed:
    """One side after its sort ticks: per column id, the bucket-major rows
    (padding removed), the bucket-major gidx, and the rows of each bucket."""
    columns: Dict[int, torch.Tensor]
    gidx: torch.Tensor
    bucket_rows: List[int]


def join_partition_count(sides: Sequence[Tuple[int, int]], unique_on_key: Sequence[bool],
                         budget_b: int = PARTITION_BUDGET_B) -> Tuple[int, float, list]:
    """(B, predicted largest pair bytes, the sides' (rows, row bytes, keys))
    for two sides of (rows, row bytes incl. the gidx's 8 B)."""
    est = []
    for s in (LEFT, RIGHT):
        rows, rb = sides[s]
        keys = rows
        other_rows = sides[1 - s][0]
        if not unique_on_key[s] and unique_on_key[1 - s] and other_rows > 0:
            keys = min(rows, other_rows)
        est.append((rows, rb, keys))
    b, pred = partition_count(est, budget_b, JOIN_B_MAX)
    return b, pred, est


def feed_side(columns: Dict[int, torch.Tensor], key_id: int, B: int,
              tick_rows: int) -> SideFeed:
    """Partitions a side's columns (the key among them) tick by tick and
    collects the buckets bucket-major."""
    ids = sorted(columns)
    key_pos = ids.index(key_id)
    n = columns[key_id].shape[0]
    ticks = []
    for start in range(0, max(n, 1), tick_rows):
        end = min(n, start + tick_rows)
        cols = [columns[c][start:end] for c in ids]
        ticks.append(partition(cols, key_pos, None, start, B))
    bucket_rows = [sum(t[0].partition_row_counts[b] for t in ticks) for b in range(B)]
    outs = {}
    for j, c in enumerate(ids + [None]):   # None: the gidx output
        per_tick = [t[j].buckets() for t in ticks]
        pieces = [buckets[b] for b in range(B) for buckets in per_tick]
        outs[c] = torch.cat(pieces) if pieces else ticks[0][j].tensor[:0]
    gidx = outs.pop(None)
    return SideFeed(outs, gidx, bucket_rows)


def streamed_join(left: Dict[int, torch.Tensor], right: Dict[int, torch.Tensor], left_key: int,
                  right_key: int, condition_list, join_type: str,
                  output_keys: Sequence[int] = (), *, unique_key_sets=([], []),
                  partition_budget_b: int = PARTITION_BUDGET_B,
                  core_budget_b: int = 256 << 20, tick_rows: int = 1 << 24,
                  name: str = "join"):
    """The join of two resident sides as the streaming engine runs it: the
    output key columns and the (left gidx, right gidx) of every result row,
    each concatenated over the core calls in call order; and the schedule
    (B, each call's bucket range)."""
    unique_on_key = [any(k and set(k) <= {key} for k in sets)
                     for key, sets in ((left_key, unique_key_sets[0]),
                                       (right_key, unique_key_sets[1]))]
    side_b = [sum(row_bytes(t) for t in side.values()) + 8 for side in (left, right)]
    side_n = [left[left_key].shape[0], right[right_key].shape[0]]
    B, _, _ = join_partition_count(list(zip(side_n, side_b)), unique_on_key,
                                   partition_budget_b)
    fl = feed_side(left, left_key, B, tick_rows)
    fr = feed_side(right, right_key, B, tick_rows)
    ends, _ = pack_core_calls(fl.bucket_rows, fr.bucket_rows, side_b, core_budget_b // 2)
    lstart = [0]
    rstart = [0]
    for b in range(B):
        lstart.append(lstart[-1] + fl.bucket_rows[b])
        rstart.append(rstart[-1] + fr.bucket_rows[b])
    results: List[List[torch.Tensor]] = []
    calls = []
    first = 0
    eq = ([left_key], [right_key])
    for end in ends:
        lc = fl.bucket_rows[first:end]
        rc = fr.bucket_rows[first:end]
        calls.append((first, end))
        if sum(lc) + sum(rc) > 0:      # an all-empty call is never run
            la, lb = lstart[first], lstart[end]
            ra, rb = rstart[first], rstart[end]
            tl = {c: t[la:lb] for c, t in fl.columns.items()}
            tr = {c: t[ra:rb] for c, t in fr.columns.items()}
            results.append(join_kernel_new(
                tl, tr, left_key, right_key, condition_list, list(output_keys), True,
                join_type, fl.gidx[la:lb], fr.gidx[ra:rb], bucket_counts=(lc, rc),
                equality_keys=eq, unique_key_sets=unique_key_sets, name=name))
        first = end
    if not results:
        dev = left[left_key].device
        empty = torch.empty(0, dtype=torch.int64, device=dev)
        cols = [t[:0] for t in (left.get(k, right.get(k)) for k in output_keys)]
        return cols + [empty, empty.clone()], (B, calls)
    joined = [torch.cat([r[i] for r in results]) for i in range(len(results[0]))]
    return joined, (B, calls)
