"""The padded per-bucket layout of a partition output and its marker type.

Mirrors ``IO/codec/batched_compress.aligned_partition_layout``
(IO/codec/batched_compress.py:27-62), ``IO/pipeline/common.PartitionedTensor``
(IO/pipeline/common.py:22-45) and ``utility/tensor_utils.row_bytes``
(utility/tensor_utils.py:7-10) of the reference: every non-empty bucket starts
on a whole ``lcm(row bytes, 512)``-byte unit, so each bucket's records start
512-byte aligned.
"""
from __future__ import annotations

from math import gcd
from typing import List, Sequence, Tuple

COMPRESSION_ALIGNMENT = 512  # constants.py:11
NULL = -64                   # constants.py:9


def row_bytes(t) -> int:
    """Bytes of one row: element size times the width (1 for 1-D)."""
    return t.element_size() * (t.shape[1] if t.dim() == 2 else 1)


def rows_per_unit(rb: int) -> int:
    """Rows of one padding unit: lcm(rb, 512) / rb."""
    return COMPRESSION_ALIGNMENT // gcd(rb, COMPRESSION_ALIGNMENT)


def aligned_partition_layout(partition_row_counts: Sequence[int], row_b: int,
                             align_unit_b: int) -> Tuple[List[int], List[int], int]:
    """(padded rows per bucket, each bucket's byte start, total padded rows)
    for buckets starting on ``align_unit_b``-byte units (a multiple of
    ``row_b``)."""
    if align_unit_b % row_b:
        raise ValueError(f"align unit {align_unit_b} is not a multiple of {row_b}-byte rows")
    per = align_unit_b // row_b
    padded, starts, at = [], [], 0
    for c in partition_row_counts:
        starts.append(at * row_b)
        p = 0 if c == 0 else -(-int(c) // per) * per
        padded.append(p)
        at += p
    return padded, starts, at


class PartitionedTensor:
    """One partition output: ``tensor`` holds the B buckets in bucket order
    in the padded layout, ``partition_row_counts[b]`` rows each."""
    __slots__ = ("tensor", "partition_row_counts")

    def __init__(self, tensor, partition_row_counts):
        self.tensor = tensor
        self.partition_row_counts = list(partition_row_counts)

    def bucket_starts(self) -> List[int]:
        """Each bucket's first row in ``tensor``."""
        rb = row_bytes(self.tensor)
        unit = rb * COMPRESSION_ALIGNMENT // gcd(rb, COMPRESSION_ALIGNMENT)
        _, starts, _ = aligned_partition_layout(self.partition_row_counts, rb, unit)
        return [s // rb for s in starts]

    def buckets(self):
        """Each bucket's rows (views), in bucket order."""
        return [self.tensor[s:s + c] for s, c in zip(self.bucket_starts(),
                                                     self.partition_row_counts)]
