"""A join's hash-partition count B from the partition budget, and the
packing of whole bucket pairs into core calls.

Mirrors the reference's ``operators/partitions.py`` (``max_load`` :29-51,
``partition_count`` :54-85) and the greedy packer of
``IO/pipeline/tick_cutter.pack`` (:62-76) used by
``operators/hashjoin.pack_core_calls`` (:567-578).  The rule has to be the
reference's (B is observable through the outputs' bucket counts), so it is
the same rule: the smallest power of two B whose largest bucket pair is
predicted to fit the budget, predicting a side's fullest bucket as rows per
key times the fullest bucket's key count at about 90% confidence.

Host-only (stdlib + numpy); tested without a GPU.
"""
from __future__ import annotations

import math
from typing import Sequence, Tuple

import numpy as np

B_MIN = 1
# The legacy partition entry point's bucket cap (pf_partition_sort); the new
# radix partition (pf_radix_partition) goes to 2^12.
JOIN_B_MAX = 512
RADIX_B_MAX = 4096


def max_load(m: float, n: int) -> float:
    """The fullest of ``n`` bins' load, at about 90%, for ``m`` balls thrown
    uniformly.  For m/n >= 32 the Raab-Steger bound m/n + sqrt(2 (m/n) ln n);
    below, the largest k such that the expected number of bins holding at
    least k balls is still >= 0.1 (binomial tail, union bound)."""
    if m <= 0:
        return 0.0
    if n <= 1:
        return float(m)
    mean = m / n
    if mean >= 32:
        return mean + math.sqrt(2.0 * mean * math.log(n))
    balls = int(math.ceil(m))
    p = 1.0 / n
    q = 1.0 - p
    # walk k upward: cdf = P(X <= k - 1), term = P(X = k - 1)
    term = q ** balls
    cdf = 0.0
    for k in range(1, balls + 1):
        cdf += term
        if n * (1.0 - cdf) < 0.1:
            return float(k - 1)
        term *= (balls - k + 1) / k * (p / q)
    return float(balls)


def partition_count(parts: Sequence[tuple], budget_b: int,
                    b_max: int) -> Tuple[int, float]:
    """Smallest power of two B in [1, b_max] whose predicted largest unit
    (the sum over ``parts`` of each part's fullest bucket, in bytes) fits
    ``budget_b``, and that prediction; when none fits, the smallest B with
    the least prediction.  ``parts``: ``(rows, row_bytes[, distinct
    keys])``, keys defaulting to rows."""
    if b_max < B_MIN or b_max & (b_max - 1):
        raise ValueError(f"b_max must be a power of two >= 1, got {b_max}")
    best_b, best = B_MIN, math.inf
    b = B_MIN
    while b <= b_max:
        unit = 0.0
        for part in parts:
            rows, rb = part[0], part[1]
            keys = min(part[2], rows) if len(part) > 2 else rows
            if rows > 0 and keys > 0:
                unit += rows / keys * max_load(keys, b) * rb
        if unit <= budget_b:
            return b, unit
        if unit < best:
            best_b, best = b, unit
        b <<= 1
    return best_b, best


def pack(points: np.ndarray, limit: int) -> np.ndarray:
    """Greedy packing of the pieces between consecutive ``points`` (the
    first is the start): each group takes as many pieces as fit in
    ``limit`` (at least one).  Returns each group's end index in
    ``points``; empty pieces join the group before them."""
    points = np.asarray(points)
    ends = []
    i, last = 0, len(points) - 1
    while i < last:
        j = int(np.searchsorted(points, points[i] + limit, side="right")) - 1
        j = max(j, i + 1)
        ends.append(j)
        i = j
    return np.asarray(ends, dtype=np.int64)


def pack_core_calls(rows_left, rows_right, side_row_bytes, budget_b):
    """Whole bucket pairs per core call, up to ``budget_b`` bytes each
    (at least one bucket a call).  Returns each call's end bucket
    (exclusive) and every pair's bytes."""
    pair_b = (np.asarray(rows_left, dtype=np.int64) * side_row_bytes[0]
              + np.asarray(rows_right, dtype=np.int64) * side_row_bytes[1])
    ends = pack(np.concatenate(([0], np.cumsum(pair_b))), budget_b).tolist()
    return ends, pair_b
