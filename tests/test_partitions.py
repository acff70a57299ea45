"""The join's partition count B and the bucket-pair packing (host logic of
``pageflow_amd.operators.partitions``), against known answers.

The known answers are data taken from the reference's
tests/operators/test_partitions.py: the balls-into-bins maximum loads, the
B of joins of given sizes, key counts and budgets, and one packing.  The
``smallest power of two`` test checks the rule itself: B is the least power
of two whose predicted largest bucket pair fits the budget, else the one
with the smallest prediction.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from pageflow_amd.operators.partitions import (B_MIN, JOIN_B_MAX, max_load, pack,
                                               pack_core_calls, partition_count)

MiB = 1 << 20

# (balls, bins) -> the fullest bin's expected load
MAX_LOADS = [
    ((25, 32), 4), ((25, 64), 3), ((25, 128), 3), ((5, 4), 3), ((1, 2), 1), ((1, 128), 1),
    ((0, 8), 0), ((7, 1), 7), ((1e6, 1), 1e6),
]

# (sides as (rows, row bytes[, distinct keys]), budget MiB) -> B
JOIN_BS = [
    ([(1000, 16), (10, 16)], 512, 1),
    ([(0, 16), (0, 24)], 512, 1),
    ([(512 * MiB // 32 + 1, 16), (512 * MiB // 32 + 1, 16)], 512, 2),
    ([(100_000_000, 16), (100_000_000, 16)], 512, 8),
    ([(100_000_000, 16), (100_000_000, 16)], 256, 16),
    ([(100_000_000, 16), (100_000_000, 16)], 1024, 4),
    ([(4 * (512 * MiB // 16), 16)], 512, 8),
    ([(4 * (512 * MiB // 16) - 4 * (512 * MiB // 16) // 100, 16)], 512, 4),
    ([(6_000_000_000, 16), (1_500_000_000, 16)], 512, 256),
    ([(24_000_000_000, 16), (6_000_000_000, 16)], 512, JOIN_B_MAX),
    ([(25, 16), (100_000_000, 16, 25)], 512, 8),
    ([(25, 16), (800_000_000, 16, 25)], 512, 256),
    ([(1, 16, 1), (80_000_000, 16, 5)], 512, 16),
    ([(5, 16), (150_000_000, 16, 5)], 512, 128),
]


@pytest.mark.parametrize("args,want", MAX_LOADS)
def test_max_load_known_answers(args, want):
    assert max_load(*args) == want


def test_max_load_heavy_regime_and_its_seam():
    mean = 4096
    assert max_load(mean * 128, 128) == pytest.approx(mean + math.sqrt(2 * mean * math.log(128)))
    assert max_load(4095, 128) == pytest.approx(max_load(4096, 128), rel=0.05)


@pytest.mark.parametrize("k", range(len(JOIN_BS)))
def test_join_partition_count_known_answers(k):
    sides, budget_mib, want = JOIN_BS[k]
    assert B_MIN == 1
    assert partition_count(sides, budget_mib * MiB, JOIN_B_MAX)[0] == want


def _predicted_pair_bytes(sides, b) -> float:
    """The rule's prediction at B = b: per side, its rows per key times the
    fullest bucket's keys times its row bytes, summed over the two sides."""
    out = 0.0
    for side in sides:
        rows, row_b = side[:2]
        keys = rows if len(side) < 3 else min(rows, side[2])
        out += 0.0 if not (rows and keys) else (rows / keys) * max_load(keys, b) * row_b
    return out


@pytest.mark.parametrize("rows,budget_mib,keys", [
    (0, 512, None), (1, 1, None), (3_000_000, 1, None), (600_000_000, 512, None),
    (600_000_000, 64, 25), (600_000_000, 2, 10_000)])
def test_smallest_power_of_two_that_fits(rows, budget_mib, keys):
    probe = (rows // 4, 16) if keys is None else (rows // 4, 16, keys)
    sides = [(rows, 24), probe]
    budget = budget_mib * MiB
    pred = {1 << e: _predicted_pair_bytes(sides, 1 << e) for e in range(10)}
    fitting = sorted(b for b, p in pred.items() if p <= budget)
    want = fitting[0] if fitting else min(pred, key=lambda b: (pred[b], b))
    assert partition_count(sides, budget, JOIN_B_MAX)[0] == want


def test_packing_keeps_pieces_whole():
    edges = np.array([0, 10, 20, 20, 45, 50])        # cumulative bytes of five pieces
    assert pack(edges, 25).tolist() == [3, 4, 5]
    assert pack(edges, 5).tolist() == [1, 2, 3, 4, 5]   # a piece over budget goes alone
    ends, pair_bytes = pack_core_calls([1, 2, 0, 4], [1, 0, 0, 2], (8, 8), 32)
    assert pair_bytes.tolist() == [16, 16, 0, 48] and ends == [3, 4]
