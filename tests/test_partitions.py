"""The partition count B and the bucket-pair packing (host logic), against
the known answers of the reference's tests/operators/test_partitions.py."""
from __future__ import annotations

import math

import numpy as np
import pytest

from pageflow_amd.operators.partitions import (B_MIN, JOIN_B_MAX, max_load, pack,
                                               pack_core_calls, partition_count)

MiB = 1 << 20


def _join_b(sides, budget):
    return partition_count(sides, budget, JOIN_B_MAX)[0]


def _largest(sides, b):
    total = 0.0
    for side in sides:
        rows, width = side[0], side[1]
        keys = min(side[2], rows) if len(side) > 2 else rows
        if rows and keys:
            total += rows / keys * max_load(keys, b) * width
    return total


def test_max_load_known_answers():
    assert max_load(4096 * 128, 128) == pytest.approx(
        4096 + math.sqrt(2 * 4096 * math.log(128)))
    assert max_load(25, 32) == 4
    assert max_load(25, 64) == max_load(25, 128) == 3
    assert max_load(5, 4) == 3
    assert max_load(1, 2) == max_load(1, 128) == 1
    assert max_load(0, 8) == 0
    assert max_load(7, 1) == 7 and max_load(1e6, 1) == 1e6
    assert max_load(4095, 128) == pytest.approx(max_load(4096, 128), rel=0.05)


def test_small_and_budget_multiples():
    assert B_MIN == 1
    assert _join_b([(1000, 16), (10, 16)], 512 * MiB) == 1
    assert _join_b([(0, 16), (0, 24)], 512 * MiB) == 1
    rows = 512 * MiB // 32 + 1
    assert _join_b([(rows, 16), (rows, 16)], 512 * MiB) == 2
    sides = [(100_000_000, 16), (100_000_000, 16)]
    assert _join_b(sides, 512 * MiB) == 8
    assert _join_b(sides, 256 * MiB) == 16
    assert _join_b(sides, 1024 * MiB) == 4
    rows = 4 * (512 * MiB // 16)
    assert _join_b([(rows, 16)], 512 * MiB) == 8
    assert _join_b([(rows - rows // 100, 16)], 512 * MiB) == 4


def test_cap_and_few_keys():
    assert _join_b([(6_000_000_000, 16), (1_500_000_000, 16)], 512 * MiB) == 256
    assert _join_b([(24_000_000_000, 16), (6_000_000_000, 16)], 512 * MiB) == JOIN_B_MAX
    nation = (25, 16)
    assert _join_b([nation, (100_000_000, 16, 25)], 512 * MiB) == 8
    assert _join_b([nation, (800_000_000, 16, 25)], 512 * MiB) == 256
    assert _join_b([(1, 16, 1), (80_000_000, 16, 5)], 512 * MiB) == 16
    assert _join_b([(5, 16), (150_000_000, 16, 5)], 512 * MiB) == 128


@pytest.mark.parametrize("rows,budget_mb,keys", [
    (0, 512, None), (1, 1, None), (3_000_000, 1, None), (600_000_000, 512, None),
    (600_000_000, 64, 25), (600_000_000, 2, 10_000)])
def test_smallest_fitting_power_of_two(rows, budget_mb, keys):
    sides = [(rows, 24), (rows // 4, 16) if keys is None else (rows // 4, 16, keys)]
    budget = budget_mb * MiB
    b = _join_b(sides, budget)
    powers = [1 << k for k in range(10)]
    fits = [p for p in powers if _largest(sides, p) <= budget]
    if fits:
        assert b == fits[0]
    else:
        least = min(_largest(sides, p) for p in powers)
        assert b == min(p for p in powers if _largest(sides, p) == least)


def test_pack_whole_pieces():
    pts = np.array([0, 10, 20, 20, 45, 50])
    assert pack(pts, 25).tolist() == [3, 4, 5]
    assert pack(pts, 5).tolist() == [1, 2, 3, 4, 5]
    ends, pair = pack_core_calls([1, 2, 0, 4], [1, 0, 0, 2], (8, 8), 32)
    assert pair.tolist() == [16, 16, 0, 48] and ends == [3, 4]
