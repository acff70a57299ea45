"""``partition`` (pf_partition_sort through
``pageflow_amd.operators.partition_sort``) against ``oracle.partition``.

Each case is a tick spec: rows, the column kinds beside the int64 key, the
key's position, B, the global offset, a mask (kept fraction, which columns
arrive compacted, whether the caller passes the kept count), a misaligned
start for the device views and the select-first thresholds.  A case passes
when every output (the columns, then the gidx) has the oracle's bucket
counts and, at each bucket's padded start, exactly the oracle's rows;
padding is not compared.

The case list re-expresses the coverage of the reference's partition tests
(tests/operators/test_partition_sort.py): tile-edge tick sizes, every B,
six columns of four widths, each width, full skew, an offset near 6e9, masks
in every kernel mode, ticks past a thousand tiles, wrong counts that raise
and leave the GPU usable, and the kernel's limits.
"""
from __future__ import annotations

from dataclasses import dataclass, field, replace
from typing import List, Optional

import numpy as np
import pytest
import torch

import oracle
import pageflow_amd.operators.partition_sort as partition_sort
from pageflow_amd.operators.partition_sort import partition
from pageflow_amd.operators.partitions import JOIN_B_MAX

pytestmark = pytest.mark.gpu

# kind -> (numpy dtype, elements a row, low, high)
KINDS = {
    "i64": (np.int64, 1, -(1 << 62), 1 << 62),
    "i32": (np.int32, 1, -(1 << 30), 1 << 30),
    "i16": (np.int16, 1, -(1 << 14), 1 << 14),
    "i8": (np.int8, 1, -128, 128),
    "s25": (np.uint8, 25, 0, 256),
    "u8x2": (np.uint8, 2, 0, 256),
    "u8x3": (np.uint8, 3, 0, 256),
    "u8x4": (np.uint8, 4, 0, 256),
    "u8x8": (np.uint8, 8, 0, 256),
    "i64x2": (np.int64, 2, -(1 << 62), 1 << 62),
}


@dataclass
class Tick:
    n: int
    kinds: List[str] = field(default_factory=lambda: ["i64"])
    B: int = 32
    key_pos: int = 0
    goff: int = 0
    seed: int = 0
    one_key: bool = False            # every row the same key: one bucket holds all
    keep: Optional[float] = None     # mask kept fraction; None: no mask
    compacted: Optional[List[bool]] = None   # per column; None: all compacted
    give_count: bool = False
    offset: int = 0                  # device views start this many rows in
    select_first_below: Optional[float] = None
    raw_select_first_below: Optional[float] = None

    def columns(self) -> List[np.ndarray]:
        rng = np.random.default_rng(self.seed)
        key = (np.full(self.n, 7, np.int64) if self.one_key
               else rng.integers(-(1 << 40), 1 << 40, self.n, dtype=np.int64))
        cols = []
        for kind in self.kinds:
            dt, width, lo, hi = KINDS[kind]
            shape = (self.n,) if width == 1 else (self.n, width)
            cols.append(rng.integers(lo, hi, shape, dtype=dt))
        cols.insert(self.key_pos, key)
        return cols

    def mask(self) -> Optional[np.ndarray]:
        if self.keep is None:
            return None
        rng = np.random.default_rng(self.seed + 7919)
        return rng.random(self.n) < self.keep


def _on_device(a: np.ndarray, offset: int) -> torch.Tensor:
    t = torch.from_numpy(np.ascontiguousarray(a))
    if offset:
        lead = torch.zeros((offset,) + tuple(t.shape[1:]), dtype=t.dtype)
        return torch.cat((lead, t)).cuda()[offset:]
    return t.cuda()


def _buckets_of(pt) -> List[np.ndarray]:
    """The rows of each bucket of a padded output, read at its padded start."""
    host = pt.tensor.cpu().numpy()
    row_b = host.dtype.itemsize * (host.shape[1] if host.ndim == 2 else 1)
    starts, total = oracle.padded_layout(pt.partition_row_counts, row_b)
    assert host.shape[0] == total
    return [host[s:s + c] for s, c in zip(starts, pt.partition_row_counts)]


def run(tick: Tick, monkeypatch=None) -> None:
    if monkeypatch is not None:
        if tick.select_first_below is not None:
            monkeypatch.setattr(partition_sort, "SELECT_FIRST_BELOW", tick.select_first_below)
        if tick.raw_select_first_below is not None:
            monkeypatch.setattr(partition_sort, "SELECT_FIRST_RAW_BELOW",
                                tick.raw_select_first_below)
    cols, mask = tick.columns(), tick.mask()
    given = cols
    if mask is not None:
        packed = tick.compacted if tick.compacted is not None else [True] * len(cols)
        given = [c[mask] if p else c for c, p in zip(cols, packed)]
    n_valid = int(mask.sum()) if (mask is not None and tick.give_count) else None
    outs = partition([_on_device(c, tick.offset) for c in given], tick.key_pos,
                     None if mask is None else _on_device(mask, tick.offset), tick.goff, tick.B,
                     n_valid=n_valid)
    counts, want = oracle.partition(cols, tick.key_pos, mask, tick.goff, tick.B)
    assert len(outs) == len(want) == len(cols) + 1
    for out, want_buckets in zip(outs, want):
        assert out.partition_row_counts == counts
        got_buckets = _buckets_of(out)
        for b in range(tick.B):
            assert np.array_equal(got_buckets[b], want_buckets[b]), f"bucket {b}"


# ----------------------------------------------------------------- cases --

PLAIN = {
    **{f"rows={n}": Tick(n, seed=n) for n in (0, 1, 25, 2047, 2048, 2049, 300_000)},
    **{f"B={b}": Tick(100_000, B=b, goff=12345, seed=b)
       for b in (1, 2, 8, 16, 32, 64, 128, 256, 512)},
    **{f"six-columns-B{b}": Tick(70_001, ["i64", "i32", "i8", "s25", "i64"], B=b, key_pos=2,
                                 seed=3) for b in (64, 512)},
    **{f"width-{k}": Tick(50_000, [k], B=16, seed=11)
       for k in ("i32", "i16", "i8", "s25", "u8x2", "u8x3", "u8x4", "u8x8", "i64x2")},
    "one-bucket-holds-all": Tick(40_000, ["i64", "s25"], B=64, one_key=True, seed=5),
    "goff-near-6e9": Tick(1 << 18, B=64, goff=6_000_000_000 - (1 << 18), seed=6),
    "past-1000-tiles-B16": Tick(5_000_011, ["i32"], B=16, goff=7, seed=96),
    "past-1000-tiles-B512": Tick(5_000_011, ["i32"], B=512, goff=7, seed=96),
}


@pytest.mark.parametrize("name", list(PLAIN))
def test_unmasked_ticks(name):
    run(PLAIN[name])


def _masked_cases():
    cases = {}
    for keep in (0.0, 0.01, 0.3, 0.63, 1.0):
        for below in (0.0, 2.0):
            cases[f"compacted-keep{keep}-below{below}"] = Tick(
                200_003, ["i64", "s25"], goff=777, seed=int(keep * 1000) + 1, keep=keep,
                give_count=True, select_first_below=below)
    for keep in (0.0, 0.095, 0.3, 1.0):
        for below in (0.0, 2.0):
            cases[f"raw-keep{keep}-below{below}"] = Tick(
                150_001, ["i64", "s25", "i16"], B=64, goff=99, seed=int(keep * 1000) + 3,
                keep=keep, compacted=[False] * 4, give_count=True, raw_select_first_below=below)
    for i, packed in enumerate(([False, True, False, True], [True, False, True, False])):
        for below in (0.0, 2.0):
            cases[f"mixed{i}-below{below}"] = Tick(
                90_000, ["i64", "s25", "i32"], B=64, goff=5, seed=41, keep=0.4,
                compacted=packed, give_count=True, raw_select_first_below=below)
    for off in (3, 5):
        for below in (0.0, 2.0):
            base = Tick(64_000, B=32, seed=51, keep=0.2, offset=off)
            cases[f"offset{off}-compacted-below{below}"] = replace(base,
                                                                 select_first_below=below)
            cases[f"offset{off}-raw-below{below}"] = replace(
                base, compacted=[False, False], give_count=True, raw_select_first_below=below)
    for n in (1, 25, 2049):
        for keep in (0.5, 0.0):
            # the count is given: a compacted column that keeps every row is
            # indistinguishable from a raw one
            cases[f"small{n}-keep{keep}"] = Tick(n, B=8, seed=n, keep=keep, give_count=True)
    for b in (16, 512):
        for mode, below, raw_below, packed in (("compacted", 0.0, 0.0, True),
                                               ("select-first", 2.0, 0.0, True),
                                               ("fold", 0.0, 0.0, False),
                                               ("raw-select-first", 0.0, 2.0, False)):
            cases[f"past-1000-tiles-{mode}-B{b}"] = Tick(
                5_000_011, ["i32"], B=b, goff=7, seed=96, keep=0.45, compacted=[packed] * 2,
                give_count=not packed, select_first_below=below,
                raw_select_first_below=raw_below)
    for b in (1, 128, 512):
        for below in (0.0, 2.0):
            for packed in (True, False):
                cases[f"B{b}-below{below}-{'compacted' if packed else 'raw'}"] = Tick(
                    80_000, ["i64", "s25"], B=b, seed=98, keep=0.37, compacted=[packed] * 3,
                    give_count=not packed, select_first_below=below,
                    raw_select_first_below=below)
    return cases


MASKED = _masked_cases()


@pytest.mark.parametrize("name", list(MASKED))
def test_masked_ticks(name, monkeypatch):
    run(MASKED[name], monkeypatch)


def test_counts_buffer_is_reused_across_calls_of_any_B():
    for i, b in enumerate((64, 8, 512, 1, 128, 256, 16)):
        run(Tick(30_000 + i, ["i32"], B=b, goff=i, seed=60 + i))


def test_a_repeated_call_is_identical():
    cols = [_on_device(c, 0) for c in Tick(100_000, seed=7).columns()]
    first, second = partition(cols, 0, None, 0, 64), partition(cols, 0, None, 0, 64)
    for a, b in zip(first, second):
        assert a.partition_row_counts == b.partition_row_counts
        assert all(np.array_equal(x, y) for x, y in zip(_buckets_of(a), _buckets_of(b)))


# ---------------------------------------------------------------- errors --

def test_raw_columns_need_the_kept_count():
    tick = Tick(12_000, seed=31, keep=0.4)
    cols = [_on_device(c, 0) for c in tick.columns()]
    with pytest.raises(ValueError, match="need the mask's kept rows"):
        partition(cols, 0, _on_device(tick.mask(), 0), 0, 16)


def test_compacted_rows_contradicting_the_count_raise():
    tick = Tick(10_000, seed=34, keep=0.5)
    mask = tick.mask()
    packed = [_on_device(c[mask], 0) for c in tick.columns()]
    with pytest.raises(RuntimeError, match="compacted columns hold"):
        partition(packed, 0, _on_device(mask, 0), 0, 32, n_valid=int(mask.sum()) + 1)


@pytest.mark.parametrize("raw_below", [0.0, 2.0])
@pytest.mark.parametrize("error", [-37, 41])
def test_a_wrong_kept_count_raises_and_the_gpu_stays_usable(raw_below, error, monkeypatch):
    tick = Tick(100_003, ["i64", "s25"], seed=91, keep=0.1, compacted=[False] * 3,
                give_count=True, raw_select_first_below=raw_below)
    monkeypatch.setattr(partition_sort, "SELECT_FIRST_RAW_BELOW", raw_below)
    mask = tick.mask()
    with pytest.raises(RuntimeError, match="the mask keeps"):
        partition([_on_device(c, 0) for c in tick.columns()], 0, _on_device(mask, 0), 0, 32,
                  n_valid=int(mask.sum()) + error)
    torch.cuda.synchronize()
    run(tick, monkeypatch)


@pytest.mark.parametrize("below", [0.0, 2.0])
def test_columns_compacted_by_another_mask_raise(below, monkeypatch):
    tick = Tick(100_003, ["i64", "s25"], B=16, seed=93, keep=0.3, select_first_below=below)
    monkeypatch.setattr(partition_sort, "SELECT_FIRST_BELOW", below)
    kept = tick.mask()
    packed = [_on_device(c[kept], 0) for c in tick.columns()]
    for other_keep in (0.2, 0.4):
        other = np.random.default_rng(95).random(tick.n) < other_keep
        with pytest.raises(RuntimeError, match="the mask keeps"):
            partition(list(packed), 0, _on_device(other, 0), 0, 16)
    torch.cuda.synchronize()
    run(tick, monkeypatch)


def _over_limits(case):
    n = 1000
    cols = [_on_device(c, 0) for c in Tick(n, ["i32", "i16", "i8", "s25"], seed=70).columns()]
    key = cols[0]
    return {
        "five widths": (cols, 32),
        "seven columns": ([key] * 7, 32),
        "int32 key": ([key.to(torch.int32)], 32),
        "B=1024": ([key], 1024),
        "B=24": ([key], 24),
        "B=0": ([key], 0),
    }[case]


@pytest.mark.parametrize("case", ["five widths", "seven columns", "int32 key", "B=1024",
                                  "B=24", "B=0"])
def test_ticks_outside_the_limits_raise(case):
    cols, b = _over_limits(case)
    with pytest.raises(ValueError, match="outside the kernel's limits"):
        partition(cols, 0, None, 0, b)


def test_the_bucket_cap_is_the_joins():
    assert partition_sort._B_MAX == JOIN_B_MAX == 512
