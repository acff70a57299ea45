"""Zero-copy reads of pinned host columns (pf_zero_copy_*, pageflow_amd
.zero_copy_reader): the gather of late materialisation and the masked read
equal the oracle's on columns cut into segments of uneven sizes, for every
element type, row width and row-space offset, with null rows, below and
above the size from which the indices are sorted; the reference test's own
case; rows outside every segment and a wrong mask count raise instead of
reading."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from pageflow_amd import zero_copy_reader as zc

pytestmark = pytest.mark.gpu

NULL = -64


def _pinned_parts(x: np.ndarray, sizes):
    parts, a = [], 0
    for s in sizes:
        t = torch.empty((s,) + x.shape[1:], dtype=torch.from_numpy(x[:1]).dtype,
                        pin_memory=True)
        t.copy_(torch.from_numpy(x[a:a + s]))
        parts.append(t)
        a += s
    return parts


def _column(rng, n, width, dtype):
    if np.issubdtype(dtype, np.floating):
        return rng.standard_normal((n, width)).astype(dtype)
    info = np.iinfo(dtype)
    return rng.integers(info.min // 2, info.max // 2, size=(n, width)).astype(dtype)


def _sizes(rng, n, k):
    cuts = np.sort(rng.choice(np.arange(1, n), k - 1, replace=False))
    return np.diff(np.concatenate(([0], cuts, [n]))).tolist()


@pytest.mark.parametrize("dtype", [np.int64, np.int32, np.int16, np.int8, np.uint8, np.float32,
                                   np.float64])
@pytest.mark.parametrize("width", [1, 3])
@pytest.mark.parametrize("n_out", [5000, 300_000])
def test_rearrange_matches_oracle(dtype, width, n_out):
    rng = np.random.default_rng([np.dtype(dtype).itemsize, int(np.dtype(dtype).kind == "f"),
                                 width, n_out])
    n = 200_000
    x = _column(rng, n, width, dtype)
    sizes = _sizes(rng, n, 7)
    parts = _pinned_parts(x, sizes)
    first = 1000
    ptrs, rows = zc.segments(parts, first_row=first)
    lo, hi = first + 12_345, first + 187_654
    idx = rng.integers(0, hi - lo, n_out)
    idx[::53] = NULL
    got = zc.zero_copy_rearrange(ptrs, rows, torch.from_numpy(idx).cuda(), lo, hi,
                                 torch.from_numpy(x[:1]).dtype, width)
    want = oracle.zero_copy_rearrange([x[a:b] for a, b in zip(np.cumsum([0] + sizes[:-1]),
                                                              np.cumsum(sizes))],
                                      first, idx, lo)
    assert got.shape == (n_out, width)
    assert np.array_equal(got.cpu().numpy(), want)


def test_rearrange_reference_case():
    # tests/IO/test_scan_regions.py:187-201: 300K rows of width 2, a slice
    # [70001, 230007), 50K indices with every 97th null
    x = torch.arange(600_000, dtype=torch.int64).reshape(300_000, 2) * 7 - 11
    parts = _pinned_parts(x.numpy(), [65536] * 4 + [300_000 - 4 * 65536])
    ptrs, rows = zc.segments(parts)
    g = torch.Generator().manual_seed(5)
    lo, hi = 70_001, 230_007
    idx = torch.randint(0, hi - lo, (50_000,), generator=g)
    idx[::97] = NULL
    want = x[lo:hi][idx.clamp(min=0)]
    want[idx == NULL] = NULL
    got = zc.zero_copy_rearrange(ptrs, rows, idx.cuda(), lo, hi, torch.int64, 2)
    assert torch.equal(got.cpu(), want)


@pytest.mark.parametrize("dtype", [np.int64, np.int32, np.uint8, np.float64])
@pytest.mark.parametrize("width", [1, 2])
def test_mask_matches_oracle(dtype, width):
    rng = np.random.default_rng(7 + width)
    n = 300_000
    x = _column(rng, n, width, dtype)
    sizes = _sizes(rng, n, 5)
    parts = _pinned_parts(x, sizes)
    ptrs, rows = zc.segments(parts)
    lo, hi = 123_457, 291_011
    mask = rng.random(hi - lo) < 0.3
    k = int(mask.sum())
    tdt = torch.from_numpy(x[:1]).dtype
    result = torch.empty((k, width) if width > 1 else (k,), dtype=tdt, device="cuda")
    zc.zero_copy_mask(ptrs, rows, torch.from_numpy(mask).cuda(), result, lo, hi)
    want = x[lo:hi][mask]
    got = result.cpu().numpy().reshape(k, width)
    assert np.array_equal(got, want)


def test_rows_outside_every_segment_raise_without_being_read():
    x = np.arange(1000, dtype=np.int64).reshape(-1, 1)
    parts = _pinned_parts(x, [400, 600])
    ptrs, rows = zc.segments(parts, first_row=100)
    idx = torch.tensor([0, 5, 1000, -3, NULL], dtype=torch.int64, device="cuda")
    with pytest.raises(RuntimeError, match="outside every segment"):
        zc.zero_copy_rearrange(ptrs, rows, idx, 100, 1100, torch.int64, 1)
    # a row space read past its last segment
    mask = torch.ones(1000, dtype=torch.bool, device="cuda")
    result = torch.empty(1000, dtype=torch.int64, device="cuda")
    with pytest.raises(RuntimeError, match="outside every segment"):
        zc.zero_copy_mask(ptrs, rows, mask, result, 150, 1150)


def test_mask_count_and_argument_checks():
    x = np.arange(100, dtype=np.int32).reshape(-1, 1)
    parts = _pinned_parts(x, [100])
    ptrs, rows = zc.segments(parts)
    mask = torch.zeros(100, dtype=torch.bool, device="cuda")
    mask[::10] = True
    with pytest.raises(RuntimeError, match="mismatch"):
        zc.zero_copy_mask(ptrs, rows, mask, torch.empty(9, dtype=torch.int32, device="cuda"),
                          0, 100)
    out = torch.empty(10, dtype=torch.int32, device="cuda")
    zc.zero_copy_mask(ptrs, rows, mask, out, 0, 100)
    assert torch.equal(out.cpu(), torch.arange(0, 100, 10, dtype=torch.int32))
    with pytest.raises(RuntimeError, match="pinned"):
        zc.device_alias(torch.zeros(4))
    with pytest.raises(RuntimeError, match="one entry per segment"):
        zc.zero_copy_rearrange(ptrs, rows[:1], torch.zeros(1, dtype=torch.int64, device="cuda"),
                               0, 100, torch.int32, 1)
    with pytest.raises(RuntimeError, match="not a range"):
        zc.zero_copy_rearrange(ptrs, rows, torch.zeros(1, dtype=torch.int64, device="cuda"),
                               50, 10, torch.int32, 1)
    empty = zc.zero_copy_rearrange(ptrs, rows, torch.zeros(0, dtype=torch.int64, device="cuda"),
                                   0, 100, torch.int32, 1)
    assert empty.shape == (0, 1)
