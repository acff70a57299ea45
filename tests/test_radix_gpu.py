"""The stable radix partition (pf_radix_partition) against the oracle: every
mode, int64 and int32 keys, 0 to 12 bits, explicit and implicit row ids,
sizes around the tile and chunk, full skew."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from pageflow_amd import radix

pytestmark = pytest.mark.gpu


def _keys(n, dtype, seed, skew=False):
    g = torch.Generator().manual_seed(seed)
    if skew:
        return torch.full((n,), 12345, dtype=dtype)
    hi = (1 << 40) if dtype == torch.int64 else (1 << 30)
    return torch.randint(-hi, hi, (n,), dtype=dtype, generator=g)


def _check(keys, bits, mode, shift=0, with_vals=True, shape=-1):
    n = keys.numel()
    vals = torch.randperm(n, generator=torch.Generator().manual_seed(n)).to(torch.int32) \
        if with_vals else None
    ko, vo, offs = radix.radix_partition(keys.cuda(), bits,
                                         vals.cuda() if vals is not None else None,
                                         shift=shift, mode=mode, shape=shape)
    v_np = vals.numpy() if vals is not None else np.arange(n, dtype=np.int32)
    wk, wv, counts = oracle.radix_partition(keys.numpy(), v_np, bits, shift,
                                            hashed=mode != radix.RAW)
    if mode == radix.HASH_OUT:
        wk = oracle.mix64(wk.astype(np.int64)).view(np.int64)
    assert np.array_equal(ko.cpu().numpy(), wk)
    assert np.array_equal(vo.cpu().numpy(), wv)
    want_offs = np.concatenate(([0], np.cumsum(counts)))
    assert np.array_equal(offs.cpu().numpy(), want_offs)


@pytest.mark.parametrize("bits", [0, 1, 4, 8, 9, 10, 12])
@pytest.mark.parametrize("dtype", [torch.int64, torch.int32])
def test_hashed_partition(bits, dtype):
    _check(_keys(300_001, dtype, bits), bits, radix.HASH_KEY)


@pytest.mark.parametrize("bits", [6, 12])
def test_hash_out_and_implicit_rows(bits):
    _check(_keys(200_000, torch.int64, 3), bits, radix.HASH_OUT, with_vals=False)
    _check(_keys(200_000, torch.int32, 4), bits, radix.HASH_OUT)


@pytest.mark.parametrize("n", [0, 1, 63, 2047, 2048, 2049, 32768, 32769, 1_000_003])
def test_sizes(n):
    _check(_keys(n, torch.int64, n), 8, radix.HASH_KEY)


@pytest.mark.parametrize("shift,bits", [(0, 8), (8, 8), (20, 12), (24, 8)])
def test_raw_digits(shift, bits):
    g = torch.Generator().manual_seed(shift)
    keys = torch.randint(0, 1 << 31, (250_000,), dtype=torch.int64, generator=g).to(torch.int32)
    _check(keys, bits, radix.RAW, shift=shift)


def test_second_level_digit_below_the_first():
    _check(_keys(400_000, torch.int64, 7), 6, radix.HASH_KEY, shift=6)


@pytest.mark.parametrize("bits", [8, 12])
def test_one_bucket_holds_every_row(bits):
    _check(_keys(500_000, torch.int64, 1, skew=True), bits, radix.HASH_KEY)


def test_large_partition_matches_oracle():
    _check(_keys(20_000_000, torch.int64, 11), 10, radix.HASH_KEY)


@pytest.mark.parametrize("shape", [0, 1, 2, 3, 5])
@pytest.mark.parametrize("bits", [3, 8])
def test_every_scatter_shape(shape, bits):
    keys = _keys(700_001, torch.int64, 20 + shape)
    vals = torch.randperm(keys.numel(), generator=torch.Generator().manual_seed(5)).to(torch.int32)
    ko, vo, offs = radix.radix_partition(keys.cuda(), bits, vals.cuda(), mode=radix.HASH_OUT,
                                         shape=shape)
    wk, wv, counts = oracle.radix_partition(keys.numpy(), vals.numpy(), bits, 0, hashed=True)
    assert np.array_equal(ko.cpu().numpy(), oracle.mix64(wk).view(np.int64))
    assert np.array_equal(vo.cpu().numpy(), wv)
    assert np.array_equal(offs.cpu().numpy(), np.concatenate(([0], np.cumsum(counts))))


@pytest.mark.parametrize("shape", [-1, 4, 1])
@pytest.mark.parametrize("mode,shift,bits", [(0, 0, 9), (0, 3, 11), (1, 0, 12), (1, 5, 10),
                                             (2, 0, 9), (2, 7, 12)])
@pytest.mark.parametrize("dtype", [torch.int64, torch.int32])
def test_wide_digits_two_passes_and_one(shape, mode, shift, bits, dtype):
    # above 8 bits: two stable passes (default or a tile shape) or one (shape 4)
    _check(_keys(400_003, dtype, bits + shift), bits, mode, shift=shift, shape=shape)


def test_wide_digit_skew_and_tiny():
    _check(_keys(100_000, torch.int64, 1, skew=True), 12, radix.HASH_OUT)
    for n in (0, 1, 5):
        _check(_keys(n, torch.int32, 2), 10, radix.HASH_KEY)
