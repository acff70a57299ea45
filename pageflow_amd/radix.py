"""Stable radix partition of (key, row id) pairs into up to 2^12 buckets
(``pf_radix_partition``): the partition of the batched and large-join tiers
and the digit pass of an LSD sort.  See include/pageflow_join.h."""
from __future__ import annotations

from typing import Optional, Tuple

import torch

from . import _lib

HASH_KEY, HASH_OUT, RAW = 0, 1, 2
MAX_BITS = 12
# Row ids are int32 (the iota ids written when ``vals`` is None included).
MAX_ROWS = (1 << 31) - 1


def workspace_bytes(n: int, bits: int) -> int:
    return int(_lib.load().pf_radix_partition_workspace_bytes(n, bits))


def radix_partition(keys: torch.Tensor, bits: int, vals: Optional[torch.Tensor] = None,
                    shift: int = 0, mode: int = HASH_KEY,
                    out: Optional[Tuple[torch.Tensor, torch.Tensor]] = None,
                    bucket_offsets: Optional[torch.Tensor] = None,
                    ws: Optional[torch.Tensor] = None, shape: int = -1):
    """(keys or hashes, row ids, bucket offsets[2^bits + 1]) of a stable
    partition of ``keys`` (int64 or int32, 1-D) by the digit ``mode``
    selects; ``vals`` are int32 row ids (None: the row index)."""
    _lib.require_cuda(keys, "keys")
    if keys.dtype not in (torch.int64, torch.int32) or keys.dim() != 1:
        raise ValueError("radix_partition: keys must be 1-D int64 or int32")
    if not 0 <= bits <= MAX_BITS:
        raise ValueError(f"radix_partition: bits must be in [0, {MAX_BITS}]")
    keys = keys.contiguous()
    n = keys.numel()
    if n > MAX_ROWS:
        raise ValueError(f"radix_partition: {n} rows; row ids are int32, so at most {MAX_ROWS}")
    dev = keys.device
    if vals is not None:
        if vals.dtype != torch.int32 or vals.numel() != n:
            raise ValueError("radix_partition: vals must be int32 with one value a key")
        vals = vals.contiguous()
    kdt = torch.int64 if mode == HASH_OUT else keys.dtype
    if out is None:
        out = (torch.empty(n, dtype=kdt, device=dev), torch.empty(n, dtype=torch.int32, device=dev))
    if bucket_offsets is None:
        bucket_offsets = torch.empty((1 << bits) + 1, dtype=torch.int64, device=dev)
    lib = _lib.load()
    nbytes = lib.pf_radix_partition_workspace_bytes(n, bits)
    if ws is None or ws.numel() < nbytes:
        ws = _lib.workspace(nbytes, dev)
    with torch.cuda.device(dev):
        _lib.check(lib.pf_radix_partition_shape(
            _lib.ptr(keys), keys.element_size(), _lib.ptr(vals), n, bits, shift, mode,
            _lib.ptr(out[0]), _lib.ptr(out[1]), _lib.ptr(bucket_offsets), _lib.ptr(ws),
            ws.numel(), shape, _lib.stream_of(dev)))
    return out[0], out[1], bucket_offsets
