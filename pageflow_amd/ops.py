"""Small device operations of the join's output side over the library:
payload gather (with the null row), matched-row marking and selection, and
the int64 group-by.  See include/pageflow_join.h for the kernels."""
from __future__ import annotations

import ctypes
from typing import Optional, Tuple

import torch

from . import _lib
from .layout import NULL

_host: Optional[torch.Tensor] = None


def _host_buffer() -> torch.Tensor:
    global _host
    if _host is None:
        _host = torch.empty(2, dtype=torch.int64, pin_memory=True)
    return _host


def _null_bits(dtype: torch.dtype) -> int:
    """The bits of the null value -64 in ``dtype`` (conversion.py:39-50
    fills a null row with -64 in the column's own type), as an int64."""
    return int(torch.tensor([NULL], dtype=dtype).view(
        {1: torch.int8, 2: torch.int16, 4: torch.int32, 8: torch.int64}[
            torch.tensor([], dtype=dtype).element_size()])[0])


def gather_rows(src: torch.Tensor, idx: torch.Tensor, check: bool = False) -> torch.Tensor:
    """``src``'s rows at ``idx`` (int64, 1-D), a null row (every element
    -64 in ``src``'s type) where ``idx`` is -64 or ``src`` has no row:
    index_select and index_with_null (conversion.py:39-50) in one kernel.
    Rows are copied as bits, so any 1/2/4/8-byte element type works.  With
    ``check``, an index outside the rows (other than -64) raises
    RuntimeError after the gather (one wait)."""
    _lib.require_cuda(src, "gather source")
    _lib.require_cuda(idx, "gather index")
    if src.element_size() not in (1, 2, 4, 8) or src.dim() not in (1, 2):
        raise ValueError("gather_rows: 1-D or 2-D rows of 1/2/4/8-byte elements")
    if src.dtype == torch.bool:
        raise ValueError("gather_rows: a bool column has no null value")
    src = src.contiguous()
    idx = idx.contiguous().to(torch.int64)
    row_elems = src.shape[1] if src.dim() == 2 else 1
    out = torch.empty((idx.numel(),) + tuple(src.shape[1:]), dtype=src.dtype, device=src.device)
    lib = _lib.load()
    bad = torch.zeros(1, dtype=torch.int64, device=src.device) if check else None
    with torch.cuda.device(src.device):
        _lib.check(lib.pf_gather_rows(_lib.ptr(src), src.shape[0], src.element_size(), row_elems,
                                      _lib.ptr(idx), idx.numel(), _lib.ptr(out), _lib.ptr(bad),
                                      _null_bits(src.dtype), _lib.stream_of(src.device)))
    if check and int(bad.item()):
        raise RuntimeError(f"gather_rows: an index lies outside the {src.shape[0]} rows")
    return out


# pageflow_join.h PF_KEY_*
_KEY_TYPE = {torch.int64: 0, torch.int32: 1, torch.int16: 2, torch.int8: 3, torch.uint8: 4,
             torch.float64: 5, torch.float32: 6, torch.bool: 7}


def join_key_bits(key: torch.Tensor, side: int, as_float: bool = False) -> torch.Tensor:
    """A key column of any scalar type as int64 join keys
    (``pf_join_key_bits``): integers by value; floats, and integers when
    ``as_float`` (joined against a float side), by float64 bit pattern with
    -0.0 folded into +0.0 and NaN made unmatched (side 0 build, 1 probe)."""
    _lib.require_cuda(key, "a join key")
    if key.dtype not in _KEY_TYPE:
        raise ValueError(f"join key of {key.dtype}: keys are integer, bool or float")
    if key.dim() == 2 and key.shape[1] == 1:
        key = key.reshape(-1)
    if key.dim() != 1:
        raise ValueError(f"a join key of shape {tuple(key.shape)}: keys are 1-D")
    key = key.contiguous()
    out = torch.empty(key.numel(), dtype=torch.int64, device=key.device)
    with torch.cuda.device(key.device):
        _lib.check(_lib.load().pf_join_key_bits(_lib.ptr(key), _KEY_TYPE[key.dtype], key.numel(),
                                                int(bool(as_float)), int(side), _lib.ptr(out),
                                                _lib.stream_of(key.device)))
    return out


def matched_rows(idx: torch.Tensor, n: int, matched: bool) -> torch.Tensor:
    """The rows of [0, n), ascending, that appear in ``idx`` (matched) or
    do not (unmatched): the bool scatter + nonzero tails of
    operators/hashjoin.py:301-348 as two kernels."""
    dev = idx.device
    flags = torch.zeros(max(n, 1), dtype=torch.uint8, device=dev)
    lib = _lib.load()
    st = _lib.stream_of(dev)
    out = torch.empty(max(n, 1), dtype=torch.int64, device=dev)
    host = _host_buffer()
    ws_bytes = lib.pf_select_rows_workspace_bytes(n)
    ws = _lib.workspace(ws_bytes, dev)
    with torch.cuda.device(dev):
        _lib.check(lib.pf_mark_rows(_lib.ptr(idx.contiguous()), idx.numel(), _lib.ptr(flags), n,
                                    1, st))
        _lib.check(lib.pf_select_rows(_lib.ptr(flags), n, int(matched), _lib.ptr(out),
                                      ctypes.c_void_p(host.data_ptr()), _lib.ptr(ws), ws_bytes,
                                      st))
    return out[:int(host[0])]


def group_sum_count(gid: torch.Tensor, val: torch.Tensor, G: int,
                    out: Optional[Tuple[torch.Tensor, torch.Tensor]] = None,
                    accumulate: bool = False) -> Tuple[torch.Tensor, torch.Tensor]:
    """Exact int64 SUM(val) and COUNT(val != -64) per group id in [0, G)."""
    _lib.require_cuda(gid, "group ids")
    if gid.dtype != torch.int32 or val.dtype != torch.int64 or gid.numel() != val.numel():
        raise ValueError("group_sum_count: int32 group ids and int64 values of one length")
    dev = gid.device
    if out is None:
        out = (torch.empty(G, dtype=torch.int64, device=dev),
               torch.empty(G, dtype=torch.int64, device=dev))
    lib = _lib.load()
    with torch.cuda.device(dev):
        _lib.check(lib.pf_group_sum_count(_lib.ptr(gid.contiguous()), _lib.ptr(val.contiguous()),
                                          gid.numel(), G, _lib.ptr(out[0]), _lib.ptr(out[1]),
                                          int(accumulate), None, _lib.stream_of(dev)))
    return out


class DimTable:
    """A resident dimension table: int32 key -> int32 group in [0, G).
    Built once on the device and reused by every request that joins to it:
    a direct byte array when the keys are dense (they span at most 4n
    values and G <= 255), an open-addressing hash table otherwise."""

    def __init__(self, keys: torch.Tensor, groups: torch.Tensor, G: int,
                 stream: Optional[torch.cuda.Stream] = None):
        _lib.require_cuda(keys, "dimension keys")
        if keys.dtype != torch.int32 or groups.dtype != torch.int32 or \
                keys.numel() != groups.numel():
            raise ValueError("DimTable: int32 keys and int32 groups of one length")
        self.G = int(G)
        self.n = keys.numel()
        self.device = keys.device
        lib = _lib.load()
        self.slots = int(lib.pf_dim_table_slots(self.n))
        self.table = torch.empty(self.slots, dtype=torch.int64, device=self.device)
        # [0] groups out of range, [1] repeated keys, [2] smallest key,
        # [3] dense key range (0: hashed)
        self.info = torch.empty(4, dtype=torch.int64, device=self.device)
        st = stream if stream is not None else torch.cuda.current_stream(self.device)
        with torch.cuda.device(self.device):
            _lib.check(lib.pf_dim_table_build(_lib.ptr(keys.contiguous()),
                                              _lib.ptr(groups.contiguous()), self.n, self.G,
                                              _lib.ptr(self.table), self.slots,
                                              _lib.ptr(self.info),
                                              ctypes.c_void_p(st.cuda_stream)))

    @property
    def nbytes(self) -> int:
        return self.slots * 8

    @property
    def dense(self) -> bool:
        """True when the table is a direct byte array (waits)."""
        return int(self.info[3]) > 0

    def check(self) -> None:
        """Raises when a group was out of range or a key repeated (waits)."""
        bad_group, dup = (int(v) for v in self.info[:2].tolist())
        if bad_group or dup:
            raise ValueError(f"DimTable: {bad_group} groups out of [0, {self.G}), "
                             f"{dup} repeated keys")

    def join_groupby(self, fk: torch.Tensor, amt: torch.Tensor,
                     out: Optional[Tuple[torch.Tensor, torch.Tensor]] = None,
                     stream: Optional[torch.cuda.Stream] = None
                     ) -> Tuple[torch.Tensor, torch.Tensor]:
        """SUM(amt), COUNT(amt) per group over the fact rows whose key is in
        the table (inner join)."""
        if fk.dtype != torch.int32 or amt.dtype != torch.int64 or fk.numel() != amt.numel():
            raise ValueError("join_groupby: int32 keys and int64 amounts of one length")
        if out is None:
            out = (torch.empty(self.G, dtype=torch.int64, device=self.device),
                   torch.empty(self.G, dtype=torch.int64, device=self.device))
        st = stream if stream is not None else torch.cuda.current_stream(self.device)
        lib = _lib.load()
        with torch.cuda.device(self.device):
            _lib.check(lib.pf_fact_join_groupby(_lib.ptr(fk), _lib.ptr(amt), fk.numel(),
                                                _lib.ptr(self.table), self.slots,
                                                _lib.ptr(self.info), self.G,
                                                _lib.ptr(out[0]), _lib.ptr(out[1]),
                                                ctypes.c_void_p(st.cuda_stream)))
        return out


__all__ = ["gather_rows", "matched_rows", "group_sum_count", "DimTable", "NULL"]
