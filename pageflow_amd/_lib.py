"""The C ABI library (include/pageflow_join.h) through ctypes.

The library is built in tree by ``pageflow_amd.build`` (or
``__graft_entry__.build()``).  There is no fallback: a missing or stale
library raises, so a call never silently runs anything but the HIP kernels.
"""
from __future__ import annotations

import ctypes
import os
import threading
from typing import Optional

import torch

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpageflow_join.so")
ABI_VERSION = 1

P = ctypes.c_void_p
I64 = ctypes.c_int64
I32 = ctypes.c_int32
SZ = ctypes.c_size_t
PP = ctypes.POINTER(ctypes.c_void_p)

# name -> (restype, argtypes)
_SIGNATURES = {
    "pf_version": (ctypes.c_int, []),
    "pf_last_error": (ctypes.c_char_p, []),
    "pf_profile": (ctypes.c_int, [I32]),
    "pf_profile_read": (I32, [ctypes.c_char_p, I64, P, P, I32]),
    "pf_mix64": (ctypes.c_int, [P, I64, P, P]),
    "pf_bucket_ids": (ctypes.c_int, [P, I64, I32, P, P]),
    "pf_join_key_bits": (ctypes.c_int, [P, I32, I64, I32, I32, P, P]),
    "pf_partition_sort_workspace_bytes": (SZ, [I64, I64, I32, I32, I32, I32]),
    "pf_partition_sort": (ctypes.c_int, [PP, PP, P, P, I32, I32, P, I64, I64, I64, I32, I32,
                                         P, P, P, SZ, P]),
    "pf_partition_by_bucket_workspace_bytes": (SZ, [I64, I32]),
    "pf_partition_by_bucket": (ctypes.c_int, [PP, PP, P, I32, P, I64, I32, P, P, SZ, P]),
    "pf_join_unique_hash_workspace_bytes": (SZ, [I64, I64, I32]),
    "pf_join_unique_hash": (ctypes.c_int, [P, P, I64, P, P, I64, I32, P, P, P, P, SZ, P]),
    "pf_join_unique_hash_noalloc": (ctypes.c_int, [P, P, I64, P, P, I64, I32, P, P, P, P]),
    "pf_join_semi_anti_workspace_bytes": (SZ, [I64, I64, I32]),
    "pf_join_semi_anti": (ctypes.c_int, [P, I64, P, I64, P, P, I32, I32, P, P, P, SZ, P]),
    "pf_semi_anti_mask_workspace_bytes": (SZ, [I64, I64, I32]),
    "pf_semi_anti_mask": (ctypes.c_int, [P, I64, P, I64, P, P, I32, I32, P, P, P, SZ, P]),
    "pf_join_expand_workspace_bytes": (SZ, [I64]),
    "pf_join_expand_count": (ctypes.c_int, [P, I64, P, P, SZ, P]),
    "pf_join_expand_write": (ctypes.c_int, [P, P, I64, I64, P, P, P, P]),
    "pf_radix_partition_workspace_bytes": (SZ, [I64, I32]),
    "pf_radix_partition": (ctypes.c_int, [P, I32, P, I64, I32, I32, I32, P, P, P, P, SZ, P]),
    "pf_radix_partition_shape": (ctypes.c_int, [P, I32, P, I64, I32, I32, I32, P, P, P, P, SZ,
                                                I32, P]),
    "pf_gather_rows": (ctypes.c_int, [P, I64, I32, I64, P, I64, P, P, I64, P]),
    "pf_mark_rows": (ctypes.c_int, [P, I64, P, I64, I32, P]),
    "pf_select_rows_workspace_bytes": (SZ, [I64]),
    "pf_select_rows": (ctypes.c_int, [P, I64, I32, P, P, P, SZ, P]),
    "pf_pair_filter_workspace_bytes": (SZ, [I64]),
    "pf_pair_filter_count": (ctypes.c_int, [P, P, I64, P, P, P, P, I32, P, I32, P, P, P, SZ,
                                            P]),
    "pf_pair_filter_write": (ctypes.c_int, [P, P, I64, P, P, P, P]),
    "pf_group_sum_count": (ctypes.c_int, [P, P, I64, I32, P, P, I32, P, P]),
    "pf_group_by_keys_workspace_bytes": (SZ, [I64, I32, I32]),
    "pf_group_by_keys": (ctypes.c_int, [P, P, I32, I64, P, P, I32, P, P, P, P, P, P, P, P, SZ,
                                        P]),
    "pf_dim_table_slots": (I64, [I64]),
    "pf_dim_table_build": (ctypes.c_int, [P, P, I64, I32, P, I64, P, P]),
    "pf_fact_join_groupby": (ctypes.c_int, [P, P, I64, P, I64, P, I32, P, P, P]),
    "pf_serve_step": (ctypes.c_int, [P, I32, I32, P]),
    "pf_device_alias": (ctypes.c_int, [P, ctypes.POINTER(P)]),
    "pf_zero_copy_rearrange_workspace_bytes": (SZ, [I64]),
    "pf_zero_copy_rearrange": (ctypes.c_int, [P, P, I32, I64, I64, P, I64, I32, I32, I64, P, P,
                                              P, SZ, P]),
    "pf_zero_copy_mask_workspace_bytes": (SZ, [I64]),
    "pf_zero_copy_mask": (ctypes.c_int, [P, P, I32, I64, I64, P, I32, I32, I64, P, I64, P, P,
                                         P, SZ, P]),
    "pf_batch_plan_create": (P, [I32, P, P, P]),
    "pf_batch_plan_destroy": (ctypes.c_int, [P]),
    "pf_batch_plan_info": (I64, [P, I32]),
    "pf_batch_join": (ctypes.c_int, [P, P, P, P, P, P, I32, P]),
    "pf_batch_capture": (ctypes.c_int, [P, P, P, P, P, P, P]),
    "pf_batch_replay": (ctypes.c_int, [P, P]),
    "pf_hash_join_workspace_bytes": (SZ, [I64, I64, I32, I32]),
    "pf_hash_join_prepare": (ctypes.c_int, [P, I64, P, I64, I32, I32, P, P, SZ, P]),
    "pf_hash_join_write": (ctypes.c_int, [I64, I64, I32, I32, I64, PP, I32, P, P, PP, P, P,
                                          SZ, P]),
    "pf_hash_join_residual_workspace_bytes": (SZ, [I64, I64, I32, I32, I32, I32]),
    "pf_hash_join_residual_prepare": (ctypes.c_int, [P, I64, P, I64, I32, I32, P, P, P, P, I32,
                                                     P, I32, P, P, SZ, P]),
    "pf_hash_join_residual_write": (ctypes.c_int, [I64, I64, I32, I32, I64, PP, I32, P, P, PP,
                                                   P, P, P, P, I32, P, I32, P, SZ, P]),
}

_lib: Optional[ctypes.CDLL] = None
_lock = threading.Lock()


class LibraryMissing(RuntimeError):
    pass


def declared_symbols():
    return list(_SIGNATURES)


def load() -> ctypes.CDLL:
    """The library, loaded once; raises LibraryMissing when it is not
    built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise LibraryMissing(
                f"{LIB_PATH} is not built: run `python -m pageflow_amd.build` "
                "(the join path has no CPU fallback)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        if lib.pf_version() != ABI_VERSION:
            raise LibraryMissing(
                f"{LIB_PATH} has ABI {lib.pf_version()}, expected {ABI_VERSION}: rebuild it")
        _lib = lib
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        msg = load().pf_last_error().decode(errors="replace")
        raise RuntimeError(msg or f"pageflow_join call failed ({rc})")


def ptr(t: Optional[torch.Tensor]):
    """A tensor's device address for the C ABI (None for None)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def stream_of(device: torch.device):
    """The current torch stream of ``device`` as a hipStream_t."""
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def require_cuda(t: torch.Tensor, what: str) -> None:
    if not t.is_cuda:
        raise ValueError(f"{what} must be a GPU tensor (the join path has no CPU fallback)")


def workspace(nbytes: int, device: torch.device) -> torch.Tensor:
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)


def void_array(ptrs):
    arr = (ctypes.c_void_p * max(len(ptrs), 1))()
    for i, p in enumerate(ptrs):
        arr[i] = p
    return arr


def i32_array(vals):
    arr = (ctypes.c_int32 * max(len(vals), 1))()
    for i, v in enumerate(vals):
        arr[i] = int(v)
    return arr


def i64_array(vals):
    arr = (ctypes.c_int64 * max(len(vals), 1))()
    for i, v in enumerate(vals):
        arr[i] = int(v)
    return arr


def profile(enable: bool) -> None:
    """Starts (clearing) or stops the library's per-kernel event timing."""
    check(load().pf_profile(int(bool(enable))))


def profile_read(max_kernels: int = 64):
    """{kernel name: (total ms, launches)} recorded since profile(True)."""
    lib = load()
    names = ctypes.create_string_buffer(max_kernels * 96)
    ms = (ctypes.c_double * max_kernels)()
    cnt = (ctypes.c_int64 * max_kernels)()
    k = lib.pf_profile_read(names, len(names), ctypes.cast(ms, ctypes.c_void_p),
                            ctypes.cast(cnt, ctypes.c_void_p), max_kernels)
    raw = names.raw.split(b"\0")
    return {raw[i].decode(): (ms[i], int(cnt[i])) for i in range(k)}
