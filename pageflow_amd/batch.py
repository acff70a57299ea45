"""Batched equi-joins (``pf_batch_*``, include/pageflow_join.h): many
independent inner joins of uneven size run as one pipeline of persistent
kernels fed by device work queues, the steady state captured as a hipGraph,
with an autotuner that picks the tunables per size class of the build.

The workload generator ``make_batch`` is the batching benchmark's: for each
of 256 joins, build rows log-uniform in 1K-4M with unique int32 keys (a
seeded permutation), probe rows 8x the build, keys uniform over the build
keys with 15% replaced by absent keys.
"""
from __future__ import annotations

import ctypes
import json
import math
import os
import time
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib

CLASSES = ("small", "medium", "large")        # build rows < 32K, < 512K, larger
# per class: (table slots per build row, probe rows an item (a power of
# two), workgroup size, probe slice KiB).  A slice size > 0 partitions the
# join's probe rows by the top bits of their table slot (radix bits =
# log2(table / slice), at most 8, one pass), and so are its build rows; a
# work item then builds the table of 2^14 slots (128 KiB) of its slice in
# LDS from the slice's build rows and answers its probe rows' lookups there
# (a larger slice is taken in several such sub-ranges, each re-reading the
# slice's probe keys).  0 probes a whole table in HBM at random.
DEFAULT_TUNE = [(2, 4096, 256, 0), (2, 16384, 256, 0), (2, 16384, 512, 128)]
CANDIDATES = [(2, 4096, 256, 0), (4, 4096, 256, 0), (2, 16384, 256, 0), (2, 16384, 512, 0),
              (4, 16384, 512, 0), (2, 65536, 512, 0),
              (2, 16384, 512, 128), (2, 16384, 512, 256), (4, 16384, 512, 128),
              (2, 65536, 512, 128)]


def class_of(nb: int) -> int:
    return 0 if nb < 32768 else (1 if nb < 524288 else 2)


def make_batch(n_joins: int = 256, seed: int = 4, lo: float = 1e3, hi: float = 4e6,
               ratio: int = 8, absent: float = 0.15, device="cuda"):
    """(build sizes, probe sizes, concatenated build keys, concatenated probe
    keys, expected matches per join) of the batching workload, on device."""
    rng = np.random.default_rng(seed)
    nb = np.rint(np.exp(rng.uniform(math.log(lo), math.log(hi), n_joins))).astype(np.int64)
    np_ = nb * ratio
    g = torch.Generator(device=device).manual_seed(seed)
    bkeys, pkeys, expect = [], [], []
    for j in range(n_joins):
        b = (torch.randperm(int(nb[j]), generator=g, device=device) * 2 + 1).to(torch.int32)
        idx = torch.randint(0, int(nb[j]), (int(np_[j]),), generator=g, device=device)
        p = b[idx]
        miss = torch.rand(int(np_[j]), generator=g, device=device) < absent
        p[miss] = (torch.randint(0, 1 << 29, (int(miss.sum()),), generator=g, device=device)
                   * 2).to(torch.int32)  # even keys: never a build key
        bkeys.append(b)
        pkeys.append(p)
        expect.append(int(np_[j]) - int(miss.sum()))
    return nb, np_, torch.cat(bkeys), torch.cat(pkeys), np.asarray(expect)


class BatchJoin:
    """A plan for a batch of joins of fixed sizes (the library owns its
    device scratch)."""

    def __init__(self, nb: Sequence[int], np_: Sequence[int],
                 tune: Optional[Sequence[Tuple[int, int, int, int]]] = None, device=None):
        self.nb = np.asarray(nb, dtype=np.int64)
        self.np = np.asarray(np_, dtype=np.int64)
        self.n = len(self.nb)
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        self.tune = [tuple(t) for t in (tune or DEFAULT_TUNE)]
        if len(self.tune) != len(CLASSES) or any(len(t) != 4 for t in self.tune):
            raise ValueError("BatchJoin: tune is (load, chunk, workgroup, slice KiB) per class")
        t = np.asarray(self.tune, dtype=np.int32).reshape(-1)
        lib = _lib.load()
        with torch.cuda.device(self.device):
            self.plan = lib.pf_batch_plan_create(
                self.n, self.nb.ctypes.data_as(ctypes.c_void_p),
                self.np.ctypes.data_as(ctypes.c_void_p), t.ctypes.data_as(ctypes.c_void_p))
        if not self.plan:
            raise RuntimeError(lib.pf_last_error().decode())
        self.p_off = np.concatenate(([0], np.cumsum(self.np)))
        self.out_probe = torch.empty(max(int(self.p_off[-1]), 1), dtype=torch.int64,
                                     device=self.device)
        self.out_build = torch.empty_like(self.out_probe)
        self.counts = torch.zeros(self.n, dtype=torch.int64, device=self.device)
        self.stream: Optional[torch.cuda.Stream] = None
        self._captured = None

    def close(self) -> None:
        if getattr(self, "plan", None):
            _lib.load().pf_batch_plan_destroy(self.plan)
            self.plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> Dict[str, int]:
        lib = _lib.load()
        names = ["build_rows", "probe_rows", "table_slots", "build_items", "probe_items",
                 "scratch_bytes", "sliced_joins", "lds_overflow"]
        return {n: int(lib.pf_batch_plan_info(self.plan, i)) for i, n in enumerate(names)}

    def run(self, bkeys: torch.Tensor, pkeys: torch.Tensor, per_join: bool = False,
            stream: Optional[torch.cuda.Stream] = None) -> None:
        """The batch, batched (one pipeline, no host wait) or one join at a
        time (the plain baseline, one host wait a join)."""
        st = stream if stream is not None else torch.cuda.current_stream(self.device)
        with torch.cuda.device(self.device):
            _lib.check(_lib.load().pf_batch_join(
                self.plan, _lib.ptr(bkeys), _lib.ptr(pkeys), _lib.ptr(self.out_probe),
                _lib.ptr(self.out_build), _lib.ptr(self.counts), int(per_join),
                ctypes.c_void_p(st.cuda_stream)))

    def capture(self, bkeys: torch.Tensor, pkeys: torch.Tensor) -> torch.cuda.Stream:
        """Records the batched pipeline as a hipGraph on a side stream
        (inputs and outputs bound); replay() launches it."""
        self.stream = torch.cuda.Stream(self.device)
        self._captured = (bkeys, pkeys)
        torch.cuda.current_stream(self.device).synchronize()
        with torch.cuda.device(self.device):
            _lib.check(_lib.load().pf_batch_capture(
                self.plan, _lib.ptr(bkeys), _lib.ptr(pkeys), _lib.ptr(self.out_probe),
                _lib.ptr(self.out_build), _lib.ptr(self.counts),
                ctypes.c_void_p(self.stream.cuda_stream)))
        return self.stream

    def replay(self) -> None:
        if self.stream is None:
            raise RuntimeError("BatchJoin: nothing captured")
        with torch.cuda.device(self.device):
            _lib.check(_lib.load().pf_batch_replay(self.plan,
                                                   ctypes.c_void_p(self.stream.cuda_stream)))

    def pairs(self, j: int) -> Tuple[torch.Tensor, torch.Tensor]:
        """Join j's (probe rows, build rows) after a run (reads its count):
        in probe-row order, or, for a join probed by slices, in probe-row
        order within each slice (slices ascending)."""
        n = int(self.counts[j])
        a = int(self.p_off[j])
        return self.out_probe[a:a + n], self.out_build[a:a + n]


def _time(fn, reps: int, device) -> float:
    fn()
    torch.cuda.synchronize(device)
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize(device)
    return (time.perf_counter() - t0) / reps


def autotune(nb, np_, bkeys, pkeys, reps: int = 3, cache: Optional[str] = None,
             device=None) -> List[Tuple[int, int, int, int]]:
    """Per size class, the candidate tunables with the shortest batched run
    over this batch (coordinate search: one class at a time, the others
    fixed at their best so far).  Results are cached in ``cache`` (JSON) by
    GPU name and batch shape."""
    device = torch.device(device) if device is not None else torch.device("cuda")
    key = f"{torch.cuda.get_device_name(device)}|{len(nb)}|{int(np.sum(nb))}|{int(np.sum(np_))}"
    if cache and os.path.exists(cache):
        with open(cache) as f:
            saved = json.load(f)
        if key in saved:
            return [tuple(t) for t in saved[key]["tune"]]
    present = {class_of(int(b)) for b in nb}
    best = list(DEFAULT_TUNE)
    trials = {}
    for c in range(len(CLASSES)):
        if c not in present:
            continue
        scores = {}
        for cand in CANDIDATES:
            tune = list(best)
            tune[c] = cand
            bj = BatchJoin(nb, np_, tune, device)
            scores[cand] = _time(lambda: bj.run(bkeys, pkeys), reps, device)
            bj.close()
        best[c] = min(scores, key=scores.get)
        trials[CLASSES[c]] = {str(k): v for k, v in scores.items()}
    if cache:
        saved = {}
        if os.path.exists(cache):
            with open(cache) as f:
                saved = json.load(f)
        saved[key] = {"tune": best, "trials_s": trials}
        with open(cache, "w") as f:
            json.dump(saved, f, indent=1)
    return best
