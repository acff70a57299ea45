"""Serving: a continuous-batching request loop for join + group-by pipelines.

Each request joins a fact slice (int32 foreign key, int64 amount; resident
in HBM) to one of a set of recurring dimension tables (int32 key -> int32
group) and returns the exact int64 SUM(amount) and COUNT(amount) per group.
This is the reference's hash join followed by KeyedAggregate pass 1 +
SUM/COUNT (operators/hashjoin.py:444-494, operators/aggregate.py:1511-1563,
operators/operator_utils.py:342-399) collapsed into one pass over the fact
rows: a probe of a resident dimension hash table feeding LDS group
accumulators (pf_serve_step, 12 bytes a fact row).

The engine:

* keeps built dimension tables resident on the device and reuses them; a
  table is built on first use, on its own stream, overlapping the steps in
  flight, and tables no admitted request uses are evicted least recently
  used first when the HBM budget needs the room;
* admits arrived requests first-come first-served while an active slot is
  free and the budget holds the request's dimension table (if not resident)
  and its accumulators;
* runs *steps*: one launch over up to PF_SERVE_MAX_ITEMS (request, row
  range) items, a row quantum to each active request in turn, so requests
  that arrive while others are on the GPU join the next step and work of
  many requests shares every launch; up to ``depth`` steps are queued ahead
  of the device;
* copies a request's 2 x G accumulators to pinned host memory after the
  step that adds its last rows, and reports it done when the host sees that
  step's event.  Latency = done - arrival.

Integer atomics make every result independent of how a request's rows are
split over steps and workgroups and of what else shares them: a request's
result is bit-identical alone and in any mix.
"""
from __future__ import annotations

import ctypes
import math
import time
from collections import OrderedDict, deque
from dataclasses import dataclass, field
from typing import Deque, Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib
from .ops import DimTable

MAX_ITEMS = 40  # PF_SERVE_MAX_ITEMS


class ServeItem(ctypes.Structure):
    """pf_serve_item (include/pageflow_join.h)."""
    _fields_ = [("fk", ctypes.c_void_p), ("amt", ctypes.c_void_p), ("n", ctypes.c_int64),
                ("table", ctypes.c_void_p), ("slots", ctypes.c_int64),
                ("info", ctypes.c_void_p), ("sum", ctypes.c_void_p), ("cnt", ctypes.c_void_p),
                ("G", ctypes.c_int32), ("reserved", ctypes.c_int32)]


@dataclass
class Request:
    rid: int
    dim: int
    fk: torch.Tensor       # int32 foreign keys (device)
    amt: torch.Tensor      # int64 amounts (device)
    arrival: float = 0.0   # seconds after the start of the run


@dataclass
class Result:
    rid: int
    sum: np.ndarray
    cnt: np.ndarray
    arrival: float
    admitted: float
    done: float

    @property
    def latency(self) -> float:
        return self.done - self.arrival


@dataclass
class _Active:
    req: Request
    table: DimTable
    acc: torch.Tensor        # int64 [2, G] on the device
    host: torch.Tensor       # int64 [2, G] pinned
    admitted: float
    pos: int = 0
    nbytes: int = 0
    sent: bool = False       # every row is in a launched step
    item: Optional[ServeItem] = None  # the request's step item, all but its row range

    @property
    def remaining(self) -> int:
        return self.req.fk.numel() - self.pos


@dataclass
class _Dim:
    keys: torch.Tensor
    groups: torch.Tensor
    G: int
    table: Optional[DimTable] = None
    ready: Optional[torch.cuda.Event] = None
    users: int = 0

    @property
    def table_bytes(self) -> int:
        return int(_lib.load().pf_dim_table_slots(self.keys.numel())) * 8 + 16


class ServingEngine:
    """A continuous-batching engine on one device (see the module text)."""

    def __init__(self, device=None, hbm_budget: int = 4 << 30, max_active: int = 32,
                 step_rows: int = 1 << 25, rows_per_wg: int = 1 << 16, depth: int = 2):
        if not 1 <= max_active <= MAX_ITEMS:
            raise ValueError(f"max_active must be in [1, {MAX_ITEMS}]")
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        self.budget = int(hbm_budget)
        self.max_active = max_active
        self.step_rows = int(step_rows)
        self.rows_per_wg = int(rows_per_wg)
        self.depth = depth
        self.dims: Dict[int, _Dim] = {}
        self.lru: "OrderedDict[int, None]" = OrderedDict()  # resident dims, oldest first
        self.used = 0          # bytes held: resident tables + admitted requests' accumulators
        self.peak_used = 0
        self.stream = torch.cuda.Stream(self.device)
        self.build_stream = torch.cuda.Stream(self.device)
        self.stats = {"steps": 0, "builds": 0, "evictions": 0, "max_items": 0}
        # one item array for every step: pf_serve_step copies the items into
        # its launch's arguments, so the array is free again when it returns
        self._items = (ServeItem * MAX_ITEMS)()
        self._items_arg = ctypes.cast(self._items, ctypes.c_void_p)
        self._stream_arg = ctypes.c_void_p(self.stream.cuda_stream)

    # ------------------------------------------------------------ dims --
    def register_dim(self, dim_id: int, keys: torch.Tensor, groups: torch.Tensor,
                     G: int) -> None:
        d = _Dim(keys.contiguous(), groups.contiguous(), int(G))
        if d.table_bytes + 16 * d.G > self.budget:
            raise ValueError(f"dimension {dim_id}: its table alone exceeds the HBM budget")
        self.dims[dim_id] = d

    def _make_room(self, need: int) -> bool:
        while self.used + need > self.budget:
            victim = next((i for i in self.lru if self.dims[i].users == 0), None)
            if victim is None:
                return False
            d = self.dims[victim]
            self.lru.pop(victim)
            d.table = None
            d.ready = None
            self.used -= d.table_bytes
            self.stats["evictions"] += 1
        return True

    def _ensure_table(self, dim_id: int) -> DimTable:
        d = self.dims[dim_id]
        if d.table is None:
            # the build overlaps the steps in flight; the step that first
            # reads the table waits for it on the device
            d.table = DimTable(d.keys, d.groups, d.G, stream=self.build_stream)
            d.ready = torch.cuda.Event()
            d.ready.record(self.build_stream)
            self.used += d.table_bytes
            self.stats["builds"] += 1
        self.lru.pop(dim_id, None)
        self.lru[dim_id] = None
        return d.table

    def check_tables(self) -> None:
        """Raises if a resident table saw a repeated key or a group out of
        range (waits for the builds)."""
        for d in self.dims.values():
            if d.table is not None:
                d.table.check()

    # ------------------------------------------------------- admission --
    def _admit(self, req: Request, now: float) -> Optional[_Active]:
        d = self.dims[req.dim]
        acc_bytes = 16 * d.G
        need = acc_bytes + (0 if d.table is not None else d.table_bytes)
        d.users += 1  # pins a resident table while room is made
        if not self._make_room(need):
            d.users -= 1
            return None
        table = self._ensure_table(req.dim)
        self.used += acc_bytes
        self.peak_used = max(self.peak_used, self.used)
        acc = torch.empty((2, d.G), dtype=torch.int64, device=self.device)
        with torch.cuda.stream(self.stream):
            acc.zero_()
        host = torch.empty((2, d.G), dtype=torch.int64, pin_memory=True)
        item = ServeItem(req.fk.data_ptr(), req.amt.data_ptr(), 0, table.table.data_ptr(),
                         table.slots, table.info.data_ptr(), acc[0].data_ptr(),
                         acc[1].data_ptr(), d.G, 0)
        return _Active(req, table, acc, host, now, nbytes=acc_bytes, item=item)

    def _release(self, st: _Active) -> None:
        self.used -= st.nbytes
        self.dims[st.req.dim].users -= 1

    # ------------------------------------------------------------ steps --
    def _launch_step(self, active: List[_Active], start: int) -> Tuple[List[_Active], int]:
        """One step over the active requests from index `start` on; returns
        the requests whose last rows it adds and the next start."""
        live = [a for a in active if not a.sent]
        if not live:
            return [], start
        n = len(live)
        order = [live[(start + i) % n] for i in range(min(n, MAX_ITEMS))]
        quantum = max(self.step_rows // len(order), self.rows_per_wg)
        items = self._items
        waits = set()
        finished = []
        for i, a in enumerate(order):
            take = min(a.remaining, quantum)
            if self.dims[a.req.dim].ready is not None:
                waits.add(a.req.dim)
            tpl = a.item
            items[i] = tpl
            it = items[i]
            it.fk = (tpl.fk or 0) + 4 * a.pos      # (an empty tensor's pointer is null)
            it.amt = (tpl.amt or 0) + 8 * a.pos
            it.n = take
            a.pos += take
            if a.remaining == 0:  # (an empty request rides along as a 0-row item)
                a.sent = True
                finished.append(a)
        for dim_id in waits:
            d = self.dims[dim_id]
            self.stream.wait_event(d.ready)
            d.ready = None  # later steps are ordered after this one
        with torch.cuda.device(self.device):
            _lib.check(_lib.load().pf_serve_step(self._items_arg, len(order), self.rows_per_wg,
                                                 self._stream_arg))
            with torch.cuda.stream(self.stream):
                for a in finished:
                    a.host.copy_(a.acc, non_blocking=True)
        self.stats["steps"] += 1
        self.stats["max_items"] = max(self.stats["max_items"], len(order))
        return finished, (start + len(order)) % max(n, 1)

    # -------------------------------------------------------------- run --
    def run(self, requests: Sequence[Request], realtime: bool = True) -> List[Result]:
        """Serves `requests` (arrival times in seconds from the start; with
        realtime=False every request has arrived at the start) and returns
        their results in request order."""
        for r in requests:
            if r.dim not in self.dims:
                raise KeyError(f"request {r.rid}: unknown dimension {r.dim}")
            if r.fk.dtype != torch.int32 or r.amt.dtype != torch.int64 or \
                    r.fk.numel() != r.amt.numel() or not r.fk.is_contiguous() or \
                    not r.amt.is_contiguous():
                raise ValueError(f"request {r.rid}: contiguous int32 keys and int64 amounts "
                                 "of one length")
        arrivals = sorted(requests, key=lambda r: (r.arrival if realtime else 0.0, r.rid))
        pending: Deque[Request] = deque()
        active: List[_Active] = []
        inflight: Deque[Tuple[torch.cuda.Event, List[_Active]]] = deque()
        results: Dict[int, Result] = {}
        nxt = 0
        start = 0
        torch.cuda.synchronize(self.device)
        t0 = time.perf_counter()
        while len(results) < len(requests):
            now = time.perf_counter() - t0
            while nxt < len(arrivals) and (not realtime or arrivals[nxt].arrival <= now):
                pending.append(arrivals[nxt])
                nxt += 1
            while pending and len(active) < self.max_active:
                st = self._admit(pending[0], now)
                if st is None:
                    break
                pending.popleft()
                active.append(st)
            launched = False
            if len(inflight) < self.depth and any(not a.sent for a in active):
                finished, start = self._launch_step(active, start)
                ev = torch.cuda.Event()
                ev.record(self.stream)
                inflight.append((ev, finished))
                launched = True
            while inflight and inflight[0][0].query():
                _, fin = inflight.popleft()
                done = time.perf_counter() - t0
                for a in fin:
                    r = a.req
                    results[r.rid] = Result(r.rid, a.host[0].numpy().copy(),
                                            a.host[1].numpy().copy(),
                                            r.arrival if realtime else 0.0, a.admitted, done)
                    active.remove(a)
                    self._release(a)
            if not launched and inflight:
                if len(inflight) >= self.depth or all(a.sent for a in active):
                    time.sleep(0)  # the device is busy: poll again
            elif not launched and not inflight and nxt < len(arrivals) and not pending:
                wait = arrivals[nxt].arrival - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(min(wait, 0.001))
        return [results[r.rid] for r in requests]

    def run_alone(self, req: Request) -> Result:
        """The request served by itself (nothing else admitted)."""
        r = Request(req.rid, req.dim, req.fk, req.amt, 0.0)
        return self.run([r], realtime=False)[0]


# ------------------------------------------------------------- workload --

@dataclass
class Workload:
    dims: Dict[int, Tuple[torch.Tensor, torch.Tensor]]
    requests: List[Request]
    G: int = 25
    dim_rows: List[int] = field(default_factory=list)


def make_workload(n_requests: int, rate: float = 0.0, seed: int = 5, n_dims: int = 8,
                  fact_lo: float = 1e5, fact_hi: float = 5e7, dim_lo: float = 1e3,
                  dim_hi: float = 2e6, slices_per_dim: int = 8, G: int = 25,
                  device=None) -> Workload:
    """The serving workload: n_dims recurring dimensions (rows log-uniform in
    [dim_lo, dim_hi], keys a seeded permutation of 1..n, groups uniform in
    [0, G)); requests round-robin over the dimensions, each with a fact
    slice (rows log-uniform in [fact_lo, fact_hi], foreign keys uniform
    over the dimension's keys, amounts uniform in [100, 10^7]).  The fact
    slices come from a pool of slices_per_dim slices a dimension (request r
    uses slice r // n_dims of its dimension's pool, cyclically), so long
    runs need no more memory.  Poisson arrivals at `rate` requests/s
    (rate 0: every arrival at 0)."""
    dev = torch.device(device) if device is not None else torch.device("cuda")
    rng = np.random.default_rng(seed)
    g = torch.Generator(device=dev).manual_seed(seed)
    dim_rows = np.rint(np.exp(rng.uniform(math.log(dim_lo), math.log(dim_hi), n_dims))
                       ).astype(np.int64)
    dims = {}
    for d in range(n_dims):
        n = int(dim_rows[d])
        keys = (torch.randperm(n, generator=g, device=dev) + 1).to(torch.int32)
        grp = torch.randint(0, G, (n,), generator=g, device=dev, dtype=torch.int32)
        dims[d] = (keys, grp)
    pool = {}
    for d in range(n_dims):
        for s in range(slices_per_dim):
            nf = int(round(math.exp(rng.uniform(math.log(fact_lo), math.log(fact_hi)))))
            fk = torch.randint(1, int(dim_rows[d]) + 1, (nf,), generator=g, device=dev,
                               dtype=torch.int32)
            amt = torch.randint(100, 10_000_001, (nf,), generator=g, device=dev,
                                dtype=torch.int64)
            pool[(d, s)] = (fk, amt)
    gaps = rng.exponential(1.0 / rate, n_requests) if rate > 0 else np.zeros(n_requests)
    arrival = np.cumsum(gaps)
    reqs = []
    for r in range(n_requests):
        d = r % n_dims
        fk, amt = pool[(d, (r // n_dims) % slices_per_dim)]
        reqs.append(Request(r, d, fk, amt, float(arrival[r])))
    return Workload(dims, reqs, G, [int(x) for x in dim_rows])


def with_arrivals(reqs: Sequence[Request], rate: float, seed: int = 5) -> List[Request]:
    """The same requests with Poisson arrivals at `rate` requests/s."""
    rng = np.random.default_rng(seed + 1)
    arrival = np.cumsum(rng.exponential(1.0 / rate, len(reqs)))
    return [Request(r.rid, r.dim, r.fk, r.amt, float(a)) for r, a in zip(reqs, arrival)]


def latency_summary(results: Sequence[Result]) -> Dict[str, float]:
    lat = np.sort(np.array([r.latency for r in results])) * 1e3
    return {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "max_ms": float(lat.max())}


__all__ = ["ServingEngine", "Request", "Result", "ServeItem", "make_workload", "with_arrivals",
           "latency_summary", "MAX_ITEMS"]
