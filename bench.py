"""Benchmark of the equi-join hot path on one MI355X.

Workload (the large plain join): build 16M unique int64 keys (a seeded
permutation of 1..16M) with two int64 payload columns; probe 640M int64 keys
uniform over the build keys, 10% replaced by absent keys drawn from
[16M+1, 2^62).  A step is one whole join over inputs resident in HBM: both
sides radix-partitioned by mix64, each bucket joined against an LDS hash
table, the matched (probe row, build row) pairs written with one gathered
build payload column.  tuples/s = (build rows + probe rows) / step time.

Prints one JSON line (see the fields below): the throughput, the roofline of
the dominant kernel (its algorithmic bytes over its average launch time,
measured with HIP events over the timed steps on the launch stream), and the
CPU baseline: the oracle's C restatement of the reference's unique-hash
join program timed on this host on a bounded sample.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def kernel_bytes(name: str, nb: int, np_: int, pairs: int, launches: float) -> float:
    """Algorithmic bytes of ONE launch of a kernel of the large join.

    A complete radix partition of a side counts once what it must move: its
    keys and row ids read once and written once, 24 B a row (the standalone
    partition's figure).  The join partitions each side in two 6-bit passes
    of a scatter kernel, so each scatter launch carries half of that; the
    histogram launches move no algorithmic bytes of their own (their reads
    are partition work already counted); the two passes are timed
    apart (hj_probe_scatter, hj_probe_scatter_2).  The count pass (only in
    the ordered mode) reads the
    partitioned probe and build hashes; the write pass reads the partitioned
    probe (hash, row id), build (hash, row id) and build payload once and
    writes the pairs and the gathered payload."""
    per_step = {
        "hj_probe_scatter": np_ * 12, "hj_probe_scatter_2": np_ * 12,
        "hj_build_scatter": nb * 12, "hj_build_scatter_2": nb * 12,
        "hj_count": np_ * 8 + nb * 8,
        "hj_write": np_ * 12 + nb * 20 + pairs * 24,
    }
    return float(per_step.get(name, 0)) / max(launches, 1.0)


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _host_threads() -> int:
    """The cores this process may run on (its affinity set)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline_c1():
    """The reference's join program on the CPU at the plain join's full size
    (build 1M rows, int32 keys a seeded permutation of 1..1M; probe 6M int32
    foreign keys, each matching once): the torch restatement of its
    sort/search core (oracle/torch_cpu.py), which is the route the reference
    takes for int32 keys, on every core this process may use; median of 3
    runs after a warm-up."""
    import torch
    from oracle import torch_cpu

    nb, np_ = 1_000_000, 6_000_000
    g = torch.Generator().manual_seed(1)
    build = (torch.randperm(nb, generator=g) + 1).to(torch.int32)
    probe = build[torch.randint(0, nb, (np_,), generator=g)]
    threads = _host_threads()
    before = torch.get_num_threads()
    torch.set_num_threads(threads)
    try:
        li, ri = torch_cpu.sort_search_pairs(build, probe)
        assert li.numel() == np_ and bool((build[li] == probe[ri]).all())
        times = []
        for _ in range(3):
            t0 = time.perf_counter()
            torch_cpu.sort_search_pairs(build, probe)
            times.append(time.perf_counter() - t0)
    finally:
        torch.set_num_threads(before)
    t = sorted(times)[1]
    return {"value": (nb + np_) / t, "unit": "tuples/s", "cores": threads,
            "os_cpu_count": os.cpu_count(), "cpu_model": _cpu_model(), "kind": "port",
            "sample": f"C1 full size: the plain inner join of a 1M-row int32 build and a 6M-row "
                      f"int32 probe (every probe row matching once), the reference's sort/search "
                      f"join program restated in torch on the CPU (sort, two searchsorted, "
                      f"expand), {threads} threads, median of 3 ({t * 1e3:.1f} ms)"}


def cpu_baseline(build_cpu, probe_cpu, sample_rows: int):
    import numpy as np
    import oracle
    bk = build_cpu.numpy()
    pk = probe_cpu[:sample_rows].numpy()
    oracle.unique_hash_pairs([bk[:1000]], [pk[:1000]])  # warm the library
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        bi, pi = oracle.unique_hash_pairs([bk], [pk])
        oracle.gather_with_null(bk, bi)
        times.append(time.perf_counter() - t0)
    t = sorted(times)[1]
    return {"value": (bk.shape[0] + pk.shape[0]) / t, "unit": "tuples/s", "cores": 1,
            "kind": "port",
            "sample": f"the full {bk.shape[0]}-row build and the first {pk.shape[0]} probe "
                      f"rows of the same inputs; oracle C restatement of the reference's "
                      f"unique-hash join (table build, probe, probe-order pairs) plus the "
                      f"payload gather, one thread, median of 3 ({t:.2f} s)"}


def cpu_sample(fn, units: float, unit: str, what: str, reps: int = 3):
    """A tier's CPU baseline: the oracle on a bounded sample of the tier's
    own inputs, one thread, median of reps runs."""
    fn()  # warm (the oracle's C library loads on first use)
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t0)
    t = sorted(times)[len(times) // 2]
    return {"value": units / t, "unit": unit, "cores": 1, "kind": "port",
            "sample": f"{what}; median of {reps} ({t:.2f} s)"}


def partition_sweep(args, dev, probe):
    """The standalone radix partition of the large join's probe keys (int64)
    with int32 row ids into 2^bits buckets, for each bits of
    ``args.part_bits``.  Algorithmic bytes: keys and row ids read once and
    written once (24 B a row)."""
    import torch
    from pageflow_amd import _lib, radix

    np_ = probe.numel()
    rows = torch.arange(np_, dtype=torch.int32, device=dev)
    kout = torch.empty_like(probe)
    vout = torch.empty_like(rows)
    sweep = {}
    cases = [(int(b), int(sh)) for b in args.part_bits.split(",")
             for sh in args.part_shapes.split(",") if int(sh) != 4 or int(b) > 8]
    for bits, shape in cases:
        offs = torch.empty((1 << bits) + 1, dtype=torch.int64, device=dev)
        ws = _lib.workspace(radix.workspace_bytes(np_, bits), dev)
        run = lambda: radix.radix_partition(probe, bits, rows, out=(kout, vout),
                                            bucket_offsets=offs, ws=ws, shape=shape)
        for _ in range(args.warmup):
            run()
        torch.cuda.synchronize()
        _lib.profile(True)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            run()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        prof = _lib.profile_read()
        _lib.profile(False)
        ms = (t1 - t0) * 1e3 / args.steps
        sc = prof["rp_scatter"][0] / args.steps
        hi = prof["rp_hist"][0] / args.steps
        # the offsets end at n, and a sample of rows sits in its bucket
        assert int(offs[-1]) == np_
        key = f"{bits}/{shape}"
        sweep[key] = {"bits": bits, "shape": shape, "ms": ms, "rows_per_s": np_ / (ms * 1e-3),
                      "GBps": np_ * 24 / (ms * 1e-3) / 1e9,
                      "frac": np_ * 24 / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS,
                      "scatter_ms_per_step": sc, "hist_ms_per_step": hi,
                      "scatter_launches_per_step": prof["rp_scatter"][1] / args.steps}
        log(f"partition bits={bits} shape={shape}: {ms:.3f} ms "
            f"({sweep[key]['GBps']:.0f} GB/s, {sweep[key]['frac']:.3f} of peak), "
            f"scatter {sc:.3f} ms, hist {hi:.3f} ms a step")
    del kout, vout, rows
    return sweep


def bench_partition(args, dev) -> None:
    import torch
    from pageflow_amd.workloads import make_large_join

    nb, np_ = args.build_rows, args.probe_rows
    log(f"generating {np_} probe keys")
    _, _, probe, _ = make_large_join(nb, np_, 2, dev)
    sweep = partition_sweep(args, dev, probe)
    best = min(sweep, key=lambda b: sweep[b]["ms"])
    agg = bench_agg_pass1(args, dev, probe)
    line = {
        "metric": "radix partition rows/s (640M int64 keys + int32 row ids)",
        "value": sweep[best]["rows_per_s"], "unit": "rows/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": sweep[best]["ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int64",
        "data": "synthetic",
        "config": {"workload": "standalone radix partition of 640M int64 keys with int32 row "
                               "ids", "bits": sweep[best]["bits"], "shape": sweep[best]["shape"]},
        "roofline": {"bound": "hbm", "kernel": "radix partition (hist + scatter)",
                     "achieved": sweep[best]["GBps"], "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": sweep[best]["frac"], "traffic": None},
        "sweep": sweep,
        "agg_pass1": agg,
    }
    print(json.dumps(line), flush=True)


def bench_agg_pass1(args, dev, h_all):
    """Pass 1 of the partitioned aggregation on one tick: the first 128M of
    the large join's probe keys as the grouping hash, with an int64 and an
    int32 value, put into the padded layout of B = 256 buckets
    (_bucket_ids, then _partition_outputs: three outputs, one kernel call).
    Algorithmic bytes: the hash read for the bucket ids and the ids written
    (8 + 4), then the ids and each output read once and written once
    (4 + 2 (8 + 8 + 4)): 56 B a row."""
    import torch
    from pageflow_amd import _lib
    from pageflow_amd.operators import aggregate as agg

    n, B = min(h_all.numel(), 128_000_000), 256
    h = h_all[:n]
    g = torch.Generator(device=dev).manual_seed(9)
    v64 = torch.randint(0, 1 << 40, (n,), generator=g, device=dev)
    v32 = torch.randint(0, 1 << 30, (n,), generator=g, device=dev, dtype=torch.int32)
    bucket = agg._bucket_ids(h, B)
    counts_t = torch.bincount(bucket.long(), minlength=B)
    counts = counts_t.tolist()

    def run():
        b = agg._bucket_ids(h, B)
        return agg._partition_outputs([h, v64, v32], b, counts, counts_t)

    outs = run()
    # each bucket's hashes all map to it, and the rows keep their order
    starts = outs[0].bucket_starts()
    for k in (0, 77, 255):
        part = outs[0].tensor[starts[k]:starts[k] + counts[k]]
        assert bool((agg._bucket_ids(part, B) == k).all())
    del outs
    torch.cuda.synchronize()
    _lib.profile(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    prof = _lib.profile_read()
    _lib.profile(False)
    sc = prof["ps_scatter"][0] / prof["ps_scatter"][1]
    gbps = 56.0 * n / (ms * 1e-3) / 1e9

    # the reference's own steps in torch on the same GPU (aggregate.py:120-152):
    # a stable sort of the bucket ids, each row's position, the padding of
    # the buckets before it, one index_copy_ an output
    def torch_steps():
        b = agg._bucket_ids(h, B)
        order = torch.sort(b, stable=True).indices
        pos = torch.empty_like(order)
        pos[order] = torch.arange(n, dtype=order.dtype, device=dev)
        res = []
        for t in (h, v64, v32):
            unit = 512 // math.gcd(t.element_size(), 512)
            pad = (counts_t + unit - 1) // unit * unit - counts_t
            shift = torch.cumsum(pad, 0) - pad
            dest = pos + shift.index_select(0, b.long())
            out = torch.empty(int((counts_t + pad).sum()), dtype=t.dtype, device=dev)
            out.index_copy_(0, dest, t)
            res.append(out)
        return res

    torch_steps()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        torch_steps()
    torch.cuda.synchronize()
    torch_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    sc_gbps = 44.0 * n / (sc * 1e-3) / 1e9  # the scatter: ids + 3 outputs in, 3 outputs out
    cpu = None
    if not args.no_cpu:
        import oracle
        m = 4_000_000
        cols = [t[:m].cpu().numpy() for t in (h, v64, v32)]
        bk = bucket[:m].cpu().numpy()
        cpu = cpu_sample(lambda: oracle.partition_outputs(cols, bk, B), m, "rows/s",
                         f"the first {m} rows of the same tick, the oracle's padded "
                         f"partition of the three outputs (numpy stable sort)")
    log(f"aggregation pass 1: {n} rows into {B} buckets {ms:.2f} ms ({gbps:.0f} GB/s), "
        f"ps_scatter {sc:.2f} ms ({sc_gbps:.0f} GB/s); the reference's steps in torch "
        f"{torch_ms:.2f} ms")
    return {"rows": n, "B": B, "ms": ms, "rows_per_s": n / (ms * 1e-3), "GBps": gbps,
            "torch_reference_steps_ms": torch_ms, "speedup_vs_torch_steps": torch_ms / ms,
            "frac": gbps / HBM_PEAK_GBPS, "ps_scatter_ms": sc, "ps_scatter_GBps": sc_gbps,
            "ps_scatter_frac": sc_gbps / HBM_PEAK_GBPS,
            "kernels_ms": {k: v[0] / args.steps for k, v in prof.items()},
            "cpu_baseline": cpu}


def bench_batch(args, dev, emit: bool = True):
    """256 uneven inner joins (int32 keys): the batched pipeline (autotuned,
    replayed from a hipGraph) against two per-join baselines: the plain
    per-join launch of the reference's path (one unique-hash join call a
    join, int64 keys, one host wait a call) and, as an ablation, the batch
    kernels run one join at a time with one host wait a join.  tuples/s =
    (build + probe rows) / time; algorithmic bytes = sum over joins of
    4 nb + 4 np + 16 matches."""
    import numpy as np
    import torch
    from pageflow_amd import _lib
    from pageflow_amd.batch import BatchJoin, autotune, make_batch

    log("generating the 256-join batch")
    nb, np_, bkeys, pkeys, expect = make_batch(device=dev)
    tuples = float(nb.sum() + np_.sum())
    cache = os.path.join(os.path.expanduser("~"), ".cache", "pageflow_amd", "batch_autotune.json")
    os.makedirs(os.path.dirname(cache), exist_ok=True)
    tune = autotune(nb, np_, bkeys, pkeys, reps=2, cache=cache, device=dev)
    log(f"autotuned per class: {tune}")
    # the ablation: the batch kernels a join at a time, at the better of the
    # batched tuning and the same without probe slices (slicing has fixed
    # costs a join that only batching amortizes)
    t_plain, per_join_tune = None, None
    for tj in (tune, [tuple(t[:3]) + (0,) for t in tune]):
        base = BatchJoin(nb, np_, tj, dev)
        ts = []
        for _ in range(args.warmup + args.steps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            base.run(bkeys, pkeys, per_join=True)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        t = sorted(ts[args.warmup:])[len(ts[args.warmup:]) // 2]
        assert (base.counts.cpu().numpy() == expect).all()
        base.close()
        if t_plain is None or t < t_plain:
            t_plain, per_join_tune = t, tj
    # the plain per-join launch of the reference's path: one unique-hash
    # join call (join_unique_hash, int64 keys, one host wait) a join; the
    # keys are widened to int64 before the timed region
    from pageflow_amd.operators.join_core import unique_hash_pairs
    bo = np.concatenate(([0], np.cumsum(nb)))
    po = np.concatenate(([0], np.cumsum(np_)))
    b64, p64 = bkeys.to(torch.int64), pkeys.to(torch.int64)
    sides = [(b64[bo[j]:bo[j + 1]], p64[po[j]:po[j + 1]]) for j in range(len(nb))]
    t_uh = []
    for _ in range(args.warmup + args.steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for b_j, p_j in sides:
            unique_hash_pairs([b_j], [p_j])
        torch.cuda.synchronize()
        t_uh.append(time.perf_counter() - t0)
    t_uh = sorted(t_uh[args.warmup:])[len(t_uh[args.warmup:]) // 2]
    checked = [int(np.argmin(nb)), int(np.argsort(nb)[len(nb) // 2]), int(np.argmax(nb))]
    plain_pairs = {j: unique_hash_pairs([sides[j][0]], [sides[j][1]]) for j in checked}
    del b64, p64, sides
    bj = BatchJoin(nb, np_, tune, dev)
    st = bj.capture(bkeys, pkeys)
    bj.replay()
    st.synchronize()
    for j, (bi, pi) in plain_pairs.items():  # the same pairs (a sliced join's in slice order)
        op, ob = bj.pairs(j)
        order = torch.argsort(op, stable=True)
        assert torch.equal(op[order], pi) and torch.equal(ob[order], bi), j
    del plain_pairs
    for _ in range(args.warmup):
        bj.replay()
    st.synchronize()
    _lib.profile(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        bj.replay()
    st.synchronize()
    t_batch = (time.perf_counter() - t0) / args.steps
    _lib.profile(False)
    assert (bj.counts.cpu().numpy() == expect).all()
    # eager run of the same pipeline, profiled per kernel
    _lib.profile(True)
    bj.run(bkeys, pkeys)
    torch.cuda.synchronize()
    prof = _lib.profile_read()
    _lib.profile(False)
    pairs = int(expect.sum())
    alg = float(4 * nb.sum() + 4 * np_.sum() + 16 * pairs)
    res = {"joins": int(len(nb)), "build_rows": int(nb.sum()), "probe_rows": int(np_.sum()),
           "pairs": pairs, "tune": [list(t) for t in tune],
           "batched_ms": t_batch * 1e3, "per_join_ms": t_plain * 1e3,
           "plain_unique_hash_ms": t_uh * 1e3, "per_join_tune": [list(t) for t in per_join_tune],
           "batched_tuples_per_s": tuples / t_batch, "per_join_tuples_per_s": tuples / t_plain,
           "plain_unique_hash_tuples_per_s": tuples / t_uh,
           "speedup_vs_plain_unique_hash": t_uh / t_batch,
           "speedup_vs_same_kernels_per_join": t_plain / t_batch,
           "speedup": t_uh / t_batch,
           "batched_GBps": alg / t_batch / 1e9, "frac": alg / t_batch / 1e9 / HBM_PEAK_GBPS,
           "kernels_ms": {k: v[0] for k, v in prof.items()}}
    log(f"batch: batched {t_batch * 1e3:.2f} ms; plain unique-hash call a join {t_uh * 1e3:.2f} "
        f"ms ({t_uh / t_batch:.2f}x); the batch kernels a join at a time {t_plain * 1e3:.2f} ms "
        f"({t_plain / t_batch:.2f}x)")
    if emit:
        dom = max(prof, key=lambda k: prof[k][0])
        line = {"metric": "batched join tuples/s (256 uneven joins)", "value": tuples / t_batch,
                "unit": "tuples/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                "ms_per_step": t_batch * 1e3, "higher_is_better": True, "scaling": "weak",
                "vs_baseline": None, "dtype": "int32", "data": "synthetic",
                "config": {"workload": "256 inner joins, build log-uniform 1K-4M unique int32, "
                                       "probe 8x with 15% absent"},
                "roofline": {"bound": "hbm", "kernel": dom, "achieved": alg / t_batch / 1e9,
                             "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                             "frac": alg / t_batch / 1e9 / HBM_PEAK_GBPS, "traffic": None},
                "batch": res}
        print(json.dumps(line), flush=True)
    bj.close()
    return res


def bench_skew(args, dev, emit: bool = True):
    """The skewed multi-match join: build 10M rows over 2.5M distinct int64
    keys (Geometric(0.25) duplicates), probe 400M rows with Zipf(1.1) key
    ranks and 20% absent keys.  Outputs: the inner-join pairs (partitioned
    LDS join, count pass then write pass), the semi-join row mask and the
    anti-join row mask.  Algorithmic bytes: inner 8 nb + 8 np + 16 pairs;
    each mask 8 nb + 8 np + np."""
    import torch
    from pageflow_amd import _lib
    from pageflow_amd.join import HashJoin
    from pageflow_amd.operators.join_core import semi_anti_mask
    from pageflow_amd.workloads import make_skewed_join

    nb, nd, np_ = args.skew_build, args.skew_distinct, args.skew_probe
    log(f"generating the skewed join: {nb} build rows over {nd} keys, {np_} probe rows")
    build, probe, want = make_skewed_join(nb, nd, np_, 3, dev)
    plan = HashJoin(nb, np_, unique=False, device=dev)
    out_p = out_b = None

    def inner():
        nonlocal out_p, out_b
        n = plan.prepare(build, probe)
        if n < 0:
            raise RuntimeError("skewed join: a build bucket is too large for the LDS path")
        if out_p is None or out_p.numel() < n:
            out_p = torch.empty(n, dtype=torch.int64, device=dev)
            out_b = torch.empty(n, dtype=torch.int64, device=dev)
        plan.write(out_p, out_b)
        return n

    def timed(fn, reps):
        for _ in range(args.warmup):
            r = fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            r = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps, r

    _lib.profile(True)
    t_inner, pairs = timed(inner, args.steps)
    t_semi, semi = timed(lambda: semi_anti_mask(build, probe, anti=False), args.steps)
    t_anti, anti = timed(lambda: semi_anti_mask(build, probe, anti=True), args.steps)
    prof = _lib.profile_read()
    _lib.profile(False)
    # full-size checks: exact pair count, every pair's keys equal, each
    # probe row's pairs distinct, masks equal to key presence
    assert pairs == want["pairs"], (pairs, want)
    op, ob = out_p[:pairs], out_b[:pairs]
    for a in range(0, pairs, 1 << 27):
        assert bool((build[ob[a:a + (1 << 27)]] == probe[op[a:a + (1 << 27)]]).all())
    assert int(torch.unique(op * nb + ob).numel()) == pairs
    present = probe <= nd
    assert int(semi.sum()) == want["present"] and bool((semi == present).all())
    assert bool((anti == ~present).all())
    log(f"skew checks passed: {pairs} pairs, {want['present']} semi rows")
    b_inner = 8.0 * nb + 8.0 * np_ + 16.0 * pairs
    b_mask = 8.0 * nb + 9.0 * np_
    res = {"build_rows": nb, "distinct_keys": nd, "probe_rows": np_, "pairs": pairs,
           "semi_rows": want["present"], "bits": plan.bits,
           "inner_ms": t_inner * 1e3, "semi_ms": t_semi * 1e3, "anti_ms": t_anti * 1e3,
           "inner_tuples_per_s": (nb + np_) / t_inner,
           "inner_GBps": b_inner / t_inner / 1e9, "inner_frac": b_inner / t_inner / 8e12,
           "semi_GBps": b_mask / t_semi / 1e9, "anti_GBps": b_mask / t_anti / 1e9,
           "kernels_ms": {k: v[0] / max(v[1], 1) for k, v in prof.items()}}
    log(f"skew: inner {t_inner * 1e3:.2f} ms ({pairs} pairs), semi {t_semi * 1e3:.2f} ms, "
        f"anti {t_anti * 1e3:.2f} ms")
    res["residual"] = bench_residual(args, dev, ob, op, nb, np_, timed, build, probe, t_inner)
    if not args.no_cpu:
        import oracle
        bk = build[:2_500_000].cpu().numpy()
        ps = probe[:2_000_000].cpu().numpy()
        res["cpu_baseline"] = cpu_sample(
            lambda: oracle.sort_search_pairs(bk, ps), bk.shape[0] + ps.shape[0], "tuples/s",
            f"the first {bk.shape[0]} build and {ps.shape[0]} probe rows joined; the "
            f"oracle's sort/search inner join (numpy sort, two searchsorted, expand)")
        res["semi_cpu_baseline"] = cpu_sample(
            lambda: oracle.semi_anti(bk, ps, False), bk.shape[0] + ps.shape[0], "tuples/s",
            f"the same sample; the oracle's C semi-join hash set")
    if emit:
        line = {"metric": "skewed multi-match join tuples/s", "value": (nb + np_) / t_inner,
                "unit": "tuples/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                "ms_per_step": t_inner * 1e3, "higher_is_better": True, "scaling": "weak",
                "vs_baseline": None, "dtype": "int64", "data": "synthetic",
                "config": {"workload": "skewed multi-match join: build 10M rows over 2.5M "
                                       "int64 keys (geometric duplicates), probe 400M Zipf(1.1) "
                                       "keys, 20% absent; pairs + semi and anti row masks"},
                "roofline": {"bound": "hbm", "kernel": "inner join (all kernels)",
                             "achieved": b_inner / t_inner / 1e9, "peak": HBM_PEAK_GBPS,
                             "unit": "GB/s", "frac": b_inner / t_inner / 8e12, "traffic": None},
                "skew": res}
        print(json.dumps(line), flush=True)
    del out_p, out_b, plan
    return res


def bench_residual(args, dev, ob, op, nb: int, np_: int, timed, build=None, probe=None,
                   t_inner=None):
    """A residual condition on the skewed join's pairs: build payload <
    probe payload (int64, values in [0, 1000)), evaluated on every pair and
    the kept pairs compacted in pair order.  The pair filter (one
    evaluation pass, a scan, one compaction pass) against the reference's
    torch steps on the same GPU (two gathers, lt, the mask AND, nonzero, two
    index selects: operators/hashjoin.py:428-439).  Algorithmic bytes: the
    evaluation 32 B a pair (two indexes, two operands) + a mask bit, the
    compaction a mask bit + 32 B a kept pair (two indexes read, two
    written).

    ``join_fused``: the same condition evaluated inside the join
    (``HashJoin.prepare(residual=...)``: the probe operand carried through
    the partition, the build operand staged in LDS, only kept pairs
    written), timed against the join followed by the filter, and checked
    equal to the filter's pairs as a multiset (the condition takes the
    range form: each build run sorted by the build operand, a probe row's
    kept rows one sub-run).  Algorithmic bytes: keys and operands of both
    sides once (16 B a row) + 16 B a kept pair."""
    import torch
    from pageflow_amd import _lib
    from pageflow_amd.operators import residual

    g = torch.Generator(device=dev).manual_seed(11)
    bpay = torch.randint(0, 1000, (nb,), generator=g, device=dev)
    ppay = torch.randint(0, 1000, (np_,), generator=g, device=dev)
    n = ob.numel()
    prog = residual.compile_conditions({11: bpay}, {12: ppay}, [['b_pay#11', '<', 'p_pay#12']])
    _lib.profile(True)
    t_fused, (kb, kp) = timed(lambda: residual.filter_pairs(prog, ob, op), args.steps)
    prof = _lib.profile_read()
    _lib.profile(False)
    kept = kb.numel()

    def torch_steps():
        mask = torch.ones(n, dtype=torch.bool, device=dev)
        mask &= torch.lt(bpay[ob], ppay[op])
        pos = torch.nonzero(mask, as_tuple=False).view(-1)
        return ob[pos], op[pos]

    t_torch, (tb, tp) = timed(torch_steps, max(1, args.steps // 2))
    assert torch.equal(kb, tb) and torch.equal(kp, tp), "fused residual differs from torch"
    del tb, tp
    log(f"residual: {n} pairs -> {kept}, fused {t_fused * 1e3:.2f} ms, "
        f"torch steps {t_torch * 1e3:.2f} ms")
    kms = {k: v[0] / max(v[1], 1) for k, v in prof.items()}
    cpu = None
    if not args.no_cpu:
        import oracle
        m = 10_000_000
        lm, rm = ob[:m].cpu().numpy(), op[:m].cpu().numpy()
        L, R = {11: bpay.cpu().numpy()}, {12: ppay.cpu().numpy()}
        cond = [['b_pay#11', '<', 'p_pay#12']]
        cpu = cpu_sample(lambda: oracle.residual_filter(L, R, cond, lm, rm), m, "pairs/s",
                         f"the first {m} pairs; the oracle's residual evaluation (numpy "
                         f"gathers, compare, nonzero, two selects)")
    b_eval = 32.0 * n + n / 8.0
    b_comp = n / 8.0 + 32.0 * kept
    out = {"pairs": n, "kept": kept, "fused_ms": t_fused * 1e3, "torch_steps_ms": t_torch * 1e3,
           "speedup": t_torch / t_fused, "pairs_per_s": n / t_fused,
           "fused_GBps": (b_eval + b_comp) / t_fused / 1e9, "kernels_ms": kms,
           "cpu_baseline": cpu}
    if "rf_eval" in kms:
        out["rf_eval_GBps"] = b_eval / (kms["rf_eval"] * 1e-3) / 1e9
        out["rf_eval_frac"] = out["rf_eval_GBps"] / HBM_PEAK_GBPS
    if "rf_compact" in kms:
        out["rf_compact_GBps"] = b_comp / (kms["rf_compact"] * 1e-3) / 1e9
    if build is not None:
        from pageflow_amd.join import HashJoin
        plan = HashJoin(nb, np_, unique=False, device=dev)
        fo = {}

        def fused():
            n = plan.prepare(build, probe, residual=prog)
            if "p" not in fo or fo["p"].numel() < n:
                fo["p"] = torch.empty(n, dtype=torch.int64, device=dev)
                fo["b"] = torch.empty(n, dtype=torch.int64, device=dev)
            plan.write(fo["p"], fo["b"])
            return n

        _lib.profile(True)
        t_join_fused, n_fused = timed(fused, args.steps)
        prof_f = _lib.profile_read()
        _lib.profile(False)
        assert n_fused == kept and plan.candidates == n, (n_fused, kept, plan.candidates, n)
        # the same pairs (the range form orders a probe row's build rows by
        # the build operand): sorted (build, probe) keys compared
        got = torch.sort(fo["b"][:kept] * np_ + fo["p"][:kept]).values
        want = torch.sort(kb * np_ + kp).values
        assert torch.equal(got, want), "the residual inside the join differs from the filter"
        del got, want
        b_jf = 16.0 * (nb + np_) + 16.0 * kept
        out["join_fused"] = {
            "ms": t_join_fused * 1e3, "join_then_filter_ms": (t_inner + t_fused) * 1e3,
            "speedup": (t_inner + t_fused) / t_join_fused, "kept": kept,
            "GBps": b_jf / t_join_fused / 1e9, "frac": b_jf / t_join_fused / 8e12,
            "kernels_ms": {k: v[0] / max(v[1], 1) for k, v in prof_f.items()}}
        log(f"residual inside the join: {t_join_fused * 1e3:.2f} ms against join + filter "
            f"{(t_inner + t_fused) * 1e3:.2f} ms")
        del plan, fo
    del kb, kp, bpay, ppay
    torch.cuda.empty_cache()
    return out


def bench_serve(args, dev, emit: bool = True):
    """The serving workload: a Poisson stream of fact->dimension join +
    group-by requests (fact rows log-uniform 100K-50M, 8 recurring
    dimensions of 1K-2M rows, 25 groups) through the continuous-batching
    engine.  Capacity is measured closed-loop (every request arrived at the
    start); the open-loop run offers 0.7 of it and reports p50/p99 latency
    (arrival -> result on the host).  Algorithmic bytes: 12 a fact row,
    8 a dimension row built, 400 a result."""
    import numpy as np
    import torch
    from pageflow_amd import _lib
    from pageflow_amd import serving

    log(f"generating the serving workload ({args.serve_requests} requests)")
    wl = serving.make_workload(args.serve_requests, rate=0.0, seed=5, device=dev)
    eng = serving.ServingEngine(dev, hbm_budget=args.serve_budget << 20, max_active=32,
                                step_rows=args.serve_step_rows, rows_per_wg=1 << 16)
    for d, (k, g) in wl.dims.items():
        eng.register_dim(d, k, g, wl.G)
    eng.run(wl.requests[:16], realtime=False)  # warm-up: builds the 8 tables
    eng.check_tables()
    fact_rows = sum(r.fk.numel() for r in wl.requests)
    t0 = time.perf_counter()
    closed = eng.run(wl.requests, realtime=False)
    t_closed = time.perf_counter() - t0
    capacity = len(wl.requests) / t_closed
    rate = args.serve_load * capacity
    log(f"serving capacity {capacity:.0f} requests/s ({fact_rows / t_closed / 1e9:.2f} G fact "
        f"rows/s); offering {rate:.0f} requests/s")
    reqs = serving.with_arrivals(wl.requests, rate, seed=5)
    steps0 = eng.stats["steps"]
    _lib.profile(True)
    t0 = time.perf_counter()
    opened = eng.run(reqs)
    t_open = time.perf_counter() - t0
    prof = _lib.profile_read()
    _lib.profile(False)
    steps = eng.stats["steps"] - steps0
    # the mix does not change a result; every fact key has a dimension row
    for a, b in zip(closed, opened):
        assert np.array_equal(a.sum, b.sum) and np.array_equal(a.cnt, b.cnt), a.rid
    for r, res in zip(wl.requests, opened):
        assert int(res.cnt.sum()) == r.fk.numel(), r.rid
    lat = serving.latency_summary(opened)
    pressure = serve_under_pressure(args, dev, wl, closed, eng)
    ss_ms, ss_n = prof.get("serve_step", (0.0, 0))
    step_bytes = 12.0 * fact_rows / max(steps, 1)
    achieved = step_bytes / (ss_ms / max(ss_n, 1) * 1e-3) / 1e9 if ss_n else None
    res = {"requests": len(reqs), "fact_rows": int(fact_rows), "dims": len(wl.dims),
           "dim_rows": wl.dim_rows, "offered_rate_rps": rate, "load": args.serve_load,
           "capacity_rps": capacity, "closed_loop_s": t_closed, "open_loop_s": t_open,
           "tuples_per_s": fact_rows / t_open, "closed_loop_tuples_per_s": fact_rows / t_closed,
           **lat, "steps": steps, "max_items_per_step": eng.stats["max_items"],
           "table_builds": eng.stats["builds"], "evictions": eng.stats["evictions"],
           "hbm_budget_MiB": args.serve_budget, "peak_engine_MiB": eng.peak_used / 2**20,
           "serve_step_avg_ms": ss_ms / max(ss_n, 1),
           "serve_step_GBps": achieved,
           "serve_step_frac": achieved / HBM_PEAK_GBPS if achieved else None,
           "pressure": pressure}
    log(f"serving: p50 {lat['p50_ms']:.2f} ms, p99 {lat['p99_ms']:.2f} ms, "
        f"{fact_rows / t_open / 1e9:.2f} G fact rows/s over {len(reqs)} requests")
    if emit:
        line = {"metric": "serving join+group-by tuples/s (p99 latency beside it)",
                "value": fact_rows / t_open, "unit": "tuples/s", "n_gpus": 1,
                "steps": len(reqs), "warmup": 16, "ms_per_step": t_open * 1e3 / len(reqs),
                "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                "dtype": "int64", "data": "synthetic",
                "config": {"workload": "Poisson stream of join+group-by requests: fact "
                                       "log-uniform 100K-50M rows, 8 resident dimensions "
                                       "1K-2M rows, 25 groups"},
                "roofline": {"bound": "hbm", "kernel": "serve_step", "achieved": achieved,
                             "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                             "frac": achieved / HBM_PEAK_GBPS if achieved else None,
                             "traffic": None},
                "serving": res}
        print(json.dumps(line), flush=True)
    return res


def serve_under_pressure(args, dev, wl, reference, base_engine):
    """The same request stream through an engine whose HBM budget holds
    only args.serve_pressure of the dimension tables' total bytes (never
    less than the largest table with every slot's accumulators): requests
    whose table is not resident wait for room, tables are evicted and built
    again.  Closed-loop capacity, then an open-loop run offering 0.7 of it,
    with its p50/p99; every result equals the unpressured engine's."""
    import numpy as np
    from pageflow_amd import serving

    dims = base_engine.dims
    tables = sum(d.table_bytes for d in dims.values())
    acc = 32 * 16 * wl.G
    largest = max(d.table_bytes for d in dims.values())
    budget = max(int(args.serve_pressure * tables), largest + acc)
    eng = serving.ServingEngine(dev, hbm_budget=budget, max_active=32,
                                step_rows=args.serve_step_rows, rows_per_wg=1 << 16)
    for d, (k, g) in wl.dims.items():
        eng.register_dim(d, k, g, wl.G)
    eng.run(wl.requests[:16], realtime=False)
    t0 = time.perf_counter()
    closed = eng.run(wl.requests, realtime=False)
    t_closed = time.perf_counter() - t0
    capacity = len(wl.requests) / t_closed
    reqs = serving.with_arrivals(wl.requests, args.serve_load * capacity, seed=5)
    ev0, b0 = eng.stats["evictions"], eng.stats["builds"]
    t0 = time.perf_counter()
    opened = eng.run(reqs)
    t_open = time.perf_counter() - t0
    for a, b, c in zip(reference, closed, opened):
        assert np.array_equal(a.sum, b.sum) and np.array_equal(a.cnt, b.cnt), a.rid
        assert np.array_equal(a.sum, c.sum) and np.array_equal(a.cnt, c.cnt), a.rid
    lat = serving.latency_summary(opened)
    fact_rows = sum(r.fk.numel() for r in wl.requests)
    res = {"hbm_budget_MiB": budget / 2**20, "tables_total_MiB": tables / 2**20,
           "budget_frac_of_tables": budget / tables, "capacity_rps": capacity,
           "offered_rate_rps": args.serve_load * capacity, **lat,
           "tuples_per_s": fact_rows / t_open,
           "evictions": eng.stats["evictions"] - ev0, "table_builds": eng.stats["builds"] - b0,
           "peak_engine_MiB": eng.peak_used / 2**20}
    log(f"serving under a {budget / 2**20:.1f} MiB budget ({budget / tables:.2f} of the "
        f"tables): capacity {capacity:.0f} requests/s, p50 {lat['p50_ms']:.2f} ms, p99 "
        f"{lat['p99_ms']:.2f} ms, {res['evictions']} evictions, {res['table_builds']} builds")
    assert res["evictions"] > 0, "the budget did not bind"
    return res


def bench_uva(args, dev, emit: bool = True):
    """The out-of-core payload gather of a join's late materialisation: an
    int64 payload column of args.uva_rows rows in pinned host memory (64 MiB
    segments), gathered on the GPU over its host link by args.uva_out row
    ids (uniform, 1% null rows -64, as an outer join emits), and the masked
    read of a scan over the same column (30% of the rows kept).  The link,
    not HBM, bounds both: the roofline peak is the measured host-to-device
    copy rate of a pinned buffer.  Useful bytes: 8 a gathered row read over
    the link (nulls are not read); the link moves whole 128-byte host lines,
    so its roofline counts the distinct lines holding a read row."""
    import numpy as np
    import torch
    from pageflow_amd import _lib
    from pageflow_amd import zero_copy_reader as zc

    n, n_out = args.uva_rows, args.uva_out
    seg = 1 << 23  # rows a 64 MiB segment
    log(f"pinning a {n * 8 / 2**30:.1f} GiB host column ({n} int64 rows)")
    parts = []
    for a in range(0, n, seg):
        t = torch.empty(min(seg, n - a), dtype=torch.int64, pin_memory=True)
        torch.arange(a, a + t.numel(), dtype=torch.int64, out=t)
        t.mul_(3).add_(7)  # row r holds 3r + 7
        parts.append(t)
    ptrs, rows = zc.segments(parts, device=dev)
    g = torch.Generator(device=dev).manual_seed(6)
    idx = torch.randint(0, n, (n_out,), generator=g, device=dev)
    idx[::100] = -64
    mask = torch.rand(n, generator=g, device=dev) < 0.3
    kept = int(mask.sum())
    mres = torch.empty(kept, dtype=torch.int64, device=dev)

    def gather():
        return zc.zero_copy_rearrange(ptrs, rows, idx, 0, n, torch.int64, 1)

    out = gather()
    want = torch.where(idx == -64, torch.full_like(idx, -64), idx * 3 + 7)
    assert torch.equal(out.view(-1), want), "gather disagrees with the column's closed form"
    zc.zero_copy_mask(ptrs, rows, mask, mres, 0, n)
    assert torch.equal(mres, torch.nonzero(mask).view(-1) * 3 + 7)
    del out, want
    torch.cuda.synchronize()
    steps = max(1, args.steps)
    _lib.profile(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        gather()
    torch.cuda.synchronize()
    t_gather = (time.perf_counter() - t0) / steps
    t0 = time.perf_counter()
    for _ in range(steps):
        zc.zero_copy_mask(ptrs, rows, mask, mres, 0, n)
    torch.cuda.synchronize()
    t_mask = (time.perf_counter() - t0) / steps
    prof = _lib.profile_read()
    _lib.profile(False)
    # the link's practical peak: a pinned host -> device copy of 1 GiB
    h = parts[0] if parts[0].numel() * 8 >= 1 << 30 else torch.empty(1 << 27, dtype=torch.int64,
                                                                      pin_memory=True)
    d = torch.empty(h.numel(), dtype=torch.int64, device=dev)
    d.copy_(h, non_blocking=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        d.copy_(h, non_blocking=True)
    torch.cuda.synchronize()
    h2d = h.numel() * 8 * 3 / (time.perf_counter() - t0) / 1e9
    read = 8.0 * (n_out - (n_out + 99) // 100)
    g_gbps = read / t_gather / 1e9
    m_gbps = 8.0 * kept / t_mask / 1e9
    # what the link must move: every 128-byte host line holding a read row
    line_rows = 16
    g_lines = int(torch.unique(idx[idx != -64] // line_rows).numel())
    m_lines = int(torch.unique(torch.nonzero(mask).view(-1) // line_rows).numel())
    g_link = g_lines * 128.0 / t_gather / 1e9
    m_link = m_lines * 128.0 / t_mask / 1e9
    kern = {k: {"total_ms": v[0] / steps, "launches": v[1] / steps} for k, v in prof.items()}
    cpu = None
    if not args.no_cpu:
        import oracle
        m = 8_000_000
        hcol = [torch.cat(parts).numpy().reshape(-1, 1)]  # one host part
        sidx = idx[:m].cpu().numpy()
        cpu = cpu_sample(lambda: oracle.zero_copy_rearrange(hcol, 0, sidx, 0), m, "rows/s",
                         f"the first {m} of the same indices over the same column in host "
                         f"memory; the oracle's gather (numpy), read by the host")
    res = {"column_rows": n, "gather_rows": n_out, "null_every": 100, "mask_kept": kept,
           "gather_ms": t_gather * 1e3, "gather_rows_per_s": n_out / t_gather,
           "gather_link_GBps": g_gbps, "mask_ms": t_mask * 1e3, "mask_link_GBps": m_gbps,
           "h2d_copy_GBps": h2d, "gather_lines_GBps": g_link, "mask_lines_GBps": m_link,
           "gather_frac_of_h2d": g_link / h2d, "mask_frac_of_h2d": m_link / h2d,
           "cpu_baseline": cpu,
           "kernels_ms": kern}
    log(f"uva: gather {n_out} rows {t_gather * 1e3:.2f} ms ({g_gbps:.1f} GB/s of rows, "
        f"{g_link:.1f} GB/s of 128-B host lines = {g_link / h2d:.2f} of a {h2d:.1f} GB/s H2D "
        f"copy); mask {kept} rows {t_mask * 1e3:.2f} ms ({m_gbps:.1f} GB/s of rows, "
        f"{m_link / h2d:.2f} of the copy)")
    if emit:
        line = {"metric": "zero-copy gather rows/s over the host link",
                "value": n_out / t_gather, "unit": "rows/s", "n_gpus": 1, "steps": steps,
                "warmup": 1, "ms_per_step": t_gather * 1e3, "higher_is_better": True,
                "scaling": "weak", "vs_baseline": None, "dtype": "int64", "data": "synthetic",
                "config": {"workload": f"gather {n_out} of {n} int64 rows of a pinned host "
                                       "column (64 MiB segments, 1% null rows)"},
                "roofline": {"bound": "host link", "kernel": "zc_gather", "achieved": g_link,
                             "peak": h2d, "unit": "GB/s", "frac": g_link / h2d,
                             "traffic": g_lines * 128},
                "uva": res}
        print(json.dumps(line), flush=True)
    return res


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--build-rows", type=int, default=16_000_000)
    ap.add_argument("--probe-rows", type=int, default=640_000_000)
    ap.add_argument("--bits", type=int, default=None)
    ap.add_argument("--ordered", action="store_true",
                    help="the large join's pairs in the deterministic partition order "
                         "(adds the count pass)")
    ap.add_argument("--cpu-sample", type=int, default=16_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=["join", "partition", "batch", "serve", "skew", "uva"],
                    default="join")
    ap.add_argument("--uva-rows", type=int, default=1 << 28)
    ap.add_argument("--uva-out", type=int, default=1 << 26)
    ap.add_argument("--skew-build", type=int, default=10_000_000)
    ap.add_argument("--skew-distinct", type=int, default=2_500_000)
    ap.add_argument("--skew-probe", type=int, default=400_000_000)
    ap.add_argument("--tiers", choices=["all", "none"], default="all",
                    help="with the join workload: also run the batch and serving tiers and "
                         "add their results to the line")
    ap.add_argument("--serve-requests", type=int, default=2048)
    ap.add_argument("--serve-load", type=float, default=0.7)
    ap.add_argument("--serve-budget", type=int, default=8192, help="MiB")
    ap.add_argument("--serve-step-rows", type=int, default=1 << 25)
    ap.add_argument("--serve-pressure", type=float, default=0.5,
                    help="the pressured serving run's budget, a fraction of the tables' bytes")
    ap.add_argument("--part-bits", default="8,9,10,11,12")
    ap.add_argument("--part-shapes", default="-1")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the join's multi-GPU sharding is a design, not a measured path)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from pageflow_amd import _lib
    from pageflow_amd.join import HashJoin
    from pageflow_amd.workloads import make_large_join

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.workload == "partition":
        bench_partition(args, dev)
        return
    if args.workload == "batch":
        bench_batch(args, dev)
        return
    if args.workload == "serve":
        bench_serve(args, dev)
        return
    if args.workload == "skew":
        bench_skew(args, dev)
        return
    if args.workload == "uva":
        bench_uva(args, dev)
        return
    nb, np_ = args.build_rows, args.probe_rows
    log(f"generating {nb} build / {np_} probe rows")
    build, pays, probe, expect = make_large_join(nb, np_, 2, dev)
    torch.cuda.synchronize()
    # pairs in any order: no count pass (the pair order is not part of the
    # join's result, which is compared as a multiset)
    plan = HashJoin(nb, np_, unique=True, bits=args.bits, device=dev,
                    ordered=args.ordered)
    out_p = torch.empty(np_, dtype=torch.int64, device=dev)
    out_b = torch.empty(np_, dtype=torch.int64, device=dev)
    out_pay = torch.empty(np_, dtype=torch.int64, device=dev)
    log(f"plan: bits={plan.bits} workspace={plan.ws_bytes / 2**30:.2f} GiB")

    def step():
        plan.prepare(build, probe)
        return plan.write(out_p, out_b, [pays[0]], [out_pay])

    for i in range(args.warmup):
        n = step()
        log(f"warmup {i}: {n} pairs")
    torch.cuda.synchronize()
    _lib.profile(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        n = step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = _lib.profile_read()
    _lib.profile(False)
    ms = (t1 - t0) * 1e3 / args.steps
    log(f"timed {args.steps} steps: {ms:.3f} ms/step, {n} pairs")

    # size-independent checks at full size: every present probe row matched
    # once, each pair's keys equal, the payload gathered by build row
    pairs = n
    assert pairs == expect, (pairs, expect)
    op, ob = out_p[:pairs], out_b[:pairs]
    assert bool((build[ob] == probe[op]).all())
    assert bool((out_pay[:pairs] == pays[0][ob]).all())
    assert int(torch.unique(op).numel()) == pairs
    log("checks passed: pair count, key equality, payload gather, unique probe rows")

    # the achievable streaming rate on this box: a 5 GB device copy
    src = probe[: min(np_, 640_000_000)]
    dst = out_b[: src.numel()]
    dst.copy_(src)
    torch.cuda.synchronize()
    tc = time.perf_counter()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    copy_gbps = 2 * src.numel() * 8 * 3 / (time.perf_counter() - tc) / 1e9
    log(f"device copy: {copy_gbps:.0f} GB/s")

    kernels = {}
    for name, (tot_ms, launches) in prof.items():
        per_step = launches / args.steps
        alg = kernel_bytes(name, nb, np_, pairs, per_step)
        kernels[name] = {"avg_ms": tot_ms / launches, "launches_per_step": per_step,
                         "algorithmic_bytes_per_launch": alg,
                         "GBps": alg / (tot_ms / launches * 1e-3) / 1e9}
    dom = max(prof, key=lambda k: prof[k][0])
    dom_ms = prof[dom][0] / prof[dom][1]
    per_launch = kernel_bytes(dom, nb, np_, pairs, prof[dom][1] / args.steps)
    achieved = per_launch / (dom_ms * 1e-3) / 1e9
    # traffic: the rocprofv3 FETCH_SIZE + WRITE_SIZE reading of this kernel
    # committed under profiles/ (counters need their own profiler run); the
    # line names the file and the commit it was measured at
    traffic, traffic_source = None, None
    tpath = os.path.join(ROOT, "profiles", f"{dom}_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            rec = json.load(f)
        traffic = rec.get("bytes_per_launch")
        traffic_source = (f"profiles/{dom}_traffic.json: rocprofv3 --pmc FETCH_SIZE and "
                          f"WRITE_SIZE runs, scaled by a device copy's readings; measured at "
                          f"commit {rec.get('commit', 'unrecorded')} ({rec.get('date', 'undated')})")
    e2e_bytes = nb * 8 + np_ * 8 + pairs * 16 + pairs * 16
    tuples = (nb + np_) / (ms * 1e-3)
    # the probe side's complete partition (histogram and both scatter
    # passes): 24 B a row (keys and row ids read once and written once)
    part_ms = sum(prof[k][0] for k in ("hj_probe_hist", "hj_probe_scatter",
                                       "hj_probe_scatter_2") if k in prof) / args.steps
    partition = {"kernels": ["hj_probe_hist", "hj_probe_scatter", "hj_probe_scatter_2"],
                 "ms_per_step": part_ms, "algorithmic_bytes": np_ * 24,
                 "GBps": np_ * 24 / (part_ms * 1e-3) / 1e9,
                 "frac": np_ * 24 / (part_ms * 1e-3) / 1e9 / HBM_PEAK_GBPS}

    bits = plan.bits
    tiers = {}
    log("standalone partition sweep of the same probe keys")
    del out_pay, src, dst, op, ob
    del plan
    torch.cuda.empty_cache()
    tiers["partition"] = partition_sweep(args, dev, probe)

    cpu = cpu_c2 = None
    if not args.no_cpu:
        log("timing the CPU baselines")
        cpu = cpu_baseline_c1()
        cpu_c2 = cpu_baseline(build.cpu(), probe[:args.cpu_sample].cpu(), args.cpu_sample)

    if args.tiers == "all":
        del build, pays, probe, out_p, out_b
        torch.cuda.empty_cache()
        tiers["skew"] = bench_skew(args, dev, emit=False)
        torch.cuda.empty_cache()
        tiers["batch"] = bench_batch(args, dev, emit=False)
        torch.cuda.empty_cache()
        tiers["serving"] = bench_serve(args, dev, emit=False)
        torch.cuda.empty_cache()
        tiers["uva"] = bench_uva(args, dev, emit=False)

    line = {
        "metric": "join tuples/s + % HBM roofline, batched vs per-join launch; p99 ms (serving)",
        "value": tuples,
        "unit": "tuples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic",
        "config": {"workload": "large plain join: build 16M unique int64 keys + 2 int64 "
                               "payloads, probe 640M int64 keys (10% absent); pairs + 1 "
                               "gathered payload",
                   "build_rows": nb, "probe_rows": np_, "pairs": pairs, "bits": bits,
                   "pair_order": "partition order" if args.ordered else "any (no count pass)"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved,
                     "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBPS,
                     "traffic": traffic, "traffic_source": traffic_source,
                     "algorithmic_bytes_per_launch": per_launch, "avg_launch_ms": dom_ms,
                     "probe_partition": partition},
        "e2e": {"algorithmic_bytes": e2e_bytes,
                "GBps": e2e_bytes / (ms * 1e-3) / 1e9,
                "frac": e2e_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS},
        "kernels": kernels,
        "copy_GBps": copy_gbps,
        "cpu_baseline": cpu,
        "cpu_baseline_c2_sample": cpu_c2,
        **tiers,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
