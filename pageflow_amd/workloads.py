"""Seeded synthetic inputs of the join workloads, generated on the device.

* ``make_large_join``: a unique int64 build (a permutation of 1..nb) with
  two int64 payload columns, and a probe uniform over the build keys with
  10% of its rows replaced by absent keys from [nb + 1, 2^62).
* ``make_skewed_join``: a multi-match build of nd distinct int64 keys (a
  permutation of 1..nd) with Geometric(0.25) duplicates (mean 4, at least
  one), adjusted to exactly nb rows and shuffled; a probe whose key ranks
  follow Zipf(s) over the nd keys (rank -> key by a fixed permutation) with
  a fraction replaced by absent keys from [nd + 1, 2^62).

The batching and serving tiers' generators are ``batch.make_batch`` and
``serving.make_workload``.
"""
from __future__ import annotations

from typing import Dict, Tuple

import numpy as np
import torch


def make_large_join(nb: int, np_: int, seed: int, device, absent: float = 0.1):
    """(build keys, (payload 0, payload 1), probe keys, matching probe rows)."""
    g = torch.Generator(device=device).manual_seed(seed)
    build = torch.randperm(nb, generator=g, device=device) + 1
    pay0 = torch.randint(-(1 << 62), 1 << 62, (nb,), generator=g, device=device)
    pay1 = torch.randint(-(1 << 62), 1 << 62, (nb,), generator=g, device=device)
    idx = torch.randint(0, nb, (np_,), generator=g, device=device)
    probe = build[idx]
    del idx
    miss = torch.rand(np_, generator=g, device=device) < absent
    n_abs = int(miss.sum())
    probe[miss] = torch.randint(nb + 1, 1 << 62, (n_abs,), generator=g, device=device)
    del miss
    return build, (pay0, pay1), probe, np_ - n_abs


def geometric_duplicates(nd: int, nb: int, p: float, rng: np.random.Generator) -> np.ndarray:
    """Rows per distinct key: Geometric(p) (support >= 1), then one row more
    for the first keys (cyclically) or one row less for the first keys with
    more than one, until the total is exactly nb (nd <= nb)."""
    if nd > nb:
        raise ValueError("more distinct keys than build rows")
    c = rng.geometric(p, nd).astype(np.int64)
    diff = nb - int(c.sum())
    while diff != 0:
        if diff > 0:
            k = min(diff, nd)
            c[:k] += 1
            diff -= k
        else:
            idx = np.nonzero(c > 1)[0][:-diff]
            c[idx] -= 1
            diff += idx.size
    return c


def make_skewed_join(nb: int, nd: int, np_: int, seed: int, device, s: float = 1.1,
                     absent: float = 0.2, chunk: int = 1 << 26
                     ) -> Tuple[torch.Tensor, torch.Tensor, Dict[str, int]]:
    """(build keys, probe keys, {"pairs": inner-join pairs, "present": probe
    rows with a match}) of the skewed multi-match workload."""
    rng = np.random.default_rng(seed)
    dup = geometric_duplicates(nd, nb, 0.25, rng)
    keys = rng.permutation(nd).astype(np.int64) + 1          # distinct key i -> keys[i]
    build_np = np.repeat(keys, dup)[rng.permutation(nb)]
    rank_key = torch.from_numpy(keys[rng.permutation(nd)]).to(device)  # Zipf rank -> key
    w = np.arange(1, nd + 1, dtype=np.float64) ** (-s)
    cdf = torch.from_numpy(np.cumsum(w) / w.sum()).to(device)
    g = torch.Generator(device=device).manual_seed(seed)
    probe = torch.empty(np_, dtype=torch.int64, device=device)
    for a in range(0, np_, chunk):
        b = min(np_, a + chunk)
        u = torch.rand(b - a, generator=g, device=device, dtype=torch.float64)
        r = torch.searchsorted(cdf, u).clamp_(max=nd - 1)
        probe[a:b] = rank_key[r]
        miss = torch.rand(b - a, generator=g, device=device) < absent
        probe[a:b][miss] = torch.randint(nd + 1, 1 << 62, (int(miss.sum()),), generator=g,
                                         device=device)
    build = torch.from_numpy(build_np).to(device)
    # exact answer sizes: build rows per key value (keys are 1..nd)
    per_key = torch.zeros(nd + 1, dtype=torch.int64, device=device)
    per_key[torch.from_numpy(keys).to(device)] = torch.from_numpy(dup).to(device)
    pairs = present = 0
    for a in range(0, np_, chunk):
        k = probe[a:a + chunk]
        hit = k <= nd
        pairs += int(per_key[k[hit]].sum())
        present += int(hit.sum())
    return build, probe, {"pairs": pairs, "present": present}


__all__ = ["make_large_join", "make_skewed_join", "geometric_duplicates"]
