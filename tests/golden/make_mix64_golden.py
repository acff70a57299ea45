"""Writes tests/golden/mix64_golden.npz: (key, mix64(key), bucket at B=512)
vectors computed by the reference's own numpy twin of the key hash
(``tests/mix64.py:12-36`` of the reference repository), so the oracle's mix64
and bucket_of are pinned to the reference rather than to a retyped formula.

Run in the build container, where the reference is mounted:

    python tests/golden/make_mix64_golden.py /root/reference

The fixture is data (inputs and the reference's outputs); the reference
itself never travels with this repository.
"""
from __future__ import annotations

import importlib.util
import os
import sys

import numpy as np


def main(ref_root: str) -> None:
    spec = importlib.util.spec_from_file_location(
        "ref_mix64", os.path.join(ref_root, "tests", "mix64.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    i64 = np.iinfo(np.int64)
    rng = np.random.default_rng(20240611)
    keys = rng.integers(i64.min, i64.max, 1000, dtype=np.int64)
    keys[:8] = [0, 1, -1, i64.min, i64.max, 1 << 33, -64, 16_000_000]
    hashed = mod.mix64_np(keys)
    import torch
    bucket512 = mod.bucket_of(torch.from_numpy(keys.copy()), 512).numpy()
    bucket1 = mod.bucket_of(torch.from_numpy(keys.copy()), 1).numpy()
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mix64_golden.npz")
    np.savez(out, keys=keys, mix64=hashed, bucket512=bucket512.astype(np.int64),
             bucket1=bucket1.astype(np.int64))
    print("wrote", out, keys.shape[0], "vectors")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
