"""Pass 1 of the partitioned aggregation (pf_partition_by_bucket,
pageflow_amd.operators.aggregate): every output's buckets equal the
oracle's stable padded layout for the buckets _bucket_ids gives, at B from
1 to 512, for outputs of every row width (more outputs than one kernel call
takes), empty ticks; the layout equals the join partition's with the hash as
the key; a wrong count or an out-of-range bucket raises."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from pageflow_amd.operators import aggregate as agg
from pageflow_amd.operators.partition_sort import partition

pytestmark = pytest.mark.gpu


def _outs(rng, n, dev):
    return [torch.from_numpy(rng.integers(-2**62, 2**62, n)).to(dev),
            torch.from_numpy(rng.integers(0, 2**31 - 1, (n, 3)).astype(np.int32)).to(dev),
            torch.from_numpy(rng.integers(0, 255, n).astype(np.uint8)).to(dev),
            torch.from_numpy(rng.standard_normal(n)).to(dev),
            torch.from_numpy(rng.integers(-100, 100, (n, 5)).astype(np.int16)).to(dev),
            torch.from_numpy(rng.integers(-2**30, 2**30, n).astype(np.int32)).to(dev),
            torch.from_numpy(rng.integers(0, 9, (n, 2)).astype(np.int64)).to(dev)]


@pytest.mark.parametrize("B", [1, 16, 512])
@pytest.mark.parametrize("n", [0, 3001, 300_000])
def test_partition_outputs_match_oracle(B, n):
    dev = torch.device("cuda")
    rng = np.random.default_rng(B * 7 + n)
    h = torch.from_numpy(rng.integers(-2**63, 2**63 - 1, n, endpoint=True)).to(dev)
    bucket = agg._bucket_ids(h, B)
    assert np.array_equal(bucket.cpu().numpy(), oracle.bucket_of(h.cpu().numpy(), B))
    counts_t = torch.bincount(bucket.long(), minlength=B)
    counts = counts_t.tolist()
    outs = _outs(rng, n, dev)
    got = agg._partition_outputs(outs, bucket, counts, counts_t)
    want_counts, want = oracle.partition_outputs([t.cpu().numpy() for t in outs],
                                                 bucket.cpu().numpy(), B)
    assert counts == want_counts
    for g, w in zip(got, want):
        assert g.partition_row_counts == counts
        assert g.tensor.shape == w.shape
        gb = [x.cpu().numpy() for x in g.buckets()]
        starts, _ = oracle.padded_layout(counts, w.dtype.itemsize * int(np.prod(w.shape[1:])))
        for k in range(B):
            assert np.array_equal(gb[k], w[starts[k]:starts[k] + counts[k]]), k


def test_layout_equals_the_join_partition_with_the_hash_as_key():
    dev = torch.device("cuda")
    rng = np.random.default_rng(1)
    n, B = 100_000, 64
    h = torch.from_numpy(rng.integers(-2**63, 2**63 - 1, n, endpoint=True)).to(dev)
    v = torch.from_numpy(rng.integers(0, 1000, n).astype(np.int32)).to(dev)
    bucket = agg._bucket_ids(h, B)
    counts_t = torch.bincount(bucket.long(), minlength=B)
    mine = agg._partition_outputs([h, v], bucket, counts_t.tolist(), counts_t)
    theirs = partition([h, v], 0, None, 0, B)
    for a, b in zip(mine, theirs[:2]):
        assert a.partition_row_counts == b.partition_row_counts
        for x, y in zip(a.buckets(), b.buckets()):
            assert torch.equal(x, y)


def test_wrong_counts_and_out_of_range_buckets_raise():
    dev = torch.device("cuda")
    v = torch.arange(1000, dtype=torch.int64, device=dev)
    bucket = (v % 8).to(torch.int32)
    counts_t = torch.bincount(bucket.long(), minlength=8)
    bad_counts = counts_t.tolist()
    bad_counts[0] += 1
    with pytest.raises(RuntimeError, match="disagree"):
        agg._partition_outputs([v], bucket, bad_counts, counts_t)
    bucket[17] = 8
    with pytest.raises(RuntimeError, match="outside"):
        agg._partition_outputs([v], bucket, counts_t.tolist(), counts_t)
    with pytest.raises(ValueError):
        agg._bucket_ids(v, 12)
