"""The skewed multi-match workload at reduced scale (geometric duplicates,
Zipf(1.1) probe keys, 20% absent): the partitioned multi-match join's pairs
equal the oracle's sort/search pairs as a multiset, the pair count equals
the generator's exact count, and the semi/anti row masks equal the
oracle's kept rows."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from pageflow_amd.join import HashJoin
from pageflow_amd.operators.join_core import semi_anti_mask
from pageflow_amd.workloads import make_skewed_join

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nb,nd,np_,bits", [(40_000, 10_000, 2_000_000, None),
                                            (40_000, 10_000, 1_000_000, 4),
                                            (200_000, 50_000, 3_000_000, 12)])
def test_skewed_join_matches_oracle(nb, nd, np_, bits):
    dev = torch.device("cuda")
    build, probe, want = make_skewed_join(nb, nd, np_, 3, dev, chunk=1 << 20)
    plan = HashJoin(nb, np_, unique=False, bits=bits, device=dev)
    n = plan.prepare(build, probe)
    assert n == want["pairs"]
    out_p = torch.empty(n, dtype=torch.int64, device=dev)
    out_b = torch.empty(n, dtype=torch.int64, device=dev)
    plan.write(out_p, out_b)
    got = np.stack([out_b.cpu().numpy(), out_p.cpu().numpy()], axis=1)
    got = got[np.lexsort((got[:, 1], got[:, 0]))]
    bk, pk = build.cpu().numpy(), probe.cpu().numpy()
    assert np.array_equal(got, oracle.join_pairs_sorted(bk, pk))
    for anti in (False, True):
        mask = semi_anti_mask(build, probe, anti=anti)
        kept = torch.nonzero(mask).flatten().cpu().numpy()
        assert np.array_equal(kept, oracle.semi_anti(bk, pk, anti))
    assert int(semi_anti_mask(build, probe, anti=False).sum()) == want["present"]
