"""The allocation-free unique-hash entry point, ``pf_join_unique_hash_noalloc``,
called through the C ABI: its pairs must equal ``pf_join_unique_hash``'s
(the workspace entry point, through ``join_core.unique_hash_pairs``) bit for
bit, on consecutive calls that share the fixed device scratch, the second
with a smaller input; and a call too large for the scratch must fail before
it launches anything."""
import ctypes

import pytest
import torch

from pageflow_amd import _lib
from pageflow_amd.operators import join_core

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 64 << 20      # PF_NOALLOC_SCRATCH_BYTES


def _keys(nb, np_, seed, absent_every=7, two=False):
    g = torch.Generator().manual_seed(seed)
    bk = torch.randperm(nb, generator=g) * 3 + 11
    pk = bk[torch.randint(0, nb, (np_,), generator=g)]
    pk[::absent_every] = -1 - torch.arange(pk[::absent_every].numel())
    build, probe = [bk], [pk]
    if two:
        b2 = torch.randint(0, 4, (nb,), generator=g)
        p2 = torch.randint(0, 4, (np_,), generator=g)
        build.append(b2)
        probe.append(p2)
    return [t.cuda() for t in build], [t.cuda() for t in probe]


def _noalloc(build, probe):
    lib = _lib.load()
    nb, np_ = build[0].numel(), probe[0].numel()
    nkeys = len(build)
    need = lib.pf_join_unique_hash_workspace_bytes(nb, np_, nkeys)
    assert need <= SCRATCH_BYTES, (nb, np_, need)
    dev = build[0].device
    bidx = torch.empty(np_, dtype=torch.int64, device=dev)
    pidx = torch.empty(np_, dtype=torch.int64, device=dev)
    host = torch.zeros(2, dtype=torch.int64, pin_memory=True)
    b1 = build[1] if nkeys == 2 else build[0]
    p1 = probe[1] if nkeys == 2 else probe[0]
    print(f"noalloc call: nb={nb} np={np_} nkeys={nkeys} scratch={need}", flush=True)
    with torch.cuda.device(dev):
        _lib.check(lib.pf_join_unique_hash_noalloc(
            _lib.ptr(build[0]), _lib.ptr(b1), nb, _lib.ptr(probe[0]), _lib.ptr(p1), np_, nkeys,
            _lib.ptr(bidx), _lib.ptr(pidx), ctypes.c_void_p(host.data_ptr()),
            _lib.stream_of(dev)))
    n = int(host[0])
    return bidx[:n], pidx[:n]


@pytest.mark.parametrize("two", [False, True])
def test_noalloc_matches_workspace_entry_on_consecutive_calls(two):
    # first call larger, second smaller: the second must not see the first's
    # table, partner or counts
    for nb, np_, seed in ((600_000, 3_000_000, 1), (50_000, 200_003, 2)):
        build, probe = _keys(nb, np_, seed, two=two)
        got_b, got_p = _noalloc(build, probe)
        want_b, want_p = join_core.unique_hash_pairs(build, probe)
        assert got_b.numel() == want_b.numel() > 0
        assert torch.equal(got_b, want_b) and torch.equal(got_p, want_p)


def test_noalloc_rejects_inputs_over_its_scratch():
    lib = _lib.load()
    nb, np_ = 4_000_000, 16_000_000
    assert lib.pf_join_unique_hash_workspace_bytes(nb, np_, 1) > SCRATCH_BYTES
    # the size check runs on the host before any pointer is touched, so
    # placeholder addresses never reach a kernel
    dev = torch.device("cuda:0")
    host = torch.zeros(2, dtype=torch.int64, pin_memory=True)
    with torch.cuda.device(dev):
        rc = lib.pf_join_unique_hash_noalloc(None, None, nb, None, None, np_, 1, None, None,
                                             ctypes.c_void_p(host.data_ptr()),
                                             _lib.stream_of(dev))
    assert rc != 0
    assert b"over the fixed" in lib.pf_last_error()
