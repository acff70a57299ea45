"""The reference's join tensor program restated with torch on the CPU.

TEST INFRASTRUCTURE ONLY: used by ``tests/`` as a checker and by
``bench.py``'s cpu_baseline leg as the timed CPU baseline.  The product
package never imports it.

``sort_search_pairs`` follows the reference's sort/search core
(operators/hashjoin.py:380-441) for one call: per bucket, ``torch.sort`` of
the left keys, two ``torch.searchsorted`` for each right row's run of equal
keys, then the expansion of every run into (left row, right row) pairs in
right-row order, which the reference does with its join_expand kernel
(join_core.cu:263-295) and is written here with ``repeat_interleave`` (the
reference test's own expand reference, tests/operators/test_join_core.py:
55-61).  It is the path the reference takes for int32 keys (its hash cores
take int64 only, operators/hashjoin.py:80-86), so it is the reference's CPU
program for the plain int32 join.

Parity: pinned in tests/test_oracle.py against ``oracle.sort_search_pairs``
(the numpy restatement) and a dense brute force, as pair multisets.
"""
from __future__ import annotations

from typing import Optional, Sequence, Tuple

import torch


def sort_search_pairs(left: torch.Tensor, right: torch.Tensor,
                      bucket_counts: Optional[Tuple[Sequence[int], Sequence[int]]] = None
                      ) -> Tuple[torch.Tensor, torch.Tensor]:
    """(left rows, right rows) of every pair with equal keys, in right-row
    order and, within a right row, sorted-key order of the left rows."""
    n_left, n_right = left.numel(), right.numel()
    lc, rc = bucket_counts if bucket_counts is not None else ([n_left], [n_right])
    sorted_ids = torch.empty(n_left, dtype=torch.int64)
    start = torch.zeros(n_right, dtype=torch.int64)
    length = torch.zeros(n_right, dtype=torch.int64)
    lo_l = lo_r = 0
    for nl, nr in zip(lc, rc):
        hi_l, hi_r = lo_l + nl, lo_r + nr
        if nl:
            keys, perm = torch.sort(left[lo_l:hi_l])
            sorted_ids[lo_l:hi_l] = perm + lo_l
            if nr:
                probe = right[lo_r:hi_r]
                first = torch.searchsorted(keys, probe)
                length[lo_r:hi_r] = torch.searchsorted(keys, probe, right=True) - first
                start[lo_r:hi_r] = first + lo_l
        lo_l, lo_r = hi_l, hi_r
    total = int(length.sum())
    right_rows = torch.repeat_interleave(torch.arange(n_right, dtype=torch.int64), length)
    run_first = torch.repeat_interleave(start, length)
    run_base = torch.repeat_interleave(torch.cumsum(length, 0) - length, length)
    within = torch.arange(total, dtype=torch.int64) - run_base
    return sorted_ids[run_first + within], right_rows
