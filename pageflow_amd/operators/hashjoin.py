"""One join core call on the gfx950 library: the drop-in for
``join_kernel_new`` of the reference's operators/hashjoin.py (:444-494) and
the helpers its tests reach (``set_join_core_mode``, ``_join_core_fallback``,
``_residual_summary``, ``evaluate_join``, ``_sort_search_pairs``,
``join_call_args``).  Names, arguments, results and the guard names a
refused fast path reports are the reference's; the code is this package's.

A call joins one tick: two dicts of column id -> device tensor, equal on
``left_key`` = ``right_key``, filtered by the residual ``condition_list``,
and returns the ``join_output_keys`` columns gathered at the result rows
(then both global indexes when ``return_global_index``).  It takes one of
three routes:

``unique``   an inner join whose hashable equality keys are covered by a
             uniqueness proof of one side: that side is built into the
             unique-hash table (``join_core.unique_hash_pairs``), unless a
             guard refuses (``_join_core_fallback`` gets its name);
``set``      a semi or anti join on one int64 key a side whose residual, if
             any, is one comparison a per-key min/max summary answers
             (``join_core.semi_anti``);
``pairs``    everything else: every matching pair from the partitioned
             multi-match LDS-table join (``pageflow_amd.join.hash_join``),
             which gives the multiset of pairs the reference's sort/search
             core gives.  The residual's top-level AND operands that compile
             to the device program are evaluated inside the join's count and
             write passes (only kept pairs are made); the rest on the pairs.

Residuals on pairs run as one compiled device program
(``operators.residual``) where the conditions compile, else on the
reference's rules with torch on the GPU (``operators.pair_terms``).  The
join-type tails and the gathers (null rows for right-outer) are library
kernels.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Dict, List, Literal, Optional, Sequence, Tuple, Union

import torch

from .. import join as _join
from .. import ops
from ..layout import NULL
from . import join_core
from . import residual as _residual
from .pair_terms import PairTerms

LEFT, RIGHT = 0, 1

JOIN_TYPES = {
    'Inner': 'inner',
    'LeftSemi': 'right-semi',
    'LeftOuter': 'right-outer',
    'LeftAnti': 'right-anti',
}

# Pairs a residual evaluation takes (operators/operator_utils.py:19).
INT_ROWS_MAX = (1 << 31) - 2

# The unique-hash gate ("legacy" sends inner joins to the pair route too;
# tests switch it) and its two size guards, with the reference's values
# (hashjoin.py:36-55): a build at most 8x the probe rows, and a scratch
# estimate (join_core.unique_hash_peak_bytes) within 2 GiB unless the call
# passes its own limit.
_JOIN_CORE_MODES = {"legacy", "unique_hash"}
_join_core_mode = "unique_hash"
_UNIQUE_HASH_MAX_BUILD_PROBE_RATIO = 8
_UNIQUE_HASH_SCRATCH_LIMIT_BYTES = 2 * 1024 ** 3


def set_join_core_mode(mode: str) -> None:
    global _join_core_mode
    if mode not in _JOIN_CORE_MODES:
        raise ValueError(f"join core mode must be one of {sorted(_JOIN_CORE_MODES)}, "
                         f"got {mode!r}")
    _join_core_mode = mode


def _join_core_fallback(reason: str):
    """The unique-hash route was refused by the guard ``reason``; returns
    None so the caller takes another route.  Tests replace this function to
    observe the reasons."""
    return None


normalize_expr_tree = _residual.normalize_expr_tree
column_ref = _residual.column_ref


def _hashable(t: torch.Tensor) -> bool:
    """The hash kernels read 1-D contiguous int64 device vectors."""
    return t.is_cuda and t.dtype == torch.int64 and t.dim() == 1 and t.is_contiguous()


# ----------------------------------------------------- unique-hash route --

@dataclass
class _UniqueRoute:
    build_left: bool
    build_ids: List[int]
    probe_ids: List[int]
    hashed: frozenset


def _plan_unique_hash(tensors_left: Dict, tensors_right: Dict, left_key, right_key,
                      join_type: str, equality_keys, unique_key_sets,
                      scratch_limit_bytes) -> Union[_UniqueRoute, str]:
    """The unique-hash route of a call, or the name of the first guard that
    refuses it.  The guards run in the reference's order, since its tests
    observe which one refuses."""
    if _join_core_mode == "legacy":
        return "gate_legacy"
    if join_type != "inner":
        return "join_type"
    if equality_keys is None or unique_key_sets is None:
        return "missing_proof_metadata"
    left_eq, right_eq = (list(k) for k in equality_keys)
    if not left_eq or len(left_eq) != len(right_eq):
        return "unaligned_equality_keys"
    # the equality pairs the table can hash: both columns present, int64
    hashed_pairs = [(lk, rk) for lk, rk in zip(left_eq, right_eq)
                    if lk in tensors_left and rk in tensors_right
                    and _hashable(tensors_left[lk]) and _hashable(tensors_right[rk])]
    # rows of a streaming tick are co-located by the partition key only, so
    # it must be one of the hashed pairs
    if (left_key, right_key) not in hashed_pairs:
        return "primary_key_layout_or_dtype"
    if len(hashed_pairs) > 2:
        return "unsupported_key_count"
    left_ids = [lk for lk, _ in hashed_pairs]
    right_ids = [rk for _, rk in hashed_pairs]

    def proven(key_sets, ids) -> bool:
        return any(set(ks) <= set(ids) for ks in (key_sets or []))

    left_ok = proven(unique_key_sets[0], left_ids)
    right_ok = proven(unique_key_sets[1], right_ids)
    if not (left_ok or right_ok):
        return "selected_keys_not_proven_unique"
    n_left, n_right = tensors_left[left_ids[0]].numel(), tensors_right[right_ids[0]].numel()
    # with both sides proven, build the smaller; a tie builds the left
    build_left = left_ok and (not right_ok or n_left <= n_right)
    n_build, n_probe = (n_left, n_right) if build_left else (n_right, n_left)
    if n_probe and n_build > _UNIQUE_HASH_MAX_BUILD_PROBE_RATIO * n_probe:
        return "build_probe_ratio_guard"
    if scratch_limit_bytes is not None and \
            join_core.unique_hash_peak_bytes(n_build, n_probe) > scratch_limit_bytes:
        return "scratch_guard"
    if build_left:
        return _UniqueRoute(True, left_ids, right_ids, frozenset(hashed_pairs))
    return _UniqueRoute(False, right_ids, left_ids, frozenset(hashed_pairs))


def _try_unique_hash_pairs(tensors_left, tensors_right, left_key, right_key, join_type,
                           equality_keys, unique_key_sets, scratch_limit_bytes):
    """(left rows, right rows, hashed equality pairs) from the unique-hash
    table, or None when a guard refuses."""
    route = _plan_unique_hash(tensors_left, tensors_right, left_key, right_key, join_type,
                              equality_keys, unique_key_sets, scratch_limit_bytes)
    if isinstance(route, str):
        return _join_core_fallback(route)
    build_side, probe_side = ((tensors_left, tensors_right) if route.build_left
                              else (tensors_right, tensors_left))
    build_rows, probe_rows = join_core.unique_hash_pairs(
        [build_side[c] for c in route.build_ids], [probe_side[c] for c in route.probe_ids])
    if route.build_left:
        return build_rows, probe_rows, route.hashed
    return probe_rows, build_rows, route.hashed


# ------------------------------------------------------------- set route --
#
# A comparison is the set of orderings of (a, b) it accepts, among a < b,
# a = b and a > b.  Negation takes the complement; swapping the operands
# mirrors < and >.

_ORDERINGS = {'=': 'E', '<>': 'LG', '<': 'L', '<=': 'LE', '>': 'G', '>=': 'EG'}
_OP_OF = {frozenset(v): k for k, v in _ORDERINGS.items()}


def _negated(op: Optional[str]) -> Optional[str]:
    if op not in _ORDERINGS:
        return None
    return _OP_OF.get(frozenset('LEG') - frozenset(_ORDERINGS[op]))


def _swapped(op: Optional[str]) -> Optional[str]:
    if op not in _ORDERINGS:
        return None
    return _OP_OF[frozenset({'L': 'G', 'G': 'L', 'E': 'E'}[o] for o in _ORDERINGS[op])]


def _plain_column(token) -> Optional[int]:
    """The id of an operand that is a bare column reference, else None."""
    token = normalize_expr_tree(token)
    if isinstance(token, str) and ' ' not in token:
        return column_ref(token)
    return None


def _residual_summary(condition, tensors_left, tensors_right):
    """(build column, probe column, op) when ``condition``, under any number
    of NOTs, compares one left (build) int64 column with one right (probe)
    int64 column by an operator the per-key summary answers
    (``join_core.SUMMARY_OPS``), written build op probe; else None."""
    tree = normalize_expr_tree(condition)
    flips = 0
    while isinstance(tree, list) and len(tree) == 2 and tree[0] == 'NOT':
        flips += 1
        tree = normalize_expr_tree(tree[1])
    if not (isinstance(tree, list) and len(tree) == 3 and isinstance(tree[1], str)):
        return None
    a, op, b = _plain_column(tree[0]), tree[1], _plain_column(tree[2])
    if flips % 2:
        op = _negated(op)
    # an id both sides hold is read on the left by the pair evaluator, so
    # the summary could answer it on the wrong side
    if (a in tensors_left and a in tensors_right) or (b in tensors_left and b in tensors_right):
        return None
    if a in tensors_left and b in tensors_right:
        build_id, probe_id = a, b
    elif a in tensors_right and b in tensors_left:
        build_id, probe_id, op = b, a, _swapped(op)
    else:
        return None
    if op not in join_core.SUMMARY_OPS:
        return None
    build, probe = tensors_left[build_id], tensors_right[probe_id]
    if not (_hashable(build) and _hashable(probe)):
        return None
    return build, probe, op


def _try_hash_semi_anti(tensors_left, tensors_right, left_key, right_key, condition_list,
                        anti):
    """The kept right rows of a semi/anti join from the hash set, or None
    for the pair route."""
    build, probe = tensors_left[left_key], tensors_right[right_key]
    if not (_hashable(build) and _hashable(probe)) or len(condition_list or []) > 1:
        return None
    summary = None
    if condition_list:
        summary = _residual_summary(condition_list[0], tensors_left, tensors_right)
        if summary is None:
            return None
    return join_core.semi_anti(build, probe, anti, summary)


# ------------------------------------------------------------- residuals --

def _conjunction(conditions: Sequence):
    """One tree true where every condition is."""
    tree = conditions[0]
    for c in conditions[1:]:
        tree = [tree, 'AND', c]
    return tree


def _conjuncts(tree) -> list:
    """The top-level AND operands of a condition tree (the tree itself when
    it is no AND)."""
    tree = normalize_expr_tree(tree)
    if isinstance(tree, list) and len(tree) == 3 and tree[1] == 'AND':
        return _conjuncts(tree[0]) + _conjuncts(tree[2])
    return [tree]


def _compiles(tensorsa, tensorsb, condition) -> bool:
    try:
        _residual.compile_conditions(tensorsa, tensorsb, [condition])
    except NotImplementedError:
        return False
    return True


def evaluate_join(device_name, tensorsa, tensorsb, tree, leftidx, rightidx, select=False):
    """The condition ``tree`` on the pairs (leftidx[i], rightidx[i]): a
    column id ``tensorsa`` (the left side) holds is read on the left rows,
    any other on the right rows (the reference's hashjoin.py:235-242).
    Returns the bool mask over the pairs, or with ``select`` the kept pairs
    themselves (left rows, right rows) in pair order.

    The tree's top-level AND operands that the device program compiles run
    as one program (one pass, and with ``select`` a fused compaction); the
    others, forms the program lacks, are evaluated on the reference's rules
    with torch on the GPU (``pair_terms``) and ANDed in."""
    parts = _conjuncts(tree)
    native = [c for c in parts if _compiles(tensorsa, tensorsb, c)]
    other = [c for c in parts if not any(c is n for n in native)]
    prog = _residual.compile_conditions(tensorsa, tensorsb, native) if native else None
    if not other:
        if select:
            return _residual.filter_pairs(prog, leftidx, rightidx)
        return _residual.evaluate_pairs(prog, leftidx, rightidx)
    mask = PairTerms(tensorsa, tensorsb, leftidx, rightidx).mask(_conjunction(other))
    if prog is not None:
        mask &= _residual.evaluate_pairs(prog, leftidx, rightidx)
    if not select:
        return mask
    kept = torch.nonzero(mask).view(-1)
    return ops.gather_rows(leftidx, kept), ops.gather_rows(rightidx, kept)


def _filter_pairs(tensors_left, tensors_right, conditions, left_rows, right_rows,
                  device=None):
    """The pairs every condition keeps, in pair order: one evaluation of
    their conjunction."""
    if not conditions or left_rows.numel() == 0:
        return left_rows, right_rows
    device = left_rows.device if device is None else device
    return evaluate_join(device.type, tensors_left, tensors_right, _conjunction(conditions),
                         left_rows, right_rows, select=True)


def _is_hashed_equality(condition, hashed_pairs) -> bool:
    """True for exactly ``#a = #b`` where (a, b) is an equality the unique
    table hashed: the probe compared the whole tuple, so every pair already
    satisfies it."""
    if not (isinstance(condition, (list, tuple)) and len(condition) == 3
            and condition[1] == '='):
        return False
    ends = {column_ref(condition[0]) if isinstance(condition[0], str) else None,
            column_ref(condition[2]) if isinstance(condition[2], str) else None}
    return any(ends == {a, b} for a, b in hashed_pairs)


def _apply_fast_path_conditions(tensors_left, tensors_right, condition_list, left_matched,
                                right_matched, device, hashed_pairs):
    """The residual after the unique-hash route, without the equalities the
    table already hashed."""
    rest = [c for c in condition_list if not _is_hashed_equality(c, hashed_pairs)]
    return _filter_pairs(tensors_left, tensors_right, rest, left_matched, right_matched, device)


# ------------------------------------------------------------ pair route --

def _join_keys(left: torch.Tensor, right: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor]:
    """Both key columns as int64 join keys, equal where the keys compare
    equal: int64 as is, other integers widened, floats (or both sides when
    one is a float) by their float64 bit patterns with -0.0 folded and NaN
    unmatched (``pf_join_key_bits``)."""
    if _hashable(left) and _hashable(right):
        return left, right
    as_float = left.is_floating_point() or right.is_floating_point()
    return (ops.join_key_bits(left, 0, as_float), ops.join_key_bits(right, 1, as_float))


def _sort_search_pairs(tensors_left, tensors_right, left, right, condition_list,
                       bucket_counts, device, name="join"):
    """Every (left row, right row) with equal keys that the residual
    conditions keep.  Stands in for the reference's sort/search core
    (hashjoin.py:380-441) with the same pair multiset: the partitioned LDS
    join over all buckets in one pass (a key's rows share a bucket, so the
    tick's ``bucket_counts`` change nothing but are checked), then the
    residual program on the pairs."""
    n_left, n_right = left.size(0), right.size(0)
    if bucket_counts is not None:
        lc, rc = bucket_counts
        if len(lc) != len(rc) or sum(lc) != n_left or sum(rc) != n_right:
            raise AssertionError(f"bucket_counts misaligned with tick tensors: {len(lc)} and "
                                 f"{len(rc)} buckets of {sum(lc)} and {sum(rc)} rows for "
                                 f"{n_left} left and {n_right} right rows")
    if n_left == 0 or n_right == 0:
        none = torch.empty(0, dtype=torch.int64, device=device)
        return none, none.clone()
    lk, rk = _join_keys(left, right)
    lk, rk = lk.contiguous(), rk.contiguous()

    def over_limit(n_pairs: int) -> None:
        if condition_list and n_pairs > INT_ROWS_MAX:
            raise RuntimeError(f"{name}: a core call of {n_left} left and {n_right} right rows "
                               f"makes {n_pairs} pairs, over the {INT_ROWS_MAX} its residual "
                               f"conditions take")

    parts = _conjuncts(_conjunction(condition_list)) if condition_list else []
    native = [c for c in parts if _compiles(tensors_left, tensors_right, c)]
    if not native:
        right_rows, left_rows, _ = _join.hash_join(lk, rk, unique=False)
        over_limit(left_rows.numel())
        return _filter_pairs(tensors_left, tensors_right, condition_list, left_rows,
                             right_rows, device)
    # the compiled terms inside the join (left = build, right = probe)
    prog = _residual.compile_conditions(tensors_left, tensors_right, native)
    plan = _join.HashJoin(n_left, n_right, unique=False, device=lk.device)
    kept = plan.prepare(lk, rk, residual=prog)
    over_limit(plan.candidates)
    right_rows = torch.empty(kept, dtype=torch.int64, device=lk.device)
    left_rows = torch.empty(kept, dtype=torch.int64, device=lk.device)
    plan.write(right_rows, left_rows)
    other = [c for c in parts if not any(c is n for n in native)]
    return _filter_pairs(tensors_left, tensors_right, other, left_rows, right_rows, device)


# ----------------------------------------------------------------- tails --

def _gather_join_result(tensors_left, tensors_right, join_output_keys, return_global_index,
                        left_out, right_out, left_global_idx, right_global_idx):
    """The output columns at the result rows of their side, then the two
    global indexes; a null left row (-64) gathers a null row."""
    cols = [ops.gather_rows(tensors_left[k], left_out) if k in tensors_left
            else ops.gather_rows(tensors_right[k], right_out) for k in join_output_keys]
    if return_global_index:
        cols += [ops.gather_rows(left_global_idx, left_out),
                 ops.gather_rows(right_global_idx, right_out)]
    return cols


def _finish_join_result(tensors_left, tensors_right, join_output_keys, return_global_index,
                        join_type, left_matched, right_matched, left_global_idx,
                        right_global_idx, n_right, device):
    """Turns matched pairs into the join type's result rows and gathers
    them: inner keeps the pairs; right-semi the right rows with a match and
    right-anti those without (ascending, no left rows); right-outer the
    pairs, then the unmatched right rows with a null left row."""
    if join_type == 'inner':
        left_out, right_out = left_matched, right_matched
    elif join_type in ('right-semi', 'right-anti'):
        right_out = ops.matched_rows(right_matched, n_right, join_type == 'right-semi')
        left_out = torch.empty(0, dtype=torch.int64, device=device)
    elif join_type == 'right-outer':
        lonely = ops.matched_rows(right_matched, n_right, False)
        right_out = torch.cat([right_matched, lonely])
        left_out = torch.cat([left_matched, torch.full((lonely.numel(),), NULL,
                                                       dtype=torch.int64, device=device)])
    else:
        raise NotImplementedError("unhandled join type")
    return _gather_join_result(tensors_left, tensors_right, join_output_keys,
                               return_global_index, left_out, right_out, left_global_idx,
                               right_global_idx)


# ------------------------------------------------------------- the call --

def join_kernel_new(tensors_left, tensors_right, left_key, right_key, condition_list,
                    join_output_keys, return_global_index,
                    type: Literal['inner', 'right-semi', 'right-anti', 'right-outer'],
                    left_global_idx, right_global_idx, bucket_counts=None, equality_keys=None,
                    unique_key_sets=None, hash_scratch_limit_bytes=None, name="join"):
    """One core call: the unique route, else (semi/anti) the set route,
    else the pair route; then the join type's tail and the gather."""
    assert type in {'inner', 'right-semi', 'right-anti', 'right-outer'}, "Invalid Join type."
    left, right = tensors_left[left_key], tensors_right[right_key]
    device = left.device if left.is_cuda else right.device
    n_right = right.size(0)
    limit = (_UNIQUE_HASH_SCRATCH_LIMIT_BYTES if hash_scratch_limit_bytes is None
             else hash_scratch_limit_bytes)

    unique = _try_unique_hash_pairs(tensors_left, tensors_right, left_key, right_key, type,
                                    equality_keys, unique_key_sets, limit)
    if unique is not None:
        left_m, right_m, hashed = unique
        left_m, right_m = _apply_fast_path_conditions(tensors_left, tensors_right,
                                                      condition_list, left_m, right_m, device,
                                                      hashed)
    else:
        if type in ('right-semi', 'right-anti'):
            kept = _try_hash_semi_anti(tensors_left, tensors_right, left_key, right_key,
                                       condition_list, anti=type == 'right-anti')
            if kept is not None:
                return _gather_join_result(tensors_left, tensors_right, join_output_keys,
                                           return_global_index,
                                           torch.empty(0, dtype=torch.int64, device=device),
                                           kept, left_global_idx, right_global_idx)
        left_m, right_m = _sort_search_pairs(tensors_left, tensors_right, left, right,
                                             condition_list, bucket_counts, device, name)
    return _finish_join_result(tensors_left, tensors_right, join_output_keys,
                               return_global_index, type, left_m, right_m, left_global_idx,
                               right_global_idx, n_right, device)


def join_call_args(args, tensor_group):
    """(left key, right key, residual conditions, core join type) of a plan
    Join (the reference's hashjoin.py:506-528): the first key pair is the
    partition key unless its left column is a float, in which case the
    second is; the other pair's equality becomes a residual condition."""
    lks, rks = args['left keys'], args['right keys']
    conditions = [args['join condition']] if 'join condition' in args else []
    pick = 0
    if len(lks) == 2:
        first_is_float = tensor_group[lks[0]].is_floating_point()
        if first_is_float:
            assert not tensor_group[lks[1]].is_floating_point(), \
                "Do not support join on two float-type keys"
            pick = 1
        other = 1 - pick
        conditions.append([f'#{lks[other]}', '=', f'#{rks[other]}'])
    assert args['join type'] in JOIN_TYPES, f"Unsupport Join Type {args['join type']}"
    return lks[pick], rks[pick], conditions, JOIN_TYPES[args['join type']]
