"""The C-ABI library: it loads, and it exports every entry point
include/pageflow_join.h declares (no compute calls: this runs without a
GPU)."""
from __future__ import annotations

import ctypes
import os
import re

from pageflow_amd import _lib

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))


def _declared():
    text = open(os.path.join(ROOT, "include", "pageflow_join.h")).read()
    return sorted(set(re.findall(r"\b(pf_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = _declared()
    assert names, "no entry point declared"
    for name in names:
        assert hasattr(lib, name), name
    assert lib.pf_version() == _lib.ABI_VERSION


def test_every_declared_symbol_has_a_binding():
    assert set(_declared()) == set(_lib.declared_symbols())


def test_serve_item_mirrors_the_header():
    from pageflow_amd import serving
    text = open(os.path.join(ROOT, "include", "pageflow_join.h")).read()
    assert int(re.search(r"#define PF_SERVE_MAX_ITEMS (\d+)", text).group(1)) == serving.MAX_ITEMS
    body = re.search(r"typedef struct pf_serve_item \{(.*?)\} pf_serve_item;", text, re.S).group(1)
    names = re.findall(r"(\w+);", body)
    assert names == [f for f, _ in serving.ServeItem._fields_]
    assert ctypes.sizeof(serving.ServeItem) == 72


def test_extension_module_has_the_reference_functions():
    import inspect
    from pageflow_amd import tqpkernels
    want = {"partition_sort": ["cols", "comp", "key_idx", "mask", "n_valid", "goff", "B",
                               "select_first", "host_out"],
            "join_expand": ["sorted_ids", "start", "len", "host_out"],
            "join_unique_hash": ["build_keys", "probe_keys", "host_out"],
            "join_semi_anti": ["build_key", "probe_key", "build_val", "probe_val", "op", "anti",
                               "host_out"]}
    for name, params in want.items():
        assert list(inspect.signature(getattr(tqpkernels, name)).parameters) == params
