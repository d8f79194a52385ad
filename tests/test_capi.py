"""The C ABI library loads and exports every symbol include/kgraph_amd.h
declares; the Python mirror keeps the reference's signatures and defaults
(python-api.cpp:424-449).  No GPU compute here."""
import ctypes
import inspect
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "kgraph_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(kg_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from kgraph_amd import _lib
    assert os.path.exists(_lib.LIB_PATH), "build the library first"
    lib = ctypes.CDLL(_lib.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s
    # and the ctypes table covers all of them
    assert set(syms) == set(_lib.SIGNATURES)


def test_version_and_defaults():
    import kgraph_amd
    from kgraph_amd import _lib
    lib = _lib.lib()
    assert b"gfx950" in lib.kg_version()
    p = _lib.IndexParams()
    lib.kg_index_params_default(p)
    # kgraph.h:14-37
    assert (p.iterations, p.L, p.K, p.S, p.R, p.controls, p.seed, p.prune, p.reverse) == (30, 100, 25, 10, 100, 100,
                                                                                        1998, 0, 0)
    assert abs(p.delta - 0.002) < 1e-9 and abs(p.recall - 0.99) < 1e-7
    s = _lib.SearchParams()
    lib.kg_search_params_default(s)
    assert (s.K, s.M, s.P, s.S, s.T, s.seed, s.init) == (25, 0, 100, 10, 1, 1998, 0)


def test_python_signatures_match_reference():
    import kgraph_amd
    b = inspect.signature(kgraph_amd.KGraph.build).parameters
    assert [b[k].default for k in ["iterations", "L", "K", "S", "controls", "delta", "recall", "prune", "reverse"]] == \
        [30, 100, 25, 10, 100, 0.002, 0.99, 0, 0]
    s = inspect.signature(kgraph_amd.KGraph.search).parameters
    assert [s[k].default for k in ["K", "P", "M", "S", "T", "threads", "withDistance", "blas"]] == \
        [25, 100, 0, 10, 1, 0, False, False]


def test_struct_layouts():
    from kgraph_amd import _lib
    assert ctypes.sizeof(_lib.IndexParams) == 44
    assert ctypes.sizeof(_lib.IndexInfo) == 28
    assert ctypes.sizeof(_lib.SearchParams) == 32
