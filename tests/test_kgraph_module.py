"""`import kgraph` gives the reference's Python interface
(python-api.cpp:412-450): the class, its keyword arguments and defaults, and
the module attributes."""
import inspect


def test_module_attributes():
    import kgraph
    import kgraph_amd
    assert kgraph.KGraph is kgraph_amd.KGraph
    assert isinstance(kgraph.git_commit, str)
    assert set(kgraph.arch) == {"name", "alignment"}


def test_signatures_match_the_reference():
    import kgraph
    b = inspect.signature(kgraph.KGraph.build).parameters
    # python-api.cpp:429-438
    want_b = dict(iterations=30, L=100, K=25, S=10, controls=100, delta=0.002, recall=0.99, prune=0, reverse=0)
    assert list(b)[1:len(want_b) + 1] == list(want_b)
    for k, v in want_b.items():
        assert b[k].default == v, k
    s = inspect.signature(kgraph.KGraph.search).parameters
    # python-api.cpp:439-448
    want_s = dict(K=25, P=100, M=0, S=10, T=1, threads=0, withDistance=False, blas=False)
    assert list(s)[1] == "query" and list(s)[2:2 + len(want_s)] == list(want_s)
    for k, v in want_s.items():
        assert s[k].default == v, k
    for m in ("load", "save", "save_text"):
        assert callable(getattr(kgraph.KGraph, m))
