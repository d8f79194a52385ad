"""BF16 screen of the graph walk's candidates (kg_set_search_screen): the
screened walk returns the oracle's ids, distances, counts and counters
bit-exactly (the screen only drops candidates that cannot enter the pool),
and the screen's error bound holds against the canonical FP32 distance on
rows chosen to stress it."""
import numpy as np
import pytest

from oracle import kgo

pytestmark = pytest.mark.gpu

L2, EUC, ANG = 0, 1, 2


def _index(kind, metric, d, n=6000, nq=200, seed=21):
    from kgraph_amd import datagen
    if kind == "uniform":
        X, Q = datagen.uniform(n, d, seed), datagen.uniform(nq, d, seed + 1)
    elif kind == "aniso":
        A = datagen.anisotropic_mixture(n + nq, d, 40, seed, seed + 1, seed + 2)
        X, Q = A[:n], A[n:]
    else:
        A = datagen.angular_mixture(n + nq, d, 50, seed, seed + 1)
        X, Q = A[:n], A[n:]
    idx = kgo.Index()
    idx.build(X, metric, K=20, L=50, S=10, R=100, iterations=12, controls=20, reverse=-1)
    return X, Q, idx


def _gpu(X, idx, metric):
    import kgraph_amd
    g = kgraph_amd.KGraph(X, "angular" if metric == ANG else "euclidean", l2sqr=(metric == L2))
    off, ids, dd, M = idx.export()
    g.set_graph(off, ids, dd, M)
    return g


@pytest.mark.parametrize("kind,metric,d,P,S,T,bitset", [
    ("uniform", EUC, 16, 100, 10, 1, False), ("uniform", L2, 37, 50, 10, 1, False),
    ("angular", ANG, 96, 100, 10, 1, False), ("angular", ANG, 96, 300, 64, 1, False),
    ("aniso", EUC, 100, 40, 30, 3, False), ("aniso", L2, 128, 200, 70, 1, False),
    ("angular", ANG, 96, 100, 10, 1, True), ("aniso", L2, 128, 60, 16, 2, True), ("aniso", EUC, 100, 40, 17, 1, True)])
def test_screened_walk_matches_oracle(kind, metric, d, P, S, T, bitset):
    # bitset: visited ids in the HBM bitset only, where batches of up to 16
    # ids test their bits and screen their rows at once
    X, Q, idx = _index(kind, metric, d)
    g = _gpu(X, idx, metric)
    g.set_search_screen(True)
    if bitset:
        g._h.call("kg_set_search_visited", 1)
    gi, gd = g.search(Q, K=10, P=P, S=S, T=T, withDistance=True, stats=True)
    rep = g.search_screen_report()
    assert rep["active"]
    ids, dd, c, comps, scans = idx.search(X, Q, metric, K=10, P=P, S=S, T=T)
    assert np.array_equal(g.last_counts, c)
    for q in range(len(Q)):
        assert np.array_equal(gi[q, :c[q]], ids[q, :c[q]]), q
        assert np.array_equal(gd[q, :c[q]], dd[q, :c[q]]), q
    assert np.array_equal(g.last_comps, comps) and np.array_equal(g.last_scans, scans)
    # the screen ruled out some candidates, and re-scored at most every one
    print("rescored %d of %d" % (rep["rescored"], comps.sum()))
    assert 0 < rep["rescored"] < comps.sum()


def _stress_rows(n, d, seed):
    rs = np.random.RandomState(seed)
    X = rs.standard_normal((n, d)) * np.exp(rs.uniform(-3, 3, (n, 1)))  # row scales over 2.6 decades
    X[: n // 4] += rs.uniform(-50, 50, (1, d))                           # a far offset shared by a quarter
    X[n // 4: n // 2] = np.round(X[n // 4: n // 2] * 300)                # large integers (not BF16-exact)
    X[n // 2: n // 2 + 10] *= 0                                          # zero rows
    return X.astype(np.float32)


@pytest.mark.parametrize("metric", [L2, EUC, ANG])
@pytest.mark.parametrize("d", [4, 20, 100, 256])
def test_screen_bound_holds(metric, d):
    import kgraph_amd
    X = _stress_rows(3000, d, 7 + d)
    Q = _stress_rows(40, d, 1000 + d)
    Q[:10] = X[:10] + np.float32(1e-3)  # queries next to data rows: tiny distances
    g = kgraph_amd.KGraph(X, "angular" if metric == ANG else "euclidean", l2sqr=(metric == L2))
    rs = np.random.RandomState(d)
    ids = rs.randint(0, len(X), (len(Q), 96)).astype(np.uint32)
    ids[:10, 0] = np.arange(10)
    ds, E = g.screen_dists(Q, ids)
    ex = np.array([[kgo.dist(metric, X[i], Q[q]) for i in ids[q]] for q in range(len(Q))], np.float64)
    err = np.abs(ds.astype(np.float64) - ex)
    assert np.isfinite(E).all() and (E >= 0).all()  # 0 only where both rows are zero
    assert (err <= E).all(), float((err / E).max())
    # the bound is not vacuous: for unit-scale cosine distances it is below 1 %
    if metric == ANG:
        assert E.max() < 0.01


def test_screen_off_where_rows_have_exact_copies():
    # integer rows in [0, 255]: walks gather the exact 8-bit copy, and the
    # screen stays off (same results either way)
    X, Q, idx = _index("uniform", L2, 16)
    Xi = np.round(X * 255).astype(np.float32)
    Qi = np.round(Q * 255).astype(np.float32)
    idx2 = kgo.Index()
    idx2.build(Xi, L2, K=20, L=50, S=10, R=100, iterations=12, controls=20, reverse=-1)
    g = _gpu(Xi, idx2, L2)
    a = g.search(Qi, K=10, P=50, withDistance=True)
    g.set_search_screen(True)
    b = g.search(Qi, K=10, P=50, withDistance=True)
    assert not g.search_screen_report()["active"]
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
