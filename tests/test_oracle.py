"""Pins of the CPU restatement oracle (oracle/kgraph_oracle.cpp).

The reference ships no tests, golden vectors or fixtures for this path, so
each check comes from an independent source: numpy's MT19937 stream, a
differently-formulated model, FP64 arithmetic, closed forms, or byte layouts
written out by hand from kgraph.cpp:272-373.
"""
import os
import struct

import numpy as np
import pytest

from oracle import kgo

L2, EUC, ANG = 0, 1, 2


# ---------------------------------------------------------------- mt19937 / GenRandom
def test_mt19937_matches_numpy_stream():
    # std::mt19937 default seed 5489: the 10000th output is 4123659995 (C++11 26.5.5).
    out = kgo.mt19937(5489, 10000)
    assert out[-1] == 4123659995
    for seed in (1, 1998, 123456789):
        ref = np.random.RandomState(seed).randint(0, 2 ** 32, size=500, dtype=np.uint64).astype(np.uint32)
        assert np.array_equal(kgo.mt19937(seed, 500), ref)


def _genrandom_from_stream(raw, size, N):
    # kgraph.cpp:44-65 applied to an externally generated raw stream
    if N == size:
        return list(range(size)), 0
    a = sorted(int(raw[i]) % (N - size) for i in range(size))
    for i in range(1, size):
        if a[i] <= a[i - 1]:
            a[i] = a[i - 1] + 1
    off = int(raw[size]) % N
    return [(x + off) % N for x in a], size + 1


@pytest.mark.parametrize("size,N", [(1, 2), (5, 7), (10, 1000), (100, 101), (100, 100000), (7, 7)])
def test_genrandom(size, N):
    seed = 1998
    got = kgo.genrandom(seed, 3, size, N)
    raw = np.random.RandomState(seed).randint(0, 2 ** 32, size=4 * (size + 1), dtype=np.uint64)
    pos = 0
    for t in range(3):
        exp, used = _genrandom_from_stream(raw[pos:], size, N)
        pos += used
        assert list(got[t]) == exp
        assert len(set(got[t].tolist())) == size and got[t].max() < N


# ---------------------------------------------------------------- UpdateKnnList
def _model_insert(entries, K, nid, nd):
    """UpdateKnnList contract (kgraph.cpp:163-175) by a different route: a
    stable sort puts nn after every entry with dist <= nd."""
    lst = entries[:K]
    for (i, d) in lst:
        if d == nd and i == nid:
            return K + 1, None
    pos = sum(1 for (_, d) in lst if d <= nd)
    return pos, lst[:pos] + [(nid, nd)] + lst[pos:]


def test_update_knn_list_model():
    rs = np.random.RandomState(0)
    for trial in range(2000):
        K = rs.randint(0, 8)
        d = np.sort(rs.randint(0, 5, size=K).astype(np.float32))
        ids = rs.permutation(20)[:K].astype(np.uint32)
        nid = int(rs.randint(0, 20))
        nd = float(rs.randint(0, 6))
        r, oids, od = kgo.update_knn_list(np.append(ids, 0), np.append(d, 0), K, nid, nd)
        er, elist = _model_insert(list(zip(ids.tolist(), d.tolist())), K, nid, nd)
        assert r == er
        if elist is not None:
            assert list(oids[:K + 1]) == [e[0] for e in elist]
            assert list(od[:K + 1]) == [e[1] for e in elist]


# ---------------------------------------------------------------- distances
def test_distance_integer_data_exact():
    rs = np.random.RandomState(1)
    for D in (16, 96, 128, 37):
        a = rs.randint(0, 256, size=D).astype(np.float32)
        b = rs.randint(0, 256, size=D).astype(np.float32)
        exact = float(((a.astype(np.float64) - b) ** 2).sum())
        assert kgo.dist(L2, a, b) == exact
        # euclidean half form: n_a + n_b - dot, all terms exact multiples of 0.5 < 2^23
        assert kgo.dist(EUC, a, b) == exact / 2


@pytest.mark.parametrize("D", [1, 15, 16, 17, 96, 128, 200])
def test_distance_fp64_bounds_and_symmetry(D):
    rs = np.random.RandomState(D)
    eps = np.finfo(np.float32).eps / 2
    for _ in range(50):
        a = rs.standard_normal(D).astype(np.float32)
        b = rs.standard_normal(D).astype(np.float32)
        a64, b64 = a.astype(np.float64), b.astype(np.float64)
        g = (D + 3) * eps / (1 - (D + 3) * eps)  # subtraction + D fmas + tree
        ssd = kgo.dist(L2, a, b)
        e = ((a64 - b64) ** 2).sum()
        assert abs(ssd - e) <= 2 * g * e + 1e-30
        assert kgo.dist(L2, a, b) == kgo.dist(L2, b, a)
        assert kgo.dist(EUC, a, b) == kgo.dist(EUC, b, a)
        assert kgo.dist(ANG, a, b) == kgo.dist(ANG, b, a)
        # python-api.cpp:66-68 evaluated in FP64: -dot / (|a||b| + FLT_EPSILON)
        den = np.sqrt(a64 @ a64) * np.sqrt(b64 @ b64) + np.finfo(np.float32).eps
        ang = -(a64 @ b64) / den
        assert abs(kgo.dist(ANG, a, b) - ang) <= 8 * g * (np.abs(a64) @ np.abs(b64)) / den + 4 * eps * abs(ang)
        assert kgo.dist(L2, a, a) == 0.0


def test_row_norms_match_metric():
    rs = np.random.RandomState(3)
    X = rs.randint(0, 256, size=(50, 128)).astype(np.float32)
    n = kgo.row_norms(EUC, X)
    assert np.array_equal(n.astype(np.float64), (X.astype(np.float64) ** 2).sum(1) / 2)


# ---------------------------------------------------------------- brute force
def test_bruteforce_vs_fp64_sort():
    rs = np.random.RandomState(4)
    X = rs.randint(0, 256, size=(3000, 32)).astype(np.float32)
    Q = rs.randint(0, 256, size=(20, 32)).astype(np.float32)
    ids, d, c = kgo.bruteforce(X, Q, 15, L2)
    D64 = ((Q[:, None, :].astype(np.float64) - X[None, :, :]) ** 2).sum(-1)
    for q in range(20):
        order = np.lexsort((np.arange(3000), D64[q]))[:15]  # (dist, id)
        assert list(ids[q]) == list(order)
        assert np.array_equal(d[q].astype(np.float64), D64[q][order])
    assert (c == 15).all()
    # epsilon cut
    eps = float(np.sort(D64[0])[5])
    ids, d, c = kgo.bruteforce(X, Q[:1], 15, L2, epsilon=eps)
    assert c[0] == (D64[0] <= eps).sum() and (d[0][:c[0]] <= eps).all()


def test_linear_knn_excludes_self():
    rs = np.random.RandomState(5)
    X = rs.randint(0, 256, size=(500, 16)).astype(np.float32)
    rows = np.array([0, 7, 499])
    ids, d = kgo.linear_knn(X, rows, 10, L2)
    D64 = ((X[rows][:, None, :].astype(np.float64) - X[None]) ** 2).sum(-1)
    for r, row in enumerate(rows):
        D64[r, row] = np.inf
        assert list(ids[r]) == list(np.lexsort((np.arange(500), D64[r]))[:10])


# ---------------------------------------------------------------- search
def complete_graph(X, metric):
    """Every node lists every other node sorted by (dist, id); M = N-1."""
    N = len(X)
    ids, d = kgo.linear_knn(X, np.arange(N), N - 1, metric)
    off = np.arange(N + 1, dtype=np.uint64) * (N - 1)
    return off, ids.ravel(), d.ravel(), np.full(N, N - 1, dtype=np.uint32)


@pytest.mark.parametrize("metric", [L2, EUC, ANG])
def test_search_on_complete_graph_is_exact(metric):
    # On a complete graph with M = deg, the node that ends at the head of the
    # pool has every node as a neighbour, so every node gets visited: the
    # result is the exact K-NN and every non-seed node costs one distance.
    rs = np.random.RandomState(6)
    X = rs.randint(0, 64, size=(300, 16)).astype(np.float32)
    Q = rs.randint(0, 64, size=(10, 16)).astype(np.float32)
    idx = kgo.Index()
    idx.import_(*complete_graph(X, metric))
    K, P = 10, 20
    ids, d, c, comps, scans = idx.search(X, Q, metric, K=K, P=P)
    bids, bd, _ = kgo.bruteforce(X, Q, K, metric)
    assert np.array_equal(d, bd)
    assert (comps == len(X) - P).all()
    assert (c == K).all()


def test_search_without_edges_returns_best_seeds():
    rs = np.random.RandomState(7)
    X = rs.randint(0, 64, size=(200, 16)).astype(np.float32)
    Q = rs.randint(0, 64, size=(5, 16)).astype(np.float32)
    idx = kgo.Index()
    idx.import_(np.zeros(201, dtype=np.uint64), np.zeros(0, dtype=np.uint32), np.zeros(0, dtype=np.float32),
                np.zeros(200, dtype=np.uint32))
    P, K = 30, 5
    ids, d, c, comps, scans = idx.search(X, Q, L2, K=K, P=P)
    seeds = kgo.genrandom(1998, 1, P, 200)[0]
    for q in range(5):
        ds = np.array([kgo.dist(L2, X[s], Q[q]) for s in seeds])
        order = np.lexsort((seeds, ds))[:K]
        assert list(ids[q]) == list(seeds[order])
    assert (comps == 0).all() and (scans == 0).all()


def test_search_epsilon_and_bruteforce_path():
    rs = np.random.RandomState(8)
    X = rs.randint(0, 64, size=(100, 8)).astype(np.float32)
    Q = rs.randint(0, 64, size=(4, 8)).astype(np.float32)
    idx = kgo.Index()
    idx.import_(*complete_graph(X, L2))
    # P >= N: exact scan (kgraph.cpp:427-433)
    ids, d, c, comps, _ = idx.search(X, Q, L2, K=7, P=100)
    bids, bd, _ = kgo.bruteforce(X, Q, 7, L2)
    assert np.array_equal(ids, bids) and (comps == 100).all()
    eps = float(bd[0][3])
    ids, d, c, _, _ = idx.search(X, Q[:1], L2, K=7, P=20, epsilon=eps)
    assert c[0] == (bd[0] <= eps).sum()


def test_search_init_contract():
    rs = np.random.RandomState(9)
    X = rs.randint(0, 64, size=(100, 8)).astype(np.float32)
    idx = kgo.Index()
    idx.import_(*complete_graph(X, L2))
    with pytest.raises(RuntimeError):
        idx.search(X, X[:1], L2, K=5, P=10, init=5, init_ids=np.zeros((1, 5), np.uint32))  # init must be < K
    with pytest.raises(RuntimeError):
        idx.search(X, X[:1], L2, K=5, P=10, T=2, init=2, init_ids=np.zeros((1, 2), np.uint32))


# ---------------------------------------------------------------- reverse / format
def test_reverse_set_semantics():
    rs = np.random.RandomState(10)
    X = rs.randint(0, 64, size=(60, 8)).astype(np.float32)
    N = 60
    ids, d = kgo.linear_knn(X, np.arange(N), 12, L2)
    M = rs.randint(1, 13, size=N).astype(np.uint32)
    idx = kgo.Index()
    idx.import_(np.arange(N + 1, dtype=np.uint64) * 12, ids.ravel(), d.ravel(), M)
    idx.reverse(-1)
    off, rids, rd, rM = idx.export()
    for i in range(N):
        exp = {}
        for j in range(M[i]):
            exp[int(ids[i, j])] = float(d[i, j])
        for u in range(N):
            for j in range(M[u]):
                if ids[u, j] == i:
                    exp[u] = float(d[u, j])
        got = list(zip(rids[off[i]:off[i + 1]].tolist(), rd[off[i]:off[i + 1]].tolist()))
        assert got == sorted(exp.items(), key=lambda t: (t[1], t[0]))
        assert rM[i] == len(exp)


def _prune2_model(lists, M):
    """Prune level 2 stated without flags or reverse lists: entry x at position
    l of node i survives iff (1) no node T kept by i at an earlier level has
    already kept x, and (2) no node T < i kept by i at an earlier level keeps
    the same x at level l.  Nodes are visited level by level in id order."""
    N = len(lists)
    Meff = [min(int(M[i]), len(lists[i])) for i in range(N)]
    kept = [[] for _ in range(N)]
    for l in range(max(Meff) if N else 0):
        now = {}
        for i in range(N):
            if l >= Meff[i]:
                continue
            x = lists[i][l]
            pruned = any(x in kept[T] for T in kept[i])
            pruned = pruned or any(T < i and now.get(T, False) and lists[T][l] == x for T in kept[i])
            now[i] = not pruned
        for i, keep in now.items():
            if keep:
                kept[i].append(lists[i][l])
    return kept


def _prune2_oracle(lists, M, level=2):
    N = len(lists)
    off = np.cumsum([0] + [len(l) for l in lists]).astype(np.uint64)
    ids = np.array([j for l in lists for j in l], dtype=np.uint32)
    d = np.arange(len(ids), dtype=np.float32)  # distances ride along with their ids
    idx = kgo.Index()
    idx.import_(off, ids, d, np.asarray(M, dtype=np.uint32))
    idx.prune(level)
    o, pi, pd, pM = idx.export()
    out = [pi[o[i]:o[i + 1]].tolist() for i in range(N)]
    for i in range(N):  # each kept entry keeps its own distance
        src = {int(a): float(b) for a, b in zip(ids[off[i]:off[i + 1]], d[off[i]:off[i + 1]])}
        assert [src[j] for j in out[i]] == pd[o[i]:o[i + 1]].tolist()
        assert pM[i] == len(out[i])
    return out


def test_prune2_triangle_known_answer():
    # 0 -> [1, 2], 1 -> [2], 2 -> []: 0 keeps 1, 1 keeps 2, so 0 -> 2 is
    # reachable through 1 and is removed
    assert _prune2_oracle([[1, 2], [2], []], [2, 1, 0]) == [[1], [2], []]
    # with 1 -> 2 outside M[1] the edge 0 -> 2 stays
    assert _prune2_oracle([[1, 2], [2], []], [2, 0, 0]) == [[1, 2], [], []]
    # same level: 1, 2 and 3 keep 0 at level 0; at level 1 node 0 keeps 3, so
    # nodes 1 and 2 (> 0, and they kept 0 earlier) lose their own level-1 entry 3
    assert _prune2_oracle([[1, 3], [0, 3], [0, 3], [0]], [2, 2, 2, 1]) == [[1, 3], [0], [0], [0]]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_prune2_matches_flag_free_model(seed):
    rs = np.random.RandomState(40 + seed)
    N = 80
    X = rs.randint(0, 16, size=(N, 4)).astype(np.float32)  # many shared neighbours
    ids, d = kgo.linear_knn(X, np.arange(N), 15, L2)
    lists = [ids[i].tolist() for i in range(N)]
    M = rs.randint(0, 18, size=N)  # M beyond the list size is clamped
    got = _prune2_oracle(lists, M)
    assert got == _prune2_model(lists, M)
    assert sum(len(k) for k in got) < sum(min(m, 15) for m in M)  # something was pruned


def _fixture_bytes(graph, M, fmt):
    b = b"KNNGRAPH" + struct.pack("<III", 2, fmt, len(graph))
    for i, lst in enumerate(graph):
        b += struct.pack("<II", M[i], len(lst))
        for (nid, nd) in lst:
            if fmt & 1:
                b += struct.pack("<I", nid)
            else:
                b += struct.pack("<If", nid, nd) + b"\0" * 4
    return b


@pytest.mark.parametrize("fmt", [0, 1])
def test_save_load_byte_layout(tmp_path, fmt):
    graph = [[(1, 0.5), (2, 1.25)], [(0, 0.5)], [(0, 1.25), (1, 3.0), (3, 4.0)], []]
    M = [1, 1, 2, 0]
    off = np.cumsum([0] + [len(g) for g in graph]).astype(np.uint64)
    ids = np.array([e[0] for g in graph for e in g], dtype=np.uint32)
    d = np.array([e[1] for g in graph for e in g], dtype=np.float32)
    idx = kgo.Index()
    idx.import_(off, ids, d, np.array(M, dtype=np.uint32))
    p = str(tmp_path / "g.kgraph")
    idx.save(p, fmt)
    assert open(p, "rb").read() == _fixture_bytes(graph, M, fmt)
    idx2 = kgo.Index()
    idx2.load(p)
    o2, i2, d2, m2 = idx2.export()
    assert np.array_equal(o2, off) and np.array_equal(i2, ids) and np.array_equal(m2, M)
    if fmt == 0:
        assert np.array_equal(d2, d)


def test_load_rejects_bad_files(tmp_path):
    p = str(tmp_path / "bad")
    open(p, "wb").write(b"KNNGRAPH" + struct.pack("<III", 3, 0, 0))
    with pytest.raises(RuntimeError, match="version"):
        kgo.Index().load(p)
    open(p, "wb").write(b"KNNGRAPH" + struct.pack("<III", 2, 0, 5))
    with pytest.raises(RuntimeError):
        kgo.Index().load(p)


# ---------------------------------------------------------------- NN-descent
def test_nndescent_recall_and_invariants():
    rs = np.random.RandomState(11)
    X = rs.random_sample((4000, 8)).astype(np.float32)
    idx = kgo.Index()
    info = idx.build(X, L2, K=10, L=40, S=10, R=100, iterations=30, controls=50, reverse=0)
    assert info.recall > 0.9
    off, ids, d, M = idx.export()
    N = len(X)
    gt_ids, gt_d = kgo.linear_knn(X, np.arange(N), 10, L2)
    rec = kgo.recall_at_k([d[off[i]:off[i] + 10] for i in range(N)], gt_d, 10)
    assert rec > 0.9
    for i in range(0, N, 97):
        lst = ids[off[i]:off[i + 1]]
        dl = d[off[i]:off[i + 1]]
        assert i not in lst and len(set(lst.tolist())) == len(lst)
        assert (np.diff(dl) >= 0).all()
        assert 1 <= M[i] <= len(lst)
        for j, dd in zip(lst, dl):
            assert dd == kgo.dist(L2, X[i], X[j])


def test_nndescent_small_dataset_guard():
    X = np.random.RandomState(0).random_sample((20, 4)).astype(np.float32)
    with pytest.raises(RuntimeError, match="K larger"):
        kgo.Index().build(X, L2, K=25)
