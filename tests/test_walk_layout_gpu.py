"""Walk layout (kg_set_search_order): graph walks read copies of the rows,
norms and padded adjacency in a breadth-first node order of the graph.  The
order must be the documented one (checked against a numpy/scipy restatement
of its rule), and walks over the reordered copies must return what walks in
dataset order return — ids, distances, counts and per-query counters — and
what the oracle returns."""
import ctypes

import numpy as np
import pytest

from oracle import kgo

pytestmark = pytest.mark.gpu

L2, EUC, ANG = 0, 1, 2


def _dataset(kind, n, nq, d, seed):
    from kgraph_amd import datagen
    if kind == "uniform":
        return datagen.uniform(n, d, seed), datagen.uniform(nq, d, seed + 1)
    if kind == "intmix":
        X = datagen.integer_mixture(n + nq, d, 50, seed, seed + 1)
        return X[:n], X[n:]
    X = datagen.angular_mixture(n + nq, d, 50, seed, seed + 1)
    return X[:n], X[n:]


_CACHE = {}


def setup(kind, metric, n=6000, nq=200, d=16):
    key = (kind, metric, n, nq, d)
    if key not in _CACHE:
        X, Q = _dataset(kind, n, nq, d, 21)
        idx = kgo.Index()
        idx.build(X, metric, K=20, L=50, S=10, R=100, iterations=12, controls=20, reverse=-1)
        _CACHE[key] = (X, Q, idx)
    return _CACHE[key]


def gpu_index(X, graph, metric):
    import kgraph_amd
    g = kgraph_amd.KGraph(X, "angular" if metric == ANG else "euclidean", l2sqr=(metric == L2))
    g.set_graph(*graph)
    return g


def order_state(g, n=None):
    active, levels, ms = ctypes.c_int(-1), ctypes.c_uint(), ctypes.c_float()
    perm = np.zeros(n, np.uint32) if n else None
    g._h.call("kg_search_order", ctypes.byref(active), ctypes.byref(levels), ctypes.byref(ms),
              perm.ctypes.data if n else None)
    return active.value, levels.value, ms.value, perm


def bfs_layout(off, ids, M):
    """The documented order, restated: weak components of the edges u -> the
    first min(M[u], deg(u)) ids of u's list; a breadth-first tree from each
    component's smallest node, a level's new nodes in order of their first
    appearance in the concatenated lists of the previous level; trees in
    order of their root, a tree's levels in order; unreached nodes last, by id."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    N = len(M)
    off = off.astype(np.int64)
    cnt = np.minimum(np.diff(off), M.astype(np.int64))

    def lists(nodes):
        lens = cnt[nodes]
        st = np.repeat(off[nodes], lens)
        rel = np.arange(lens.sum()) - np.repeat(np.cumsum(lens) - lens, lens)
        return ids[st + rel].astype(np.int64), np.repeat(nodes, lens)

    dst, src = lists(np.arange(N))
    A = coo_matrix((np.ones(len(src)), (src, dst)), shape=(N, N))
    nc, lab = connected_components(A, directed=True, connection="weak")
    root_of = np.full(nc, N, np.int64)
    np.minimum.at(root_of, lab, np.arange(N))
    roots = np.sort(root_of)
    seen = np.zeros(N, bool)
    seen[roots] = True
    tree = np.full(N, -1, np.int64)
    tree[roots] = roots
    order = [roots]
    fr = roots
    while len(fr):
        nb, par = lists(fr)
        keep = ~seen[nb]
        nb, par = nb[keep], par[keep]
        _, first = np.unique(nb, return_index=True)
        first = np.sort(first)
        fr = nb[first]
        tree[fr] = tree[par[first]]
        seen[fr] = True
        order.append(fr)
    order = np.concatenate(order)
    order = order[np.argsort(tree[order], kind="stable")]
    return np.concatenate([order, np.flatnonzero(~seen)]).astype(np.uint32)


def walk(g, Q, **kw):
    ids, dd = g.search(Q, withDistance=True, stats=True, **kw)
    return ids, dd, g.last_counts.copy(), g.last_comps.copy(), g.last_scans.copy()


def check_oracle(r, X, Q, idx, metric, **kw):
    ids, dd, c, comps, scans = idx.search(X, Q, metric, **kw)
    assert np.array_equal(r[2], c)
    for q in range(len(Q)):
        assert np.array_equal(r[0][q, :c[q]], ids[q, :c[q]]), q
        assert np.array_equal(r[1][q, :c[q]], dd[q, :c[q]]), q
    assert np.array_equal(r[3], comps) and np.array_equal(r[4], scans)


@pytest.mark.parametrize("kind,metric,d", [("intmix", L2, 128), ("uniform", EUC, 37), ("angular", ANG, 96)])
def test_layout_is_the_breadth_first_order(kind, metric, d):
    X, Q, idx = setup(kind, metric, d=d)
    graph = idx.export()
    g = gpu_index(X, graph, metric)
    assert order_state(g)[0] == 0  # nothing walked yet
    walk(g, Q, K=10, P=40)
    active, levels, ms, perm = order_state(g, len(X))
    print("%s: %d levels, %.2f ms" % (kind, levels, ms))
    assert active == 1 and levels >= 2 and ms > 0
    assert np.array_equal(np.sort(perm), np.arange(len(X)))
    assert np.array_equal(perm, bfs_layout(graph[0], graph[1], graph[3]))


@pytest.mark.parametrize("kind,metric,d", [("intmix", L2, 128), ("uniform", EUC, 37), ("angular", ANG, 96)])
def test_layout_gives_identical_walks(kind, metric, d):
    # random starts, T > 1, M > 0, epsilon, entry points, a mixed K/P batch,
    # the LDS set and the bitset, every row form, the BF16 screen, two queries
    # per wave: dataset order and the graph order return the same
    X, Q, idx = setup(kind, metric, d=d)
    g = gpu_index(X, idx.export(), metric)
    g.set_entry_sample(512, seed=2)
    rs = np.random.RandomState(4)
    Ks = rs.choice([1, 10, 25], size=len(Q)).astype(np.uint32)
    Ps = rs.choice([10, 40, 100], size=len(Q)).astype(np.uint32)
    cases = [dict(K=10, P=40), dict(K=10, P=200, S=64), dict(K=10, P=30, T=3), dict(K=10, P=64, M=40),
             dict(K=25, P=50, entries=9), dict(K=10, Ks=Ks, Ps=Ps), dict(K=10, P=100, S=16),
             dict(K=10, P=50, epsilon=1e30)]
    settings = [("visited", 0), ("visited", 1), ("qpw", 2)]
    settings += [("rows", 1), ("rows", 2)] if kind == "intmix" else [("screen", 1)]
    for what, val in [(None, None)] + settings:
        if what == "visited":
            g._h.call("kg_set_search_visited", val)
        elif what == "rows":
            g._h.call("kg_set_row_storage", val)
        elif what == "screen":
            g._h.call("kg_set_search_screen", val)
        elif what == "qpw":
            g.set_search_queries_per_wave(val)
        for kw in cases:
            out = {}
            for order in (0, 1):
                g._h.call("kg_set_search_order", order)
                out[order] = walk(g, Q, **kw)
                # (a search M beyond the padded rows walks the CSR lists in dataset order)
                padded = ctypes.c_int(-1)
                g._h.call("kg_search_adjacency", ctypes.byref(padded), None)
                assert order_state(g)[0] == order * padded.value, (what, val, kw)
                assert padded.value == 1 or "M" in kw
            for a_, b_ in zip(out[0], out[1]):
                assert np.array_equal(a_, b_), (what, val, kw)
        g._h.call("kg_set_search_visited", 0)
        g._h.call("kg_set_row_storage", 0)
        g._h.call("kg_set_search_screen", 0)
        g.set_search_queries_per_wave(1)
    g._h.call("kg_set_search_order", 1)
    check_oracle(walk(g, Q, K=10, P=100), X, Q, idx, metric, K=10, P=100)


def test_layout_with_components_and_unreached_nodes():
    # a hand-made graph: three components, 60 nodes nobody links to (unreached
    # by the trees, placed last), lists repeating an id, an empty list; the
    # order is the documented one and walks match dataset order and the oracle
    rs = np.random.RandomState(7)
    N, d = 3000, 24
    X = rs.rand(N, d).astype(np.float32)
    Q = rs.rand(150, d).astype(np.float32)
    comp = np.repeat(np.arange(3), [1500, 900, 600])
    starts = np.array([0, 1500, 2400])
    sizes = np.array([1500, 900, 600])
    unreached = set(rs.choice(np.arange(1, 1500), 60, replace=False).tolist())
    lists = []
    for u in range(N):
        c = comp[u]
        pool = np.arange(starts[c], starts[c] + sizes[c])
        pool = pool[[p not in unreached for p in pool]]
        k = 0 if u == 2999 else int(rs.randint(4, 16))
        nb = rs.choice(pool[pool != u], k, replace=False) if k else np.zeros(0, np.int64)
        if u % 97 == 0 and k > 2:
            nb[2] = nb[0]  # a repeated id
        dist = ((X[nb] - X[u]) ** 2).sum(1).astype(np.float32)
        o = np.argsort(dist, kind="stable")
        lists.append((nb[o].astype(np.uint32), dist[o]))
    off = np.zeros(N + 1, np.uint64)
    off[1:] = np.cumsum([len(l[0]) for l in lists])
    ids = np.concatenate([l[0] for l in lists]).astype(np.uint32)
    dists = np.concatenate([l[1] for l in lists]).astype(np.float32)
    M = np.array([max(1, len(l[0]) - int(rs.randint(0, 3))) for l in lists], np.uint32)
    graph = (off, ids, dists, M)
    g = gpu_index(X, graph, L2)
    out = {}
    for order in (0, 1):
        g._h.call("kg_set_search_order", order)
        out[order] = walk(g, Q, K=10, P=30)
    for a_, b_ in zip(out[0], out[1]):
        assert np.array_equal(a_, b_)
    active, levels, ms, perm = order_state(g, N)
    want = bfs_layout(off, ids, M)
    assert np.array_equal(perm, want)
    pos = np.empty(N, np.int64)
    pos[perm] = np.arange(N)
    assert pos[sorted(unreached)].min() > pos[np.setdiff1d(np.arange(N), sorted(unreached))].max() - 5
    idx = kgo.Index()
    idx.import_(off, ids, dists, M)
    check_oracle(out[1], X, Q, idx, L2, K=10, P=30)


def test_layout_follows_graph_and_data_changes():
    # the layout and its row copies are remade after the graph or the rows
    # change: walks after set_graph (another graph), prune and new rows of the
    # same shape return the oracle's results for the new state
    from kgraph_amd import _lib
    X, Q, idx = setup("uniform", EUC, d=37)
    g = gpu_index(X, idx.export(), EUC)
    walk(g, Q, K=10, P=40)
    assert order_state(g)[0] == 1
    # another graph
    idx2 = kgo.Index()
    idx2.build(X, EUC, K=12, L=30, S=10, R=100, iterations=6, controls=20, reverse=0, seed=7)
    g.set_graph(*idx2.export())
    r = walk(g, Q, K=10, P=40)
    assert order_state(g)[0] == 1
    check_oracle(r, X, Q, idx2, EUC, K=10, P=40)
    # pruned on the device (level 1 keeps the first min(M, K) edges of each list)
    g._h.call("kg_prune", 1)
    r = walk(g, Q, K=10, P=40)
    idx3 = kgo.Index()
    idx3.import_(*g.graph())
    check_oracle(r, X, Q, idx3, EUC, K=10, P=40)
    # new rows of the same shape under the same graph
    X2 = np.ascontiguousarray(X[::-1])
    g._h.call("kg_set_data", X2.ctypes.data, X2.shape[0], X2.shape[1], X2.shape[1], EUC)
    r = walk(g, Q, K=10, P=40)
    check_oracle(r, X2, Q, idx3, EUC, K=10, P=40)


@pytest.mark.parametrize("shape", ["no-edges", "two-nodes", "path", "star"])
def test_layout_on_small_and_degenerate_graphs(shape):
    # graphs with no edges (every node its own component), two nodes, a long
    # path (one level per node) and a star: the order is the documented one
    # and walks match dataset order and the oracle
    rs = np.random.RandomState(3)
    N = {"no-edges": 50, "two-nodes": 2, "path": 600, "star": 300}[shape]
    X = rs.rand(N, 8).astype(np.float32)
    Q = rs.rand(20, 8).astype(np.float32)
    if shape == "no-edges":
        lists = [np.zeros(0, np.int64) for _ in range(N)]
    elif shape == "two-nodes":
        lists = [np.array([1]), np.array([0])]
    elif shape == "path":  # i -> i + 1 (and the last node -> 0): N levels from node 0
        lists = [np.array([(i + 1) % N]) for i in range(N)]
    else:  # every node -> the hub 0, the hub -> everyone
        lists = [np.arange(1, N)] + [np.array([0]) for _ in range(1, N)]
    off = np.zeros(N + 1, np.uint64)
    off[1:] = np.cumsum([len(l) for l in lists])
    ids = np.concatenate(lists).astype(np.uint32) if off[-1] else np.zeros(0, np.uint32)
    dists = np.concatenate([((X[l] - X[i]) ** 2).sum(1) for i, l in enumerate(lists)]).astype(np.float32) \
        if off[-1] else np.zeros(0, np.float32)
    M = np.array([max(1, len(l)) for l in lists], np.uint32)
    g = gpu_index(X, (off, ids, dists, M), L2)
    P = 1 if N == 2 else 5
    out = {}
    for order in (0, 1):
        g._h.call("kg_set_search_order", order)
        out[order] = walk(g, Q, K=1, P=P, S=10)
    for a_, b_ in zip(out[0], out[1]):
        assert np.array_equal(a_, b_)
    active, levels, ms, perm = order_state(g, N)
    assert np.array_equal(perm, bfs_layout(off, ids, M)), shape
    idx = kgo.Index()
    idx.import_(off, ids, dists, M)
    check_oracle(out[1], X, Q, idx, L2, K=1, P=P, S=10)


@pytest.mark.parametrize("metric", [L2, EUC, ANG])
def test_lean_walk_gives_identical_walks(metric):
    # the lean walk (integer queries over the 8-bit rows from selected entry
    # points, one trial; 5 waves per SIMD with its own visited-set size) and
    # the general walk return the same; a batch with one non-integer query
    # is walked by the general launch of the pair
    X, Q, idx = setup("intmix", metric, d=128)
    g = gpu_index(X, idx.export(), metric)
    g.set_entry_sample(512, seed=2)
    Qf = Q.copy()
    Qf[7, 3] += 0.5  # not an integer
    rs = np.random.RandomState(4)
    Ks = rs.choice([1, 10, 25], size=len(Q)).astype(np.uint32)
    Ps = rs.choice([10, 40, 100], size=len(Q)).astype(np.uint32)
    cases = [dict(K=10, P=50, S=128, entries=9), dict(K=25, P=20, S=10, entries=9),
             dict(K=10, P=200, S=64, entries=9), dict(K=10, Ks=Ks, Ps=Ps, entries=9)]
    for qs in (Q, Qf):
        for kw in cases:
            out = {}
            for lean in (0, 1):
                g._h.call("kg_set_search_lean", lean)
                out[lean] = walk(g, qs, **kw)
                vc, sp, bits = ctypes.c_uint(), ctypes.c_uint(), ctypes.c_uint()
                g._h.call("kg_search_visited_stats", ctypes.byref(vc), ctypes.byref(sp))
                g._h.call("kg_search_visited_bits", ctypes.byref(bits))
                print(metric, kw.get("P"), "lean", lean, "non-integer" if qs is Qf else "integer",
                      "set", vc.value, bits.value, "spilled", sp.value)
            for a_, b_ in zip(out[0], out[1]):
                assert np.array_equal(a_, b_), (kw, qs is Qf)
    g._h.call("kg_set_search_lean", 1)
    ids = g.select_entries(Q, 9)
    r = walk(g, Q, K=10, P=50, S=128, entries=9)
    check_oracle(r, X, Q, idx, metric, K=10, P=50, S=128, init=9, init_ids=ids)
