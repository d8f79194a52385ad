"""GPU search parity: the HIP graph walk (through the C ABI) returns the
oracle's ids, distances, result counts and per-query counters bit-exactly on
the same index, the same queries and the same entry points."""
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


def setup(kind, metric, n=6000, nq=200, d=16, build=None):
    key = (kind, metric, n, nq, d)
    if key not in _CACHE:
        X, Q = _dataset(kind, n, nq, d, 11)
        idx = kgo.Index()
        kw = dict(K=20, L=50, S=10, R=100, iterations=12, controls=20, reverse=-1)
        kw.update(build or {})
        idx.build(X, metric, **kw)
        _CACHE[key] = (X, Q, idx)
    return _CACHE[key]


def gpu_index(X, idx, metric):
    import kgraph_amd
    g = kgraph_amd.KGraph(X, "angular" if metric == ANG else "euclidean", l2sqr=(metric == L2))
    off, ids, d, M = idx.export()
    g.set_graph(off, ids, d, M)
    return g


def compare(X, Q, idx, metric, **kw):
    g = gpu_index(X, idx, metric)
    K = kw.get("K", 25)
    ids, d, c, comps, scans = idx.search(X, Q, metric, **{k: v for k, v in kw.items()})
    gkw = {k: v for k, v in kw.items() if k not in ("init", "init_ids")}
    gi, gd = g.search(Q, withDistance=True, stats=True, **gkw)
    gc, gcomp, gscan = g.last_counts, g.last_comps, g.last_scans
    assert np.array_equal(gc, c)
    for q in range(len(Q)):
        n = c[q]
        assert np.array_equal(gi[q, :n], ids[q, :n]), q
        assert np.array_equal(gd[q, :n], d[q, :n]), q
    assert np.array_equal(gcomp, comps)
    assert np.array_equal(gscan, scans)
    return gi, gd, c


@pytest.mark.parametrize("kind,metric,d", [("uniform", EUC, 16), ("uniform", L2, 16), ("intmix", L2, 128),
                                           ("angular", ANG, 96), ("uniform", EUC, 37)])
def test_search_parity_default_params(kind, metric, d):
    X, Q, idx = setup(kind, metric, d=d)
    compare(X, Q, idx, metric, K=10, P=100)


@pytest.mark.parametrize("P,S,M", [(10, 10, 0), (50, 1, 0), (200, 70, 0), (64, 10, 40), (300, 10, 0)])
def test_search_parity_params(P, S, M):
    X, Q, idx = setup("intmix", L2, d=128)
    compare(X, Q, idx, L2, K=10, P=P, S=S, M=M)


def test_search_parity_trials():
    X, Q, idx = setup("uniform", EUC)
    compare(X, Q, idx, EUC, K=10, P=30, T=3)


def test_search_parity_epsilon():
    X, Q, idx = setup("uniform", L2)
    bids, bd, _ = kgo.bruteforce(X, Q, 10, L2)
    eps = float(np.median(bd[:, 4]))
    gi, gd, c = compare(X, Q, idx, L2, K=10, P=50, epsilon=eps)
    assert (c < 10).any() and (gd[np.arange(10)[None, :] < c[:, None]] <= eps).all()


@pytest.mark.parametrize("kind,metric,d,lossless", [("intmix", L2, 37, True), ("intmix", ANG, 48, True),
                                                     ("intmix", EUC, 128, True), ("uniform", L2, 16, False)])
def test_row_storage_forms_give_identical_walks(kind, metric, d, lossless):
    # integer rows in [0, 255] have an 8-bit copy and a lossless BF16 copy:
    # walks gather the 8-bit one by default (mode 2: the BF16 one) and must
    # return the oracle's ids, distances and counters; uniform floats have
    # neither, and every mode gathers the FP32 rows
    import ctypes
    X, Q, idx = setup(kind, metric, d=d)
    g = gpu_index(X, idx, metric)
    active, exact = ctypes.c_int(-1), ctypes.c_int(-1)
    g._h.call("kg_row_storage", ctypes.byref(active), ctypes.byref(exact))
    assert active.value == (2 if lossless else 0) and exact.value == (3 if lossless else 0)
    res = {}
    # mode 3: the 8-bit copy with FP32 arithmetic (mode 0 takes the exact
    # integer path over it when the queries are integers in [0, 255])
    for mode, want in ((0, 2), (1, 0), (2, 1), (3, 2)):
        g._h.call("kg_set_row_storage", mode)
        g._h.call("kg_row_storage", ctypes.byref(active), None)
        assert active.value == (want if lossless else 0)
        gi, gd = g.search(Q, K=10, P=40, withDistance=True, stats=True)
        res[mode] = (gi, gd, g.last_counts.copy(), g.last_comps.copy(), g.last_scans.copy())
    g._h.call("kg_set_row_storage", 0)
    for mode in (1, 2, 3):
        for a, b in zip(res[0], res[mode]):
            assert np.array_equal(a, b)
    ids, dd, c, comps, scans = idx.search(X, Q, metric, K=10, P=40)
    assert np.array_equal(res[0][2], c)
    for q in range(len(Q)):
        assert np.array_equal(res[0][0][q, :c[q]], ids[q, :c[q]]), q
        assert np.array_equal(res[0][1][q, :c[q]], dd[q, :c[q]]), q
    assert np.array_equal(res[0][3], comps) and np.array_equal(res[0][4], scans)


@pytest.mark.parametrize("P,S,T", [(40, 10, 1), (300, 10, 1), (20, 64, 3), (500, 30, 1)])
def test_lds_visited_set_gives_identical_walks(P, S, T):
    # the LDS visited set (16-bit entries here, N = 6000, or 32-bit; spilling
    # into the HBM bitset once half full) and the bitset alone give the
    # oracle's ids, distances and counters; a 256-entry set makes queries
    # spill mid-walk, T > 1 reads the set across trials
    import ctypes
    X, Q, idx = setup("intmix", L2, d=128)
    g = gpu_index(X, idx, L2)
    res = {}
    for mode, size in [(0, 0), (2, 0), (1, 0), (0, 256), (2, 256)]:
        g._h.call("kg_set_search_visited", mode)
        g._h.call("kg_set_search_visited_size", size)
        gi, gd = g.search(Q, K=10, P=P, S=S, T=T, withDistance=True, stats=True)
        vc, sp, bits = ctypes.c_uint(), ctypes.c_uint(), ctypes.c_uint()
        g._h.call("kg_search_visited_stats", ctypes.byref(vc), ctypes.byref(sp))
        g._h.call("kg_search_visited_bits", ctypes.byref(bits))
        res[mode, size] = (gi, gd, g.last_counts.copy(), g.last_comps.copy(), g.last_scans.copy(), vc.value,
                           sp.value, bits.value)
        print("P=%d mode %d size %d: LDS set %d entries of %d bits, %d of %d queries spilled" % (
            P, mode, size, vc.value, bits.value, sp.value, len(Q)))
    g._h.call("kg_set_search_visited", 0)
    g._h.call("kg_set_search_visited_size", 0)
    assert res[1, 0][5] == 0 and res[1, 0][6] == 0
    if P <= 300:
        assert res[0, 0][5] >= 256  # a P = 500 pool leaves no LDS room for a set: bitset only
        assert res[0, 0][7] == 16 and res[2, 0][7] == 32
    assert res[0, 256][5] == 256 and res[0, 256][7] == 16 and res[0, 256][6] > 0
    assert res[2, 256][5] == 256 and res[2, 256][7] == 32 and res[2, 256][6] > 0
    for key in res:
        for a, b in zip(res[0, 0][:5], res[key][:5]):
            assert np.array_equal(a, b), key
    ids, dd, c, comps, scans = idx.search(X, Q, L2, K=10, P=P, S=S, T=T)
    r = res[0, 0]
    assert np.array_equal(r[2], c)
    for q in range(len(Q)):
        assert np.array_equal(r[0][q, :c[q]], ids[q, :c[q]]), q
        assert np.array_equal(r[1][q, :c[q]], dd[q, :c[q]]), q
    assert np.array_equal(r[3], comps) and np.array_equal(r[4], scans)


@pytest.mark.parametrize("P,S,T", [(40, 10, 1), (300, 10, 1), (20, 64, 3)])
def test_visited_sets_of_three_times_a_power_of_two(P, S, T):
    # sets of 3 * 2^k entries (home slot by multiply-shift, probes wrap at the
    # size; 16-bit entries keep a remainder of S - 1 bits): 384 entries make
    # queries spill mid-walk, 3072 hold them; 16- and 32-bit entries; the
    # same walks as the bitset alone and the oracle
    import ctypes
    X, Q, idx = setup("intmix", L2, d=128)
    g = gpu_index(X, idx, L2)
    res = {}
    for mode, size in [(1, 0), (0, 384), (0, 3072), (2, 384), (2, 3072)]:
        g._h.call("kg_set_search_visited", mode)
        g._h.call("kg_set_search_visited_size", size)
        gi, gd = g.search(Q, K=10, P=P, S=S, T=T, withDistance=True, stats=True)
        vc, sp, bits = ctypes.c_uint(), ctypes.c_uint(), ctypes.c_uint()
        g._h.call("kg_search_visited_stats", ctypes.byref(vc), ctypes.byref(sp))
        g._h.call("kg_search_visited_bits", ctypes.byref(bits))
        res[mode, size] = (gi, gd, g.last_counts.copy(), g.last_comps.copy(), g.last_scans.copy(), vc.value,
                           sp.value, bits.value)
        print("P=%d mode %d size %d: LDS set %d entries of %d bits, %d of %d queries spilled" % (
            P, mode, size, vc.value, bits.value, sp.value, len(Q)))
    g._h.call("kg_set_search_visited", 0)
    g._h.call("kg_set_search_visited_size", 0)
    for size in (384, 3072):
        if P <= 40:  # (a P = 300 pool leaves no LDS room for 3072 32-bit entries at 16 waves)
            assert res[0, size][5] == size and res[0, size][7] == 16
            assert res[2, size][5] == size and res[2, size][7] == 32
    assert res[0, 384][6] > 0
    for key in res:
        for a_, b_ in zip(res[1, 0][:5], res[key][:5]):
            assert np.array_equal(a_, b_), key
    ids, dd, c, comps, scans = idx.search(X, Q, L2, K=10, P=P, S=S, T=T)
    r = res[0, 3072]
    assert np.array_equal(r[2], c)
    for q in range(len(Q)):
        assert np.array_equal(r[0][q, :c[q]], ids[q, :c[q]]), q
        assert np.array_equal(r[1][q, :c[q]], dd[q, :c[q]]), q
    assert np.array_equal(r[3], comps) and np.array_equal(r[4], scans)


@pytest.mark.parametrize("P,S,M,T", [(40, 10, 0, 1), (10, 64, 0, 1), (100, 80, 0, 1), (20, 10, 30, 3),
                                     (30, 10, 1000, 1)])
def test_padded_adjacency_gives_identical_walks(P, S, M, T):
    # walks reading the padded rows (a node's ids and M loaded at its first
    # expansion) and the CSR lists with per-candidate records return the
    # oracle's ids, distances and counters; a search M above the row width
    # reads the CSR lists
    import ctypes
    X, Q, idx = setup("intmix", L2, d=128)
    g = gpu_index(X, idx, L2)
    res = {}
    for mode in (0, 1):
        g._h.call("kg_set_search_adjacency", mode)
        gi, gd = g.search(Q, K=10, P=P, S=S, M=M, T=T, withDistance=True, stats=True)
        padded, width = ctypes.c_int(-1), ctypes.c_uint()
        g._h.call("kg_search_adjacency", ctypes.byref(padded), ctypes.byref(width))
        res[mode] = (gi, gd, g.last_counts.copy(), g.last_comps.copy(), g.last_scans.copy(), padded.value,
                     width.value)
    g._h.call("kg_set_search_adjacency", 0)
    assert res[0][6] >= 16 and res[0][6] % 16 == 0
    assert res[0][5] == int(M <= res[0][6]) and res[1][5] == 0
    for a, b in zip(res[0][:5], res[1][:5]):
        assert np.array_equal(a, b)
    ids, dd, c, comps, scans = idx.search(X, Q, L2, K=10, P=P, S=S, M=M, T=T)
    assert np.array_equal(res[0][2], c)
    for q in range(len(Q)):
        assert np.array_equal(res[0][0][q, :c[q]], ids[q, :c[q]]), q
        assert np.array_equal(res[0][1][q, :c[q]], dd[q, :c[q]]), q
    assert np.array_equal(res[0][3], comps) and np.array_equal(res[0][4], scans)


def test_compact_visited_set_on_a_larger_index():
    # 300,000 nodes (19-bit ids): 16-bit entries keep 9 remainder bits at
    # 1024 entries (displacements up to 126); a 256-entry set would need 11
    # and takes 32-bit entries; every form walks identically to the bitset
    import ctypes
    import kgraph_amd
    from kgraph_amd import datagen
    X = datagen.integer_mixture(300000, 32, 300, 21, 22)
    Q = datagen.integer_mixture(500, 32, 300, 21, 23)
    g = kgraph_amd.KGraph(X, "euclidean", l2sqr=True)
    g.build(K=16, L=32, iterations=8, controls=20, reverse=-1)
    out = {}
    for mode, size in [(1, 0), (0, 0), (0, 1024), (0, 256), (2, 256)]:
        g._h.call("kg_set_search_visited", mode)
        g._h.call("kg_set_search_visited_size", size)
        gi, gd = g.search(Q, K=10, P=60, S=20, withDistance=True, stats=True)
        bits = ctypes.c_uint()
        g._h.call("kg_search_visited_bits", ctypes.byref(bits))
        out[mode, size] = (gi, gd, g.last_counts.copy(), g.last_comps.copy(), g.last_scans.copy(), bits.value)
    g._h.call("kg_set_search_visited", 0)
    g._h.call("kg_set_search_visited_size", 0)
    assert out[0, 0][5] == 16 and out[0, 1024][5] == 16 and out[0, 256][5] == 32 and out[2, 256][5] == 32
    for key in out:
        for a, b in zip(out[1, 0][:5], out[key][:5]):
            assert np.array_equal(a, b), key


def test_search_parity_mixed_batch():
    X, Q, idx = setup("intmix", L2, d=128)
    rs = np.random.RandomState(3)
    Ks = rs.choice([1, 10, 100], size=len(Q), p=[0.5, 0.4, 0.1]).astype(np.uint32)
    Ps = rs.choice([50, 100, 300, 500], size=len(Q)).astype(np.uint32)
    g = gpu_index(X, idx, L2)
    gi, gd = g.search(Q, K=10, P=100, Ks=Ks, Ps=Ps, withDistance=True, stats=True)
    for q in range(len(Q)):
        ids, d, c, comps, scans = idx.search(X, Q[q:q + 1], L2, K=int(Ks[q]), P=int(Ps[q]))
        assert g.last_counts[q] == c[0]
        assert np.array_equal(gi[q, :c[0]], ids[0, :c[0]])
        assert np.array_equal(gd[q, :c[0]], d[0, :c[0]])
        assert g.last_comps[q] == comps[0] and g.last_scans[q] == scans[0]


def test_search_parity_mixed_batch_with_entries():
    # per-query K / P with selected starting points: query q walks from
    # min(m, K_q + P_q - 1) of them (K = 1 queries included)
    X, Q, idx = setup("intmix", L2, d=128)
    rs = np.random.RandomState(5)
    Ks = rs.choice([1, 10, 100], size=len(Q), p=[0.5, 0.4, 0.1]).astype(np.uint32)
    Ps = rs.choice([5, 50, 300], size=len(Q)).astype(np.uint32)
    g = gpu_index(X, idx, L2)
    E, m = 512, 9
    g.set_entry_sample(E, seed=77)
    gi, gd = g.search(Q, K=10, P=100, Ks=Ks, Ps=Ps, entries=m, withDistance=True, stats=True)
    samp = _entry_sample_ids(77, E, len(X))
    si, _, _ = kgo.bruteforce(X[samp], Q, m, L2)
    init = np.zeros((len(Q), int(Ks.max())), np.uint32)
    init[:, :m] = samp[si]
    oi, od, oc, ocomp, oscan = idx.search(X, Q, L2, K=10, P=100, Ks=Ks, Ps=Ps, init=m, init_ids=init)
    assert np.array_equal(g.last_counts, oc)
    for q in range(len(Q)):
        n = oc[q]
        assert np.array_equal(gi[q, :n], oi[q, :n]) and np.array_equal(gd[q, :n], od[q, :n]), q
    assert np.array_equal(g.last_comps, ocomp) and np.array_equal(g.last_scans, oscan)


def test_search_bruteforce_path_when_P_ge_N():
    X, Q, idx = setup("uniform", EUC, n=500, nq=20, build=dict(K=10, L=20))
    compare(X, Q, idx, EUC, K=10, P=500)


def test_search_init_starting_points():
    X, Q, idx = setup("uniform", L2)
    import kgraph_amd
    from kgraph_amd import _lib
    g = gpu_index(X, idx, L2)
    init = np.random.RandomState(5).randint(0, len(X), size=(len(Q), 5)).astype(np.uint32)
    ids, d, c, comps, scans = idx.search(X, Q, L2, K=10, P=40, init=5, init_ids=init)
    gids = np.zeros((len(Q), 10), dtype=np.uint32)
    gids[:, :5] = init
    gd = np.zeros((len(Q), 10), dtype=np.float32)
    gc = np.zeros(len(Q), dtype=np.uint32)
    Qc = _lib.f32c(Q)
    sp = _lib.SearchParams(10, 0, 40, 10, 1, 1e30, 1998, 5)
    g._h.call("kg_search", _lib.ptr(Qc), len(Q), Qc.shape[1], sp, None, None, _lib.ptr(gids), _lib.ptr(gd), 10,
              _lib.ptr(gc), None, None)
    assert np.array_equal(gc, c) and np.array_equal(gids, ids) and np.array_equal(gd, d)
    # an out-of-range starting id (checked on the device before the walk)
    # raises and writes nothing; the next call starts clean
    bad = np.zeros((len(Q), 10), dtype=np.uint32)
    bad[:, :5] = init
    bad[len(Q) // 2, 3] = len(X)
    bd = np.full((len(Q), 10), -1.0, dtype=np.float32)
    bc = np.full(len(Q), 7, dtype=np.uint32)
    with pytest.raises(ValueError, match="initial id out of range"):
        g._h.call("kg_search", _lib.ptr(Qc), len(Q), Qc.shape[1], sp, None, None, _lib.ptr(bad), _lib.ptr(bd), 10,
                  _lib.ptr(bc), None, None)
    assert (bd == -1.0).all() and (bc == 7).all() and bad[len(Q) // 2, 3] == len(X)
    gids[:] = 0
    gids[:, :5] = init
    g._h.call("kg_search", _lib.ptr(Qc), len(Q), Qc.shape[1], sp, None, None, _lib.ptr(gids), _lib.ptr(gd), 10,
              _lib.ptr(gc), None, None)
    assert np.array_equal(gc, c) and np.array_equal(gids, ids) and np.array_equal(gd, d)


def test_bruteforce_parity():
    import kgraph_amd
    # intmix under 'angular' has large norms: the row norms must be correctly
    # rounded square roots, as the reference's std::sqrt is
    for kind, metric, d in [("uniform", EUC, 16), ("intmix", L2, 128), ("angular", ANG, 96), ("intmix", ANG, 32),
                            ("intmix", EUC, 48)]:
        X, Q = _dataset(kind, 20000, 50, d, 3)
        g = kgraph_amd.KGraph(X, "angular" if metric == ANG else "euclidean", l2sqr=(metric == L2))
        gi, gd = g.search(Q, K=100, withDistance=True)  # no index: exact scan
        bi, bd, bc = kgo.bruteforce(X, Q, 100, metric)
        assert np.array_equal(gi, bi) and np.array_equal(gd, bd), kind


def test_bruteforce_parity_many_buffer_flushes():
    # a block scans 40,000 rows per query and its candidate buffer fills and
    # is folded in many times (every row is a candidate until K are found):
    # every one of 3,000 queries must equal the oracle's exact top-K
    import kgraph_amd
    from kgraph_amd import datagen
    X = datagen.uniform(40000, 16, 13)
    Q = datagen.uniform(3000, 16, 14)
    for metric, K in ((EUC, 25), (L2, 10)):
        g = kgraph_amd.KGraph(X, "euclidean", l2sqr=(metric == L2))
        for rep in range(2):
            gi, gd = g.search(Q, K=K, withDistance=True)
            bi, bd, bc = kgo.bruteforce(X, Q, K, metric)
            bad = np.nonzero((gi != bi).any(1) | (gd != bd).any(1))[0]
            assert len(bad) == 0, "queries %s differ" % bad[:10]


def test_save_load_roundtrip_then_search(tmp_path):
    X, Q, idx = setup("uniform", EUC)
    p = str(tmp_path / "oracle.kgraph")
    idx.save(p)
    import kgraph_amd
    g = kgraph_amd.KGraph(X, "euclidean")
    g.load(p)
    p2 = str(tmp_path / "gpu.kgraph")
    g.save(p2)
    assert open(p, "rb").read() == open(p2, "rb").read()
    gi = g.search(Q, K=10, P=100)
    ids, d, c, _, _ = idx.search(X, Q, EUC, K=10, P=100)
    assert np.array_equal(gi, ids)


def _entry_sample_ids(seed, E, N):
    return np.sort(kgo.genrandom(seed, 1, E, N)[0])


# integer rows and queries take the exact MFMA selection (BF16 tiles equal the
# FP32 distances there); FP32 data the canonical FP32 scan.  Small samples
# (every lane sees 4 or 8 rows) make the 4-key first pass hand many queries to
# the second pass.
@pytest.mark.parametrize("kind,metric,d,m,E", [("intmix", L2, 128, 7, 512), ("uniform", EUC, 16, 7, 512),
                                               ("angular", ANG, 96, 7, 512), ("intmix", EUC, 64, 9, 512),
                                               ("intmix", ANG, 32, 16, 512), ("intmix", L2, 24, 3, 512),
                                               ("intmix", L2, 128, 16, 64), ("intmix", EUC, 64, 12, 128),
                                               ("intmix", L2, 128, 9, 100), ("intmix", L2, 128, 9, 20),
                                               ("uniform", L2, 16, 9, 100), ("intmix", L2, 200, 9, 100),
                                               ("intmix", EUC, 20, 5, 40), ("intmix", ANG, 72, 9, 4100)])
def test_entry_selection_parity(kind, metric, d, m, E):
    X, Q, idx = setup(kind, metric, d=d)
    g = gpu_index(X, idx, metric)
    g.set_entry_sample(E, seed=77)
    got = g.select_entries(Q, m)
    samp = _entry_sample_ids(77, E, len(X))
    si, sd, _ = kgo.bruteforce(X[samp], Q, m, metric)
    assert np.array_equal(got, samp[si])
    # the walk from those starting points equals the oracle's (kgraph.cpp:461-467)
    K = max(10, m + 1)
    gi, gd = g.search(Q, K=K, P=40, entries=m, withDistance=True, stats=True)
    oi, od, oc, ocomp, oscan = idx.search(X, Q, metric, K=K, P=40, init=m, init_ids=samp[si])
    assert np.array_equal(g.last_counts, oc)
    assert np.array_equal(gi, oi) and np.array_equal(gd, od)
    assert np.array_equal(g.last_comps, ocomp) and np.array_equal(g.last_scans, oscan)


def test_entry_selection_query_forms_in_turn():
    # integer rows: integer queries take the exact MFMA selection, others (a
    # fraction, or values past 256) the canonical FP32 scan; one handle serves
    # them in turn, and each selection equals the exact one
    X, Q, idx = setup("intmix", L2, d=128)
    g = gpu_index(X, idx, L2)
    E, m = 512, 7
    g.set_entry_sample(E, seed=77)
    samp = _entry_sample_ids(77, E, len(X))
    big = Q.copy()
    big[:, 5] = 300.0
    for Qv in (Q, Q + 0.25, Q, big, Q):
        got = g.select_entries(Qv, m)
        si, _, _ = kgo.bruteforce(X[samp], Qv, m, L2)
        assert np.array_equal(got, samp[si])


@pytest.mark.parametrize("kind,metric,d", [("angular", ANG, 96), ("uniform", EUC, 16)])
def test_entry_screen_bf16_reranks_exactly(kind, metric, d):
    # BF16 MFMA screen of FP32 rows: a 16-node shortlist per query re-ranked
    # with the exact distance -- returned in exact (dist, id) order, and equal
    # to the exact selection for nearly every query
    from kgraph_amd import PRECISION_BF16, PRECISION_FP32
    X, Q, idx = setup(kind, metric, d=d)
    g = gpu_index(X, idx, metric)
    E, m = 2048, 7
    g.set_entry_sample(E, seed=77)
    g._h.call("kg_set_precision", PRECISION_BF16)
    got = g.select_entries(Q, m)
    g._h.call("kg_set_precision", PRECISION_FP32)
    samp = _entry_sample_ids(77, E, len(X))
    si, sd, _ = kgo.bruteforce(X[samp], Q, m, metric)
    same = np.mean([np.array_equal(a, b) for a, b in zip(got, samp[si])])
    assert same >= 0.97, same
    for q in range(len(Q)):
        dd = [kgo.dist(metric, X[i], Q[q]) for i in got[q]]
        assert dd == sorted(dd)


def test_entry_selection_lifts_recall_on_separated_clusters():
    # 200 well separated clusters: the k-NN graph has no inter-cluster edges, so
    # random starts reach a query's cluster with probability ~P/200, while the
    # nearest sampled nodes lie in it.
    from kgraph_amd import datagen
    X = datagen.integer_mixture(20000, 64, 200, 5, 6, sigma=6.0, center_hi=250.0)
    Q = datagen.integer_mixture(300, 64, 200, 5, 7, sigma=6.0, center_hi=250.0)
    import kgraph_amd
    g = kgraph_amd.KGraph(X, "euclidean", l2sqr=True)
    g.build(K=20, L=40, iterations=12, controls=50, reverse=-1)
    _, gt_d, _ = kgo.bruteforce(X, Q, 10, L2)
    _, d_plain = g.search(Q, K=10, P=20, withDistance=True)
    g.set_entry_sample(2000, seed=3)
    _, d_ent = g.search(Q, K=10, P=20, entries=9, withDistance=True)
    r_plain = kgo.recall_at_k(d_plain, gt_d, 10)
    r_ent = kgo.recall_at_k(d_ent, gt_d, 10)
    print("recall plain %.3f entries %.3f" % (r_plain, r_ent))
    assert r_plain < 0.5 and r_ent > 0.95


def test_autotuned_launch_shape_keeps_results():
    # every launch shape the autotuner tries returns the same walk; the chosen
    # shape stays within the kernel's launch bound (<= 4 waves per block)
    X, Q, idx = setup("uniform", EUC)
    g = gpu_index(X, idx, EUC)
    before = g.search(Q, K=10, P=50, withDistance=True)
    wpb, bpc = g.autotune_search(Q, K=10, P=50)
    assert wpb in (1, 2, 4) and bpc in (4, 8, 16, 32)
    after = g.search(Q, K=10, P=50, withDistance=True)
    assert np.array_equal(before[0], after[0]) and np.array_equal(before[1], after[1])


@pytest.mark.parametrize("kind,metric,d,entries", [("uniform", EUC, 16, 0), ("angular", ANG, 96, 7)])
def test_walk_autotuner_keeps_results(kind, metric, d, entries):
    # every setting the walk autotuner tries (visited set, adjacency rows, walk
    # layout, BF16 screen, launch shape) returns the same walk, with random starts
    # and with selected entry points
    X, Q, idx = setup(kind, metric, d=d)
    g = gpu_index(X, idx, metric)
    if entries:
        g.set_entry_sample(512, seed=2)
    before = g.search(Q, K=10, P=50, withDistance=True, entries=entries)
    got = g.autotune_walk(Q, K=10, P=50, entries=entries)
    print(got)
    assert got["waves_per_block"] in (1, 2, 4) and got["blocks_per_cu"] in (0, 4, 8, 16, 32)
    assert got["visited_entries"] in (0, 256, 512, 1024, 2048, 4096, 8192)
    assert got["adjacency"] in ("padded", "csr")
    assert got["screen"] in (False, True)
    assert got["layout"] in ("breadth-first", "dataset")
    after = g.search(Q, K=10, P=50, withDistance=True, entries=entries)
    assert np.array_equal(before[0], after[0]) and np.array_equal(before[1], after[1])


@pytest.mark.parametrize("kind,metric,d", [("intmix", L2, 128), ("uniform", EUC, 37), ("angular", ANG, 96)])
def test_two_queries_per_wave_give_identical_walks(kind, metric, d):
    # each query on half a wave (kg_set_search_queries_per_wave 2) against the
    # whole-wave walk and the oracle: random starts, T > 1, M > 0, epsilon,
    # entry points, a mixed K/P batch, the LDS set and the bitset, every row form
    import ctypes
    X, Q, idx = setup(kind, metric, d=d)
    g = gpu_index(X, idx, metric)
    g.set_entry_sample(512, seed=2)
    rs = np.random.RandomState(4)
    Ks = rs.choice([1, 10, 25], size=len(Q)).astype(np.uint32)
    Ps = rs.choice([10, 40, 100], size=len(Q)).astype(np.uint32)
    cases = [dict(K=10, P=40), dict(K=10, P=200, S=64), dict(K=10, P=30, T=3), dict(K=10, P=64, M=40),
             dict(K=25, P=50, entries=9), dict(K=10, Ks=Ks, Ps=Ps), dict(K=10, P=100, S=16), dict(K=10, P=60, S=8)]
    for vis in (0, 1):
        g._h.call("kg_set_search_visited", vis)
        for rows in ((0, 1, 2) if kind == "intmix" else (0,)):
            g._h.call("kg_set_row_storage", rows)
            for kw in cases:
                out = {}
                for q in (1, 2):
                    g.set_search_queries_per_wave(q)
                    ids, dd = g.search(Q, withDistance=True, stats=True, **kw)
                    out[q] = (ids, dd, g.last_counts.copy(), g.last_comps.copy(), g.last_scans.copy())
                g.set_search_queries_per_wave(1)
                for a_, b_ in zip(out[1], out[2]):
                    assert np.array_equal(a_, b_), (vis, rows, kw)
    g._h.call("kg_set_search_visited", 0)
    g._h.call("kg_set_row_storage", 0)
    g.set_search_queries_per_wave(2)
    compare_g = g
    ids, dd, c, comps, scans = idx.search(X, Q, metric, K=10, P=100)
    gi, gd = compare_g.search(Q, K=10, P=100, withDistance=True, stats=True)
    assert np.array_equal(compare_g.last_counts, c)
    for q in range(len(Q)):
        assert np.array_equal(gi[q, :c[q]], ids[q, :c[q]]) and np.array_equal(gd[q, :c[q]], dd[q, :c[q]]), q
    assert np.array_equal(compare_g.last_comps, comps) and np.array_equal(compare_g.last_scans, scans)
    g.set_search_queries_per_wave(1)


def _walk_cases():
    import json
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "walk_cases.json")
    return json.load(open(path))["cases"]


@pytest.mark.parametrize("case", _walk_cases(), ids=[c["name"] for c in _walk_cases()])
def test_hand_traced_walks(case):
    # the hand-traced walks of tests/golden/walk_cases.json (rewind, S
    # batches, M below the degree, two trials merged) through the HIP walk
    import kgraph_amd
    from kgraph_amd import _lib
    x = np.asarray(case["x"], np.float32)[:, None]
    gl = case["graph"]
    off = np.zeros(len(gl) + 1, np.uint64)
    off[1:] = np.cumsum([len(v) for v in gl])
    ids = np.array([j for v in gl for j in v], np.uint32)
    d = np.array([(x[i, 0] - x[j, 0]) ** 2 for i, v in enumerate(gl) for j in v], np.float32)
    g = kgraph_amd.KGraph(x, "euclidean", l2sqr=True)
    g.set_graph(off, ids, d, np.asarray(case["M"], np.uint32))
    p, e = case["params"], case["expect"]
    K = p["K"]
    init = p.get("init") or []
    q = np.zeros((1, 1), np.float32)
    gi = np.zeros((1, K), np.uint32)
    gi[0, :len(init)] = init
    gd = np.zeros((1, K), np.float32)
    gc, gcomp, gscan = (np.zeros(1, np.uint32) for _ in range(3))
    sp = _lib.SearchParams(K, p["M"], p["P"], p["S"], p["T"], 1e30, p.get("seed", 1998), len(init))
    g._h.call("kg_search", _lib.ptr(q), 1, 1, sp, None, None, _lib.ptr(gi), _lib.ptr(gd), K, _lib.ptr(gc),
              _lib.ptr(gcomp), _lib.ptr(gscan))
    n = int(gc[0])
    assert gi[0, :n].tolist() == e["ids"] and gd[0, :n].tolist() == [float(v) for v in e["dists"]]
    assert (int(gcomp[0]), int(gscan[0])) == (e["comps"], e["scans"])


def test_search_parity_at_the_headline_scale():
    """The 1M x 128 integer-mixture workload at its full size: the index the
    GPU builds with the workload's IndexParams (K=100, L=150, S=10, R=100, 15
    iterations, reverse -1; the exact 8-bit join), searched at the benchmark's
    picked setting (K=10, P=50, S=128, starting from each query's 9 nearest
    nodes of a 4,096-node entry sample).  On 300 of the workload's queries the
    GPU walk returns the oracle's ids, distances, result counts, comps and
    scans bit-exactly over the same graph and the same starting points, and
    the GPU's entry selection equals the exact nearest sampled nodes."""
    import kgraph_amd
    from kgraph_amd import PRECISION_BF16, datagen
    N, E, m = 1_000_000, 4096, 9
    X = datagen.integer_mixture(N, 128, 1000, 3, 4)
    Q = np.ascontiguousarray(datagen.integer_mixture(10_000, 128, 1000, 3, 5)[:300])
    g = kgraph_amd.KGraph(X, "euclidean", l2sqr=True)
    info = g.build(K=100, L=150, S=10, R=100, iterations=15, controls=100, reverse=-1, precision=PRECISION_BF16)
    assert g.build_counters()["join_form"] == "int8-exact" and info.recall > 0.95
    g.set_entry_sample(E, seed=1998)
    sel = g.select_entries(Q, m)
    samp = _entry_sample_ids(1998, E, N)
    si, _, _ = kgo.bruteforce(X[samp], Q, m, L2)
    assert np.array_equal(sel, samp[si])
    gi, gd = g.search(Q, K=10, P=50, S=128, entries=m, withDistance=True, stats=True)
    idx = kgo.Index()
    idx.import_(*g.graph())
    init = np.zeros((len(Q), 10), np.uint32)
    init[:, :m] = sel
    oi, od, oc, ocomp, oscan = idx.search(X, Q, L2, K=10, P=50, S=128, init=m, init_ids=init)
    assert np.array_equal(g.last_counts, oc)
    for q in range(len(Q)):
        n = oc[q]
        assert np.array_equal(gi[q, :n], oi[q, :n]) and np.array_equal(gd[q, :n], od[q, :n]), q
    assert np.array_equal(g.last_comps, ocomp) and np.array_equal(g.last_scans, oscan)
    print("1M-node index: %d queries bit-exact, mean comps %.0f, scans %.0f" % (len(Q), ocomp.mean(), oscan.mean()))
