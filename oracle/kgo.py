"""ctypes binding of the CPU restatement oracle (oracle/kgraph_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, never by the kgraph_amd package.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _has_avx512():
    try:
        with open("/proc/cpuinfo") as f:
            flags = f.read()
        return " avx512f" in flags and " avx512bw" in flags and " avx512vl" in flags and " avx512dq" in flags
    except OSError:
        return False


def lib_path():
    name = "libkgraph_oracle_v4.so" if _has_avx512() else "libkgraph_oracle_v3.so"
    return os.path.join(HERE, name)


def build():
    subprocess.run(["make", "-s", "-C", HERE, "-j2"], check=True)


class kgo_index_params(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_uint), ("L", ctypes.c_uint), ("K", ctypes.c_uint),
                ("S", ctypes.c_uint), ("R", ctypes.c_uint), ("controls", ctypes.c_uint),
                ("seed", ctypes.c_uint), ("delta", ctypes.c_float), ("recall", ctypes.c_float),
                ("prune", ctypes.c_uint), ("reverse", ctypes.c_int)]


class kgo_index_info(ctypes.Structure):
    _fields_ = [("stop_condition", ctypes.c_int), ("iterations", ctypes.c_uint),
                ("cost", ctypes.c_float), ("recall", ctypes.c_float), ("accuracy", ctypes.c_float),
                ("delta", ctypes.c_float), ("M", ctypes.c_float)]


class kgo_search_params(ctypes.Structure):
    _fields_ = [("K", ctypes.c_uint), ("M", ctypes.c_uint), ("P", ctypes.c_uint),
                ("S", ctypes.c_uint), ("T", ctypes.c_uint), ("epsilon", ctypes.c_float),
                ("seed", ctypes.c_uint), ("init", ctypes.c_uint)]


_P = ctypes.c_void_p
_U = ctypes.c_uint
_SZ = ctypes.c_size_t
_I = ctypes.c_int
_SIG = {
    "kgo_dist": (ctypes.c_float, [_I, _P, _P, _U]),
    "kgo_row_norms": (None, [_I, _P, _U, _U, _SZ, _P]),
    "kgo_mt19937": (None, [_U, _U, _P]),
    "kgo_genrandom": (None, [_U, _U, _U, _U, _P]),
    "kgo_update_knn_list": (_U, [_P, _P, _U, _U, ctypes.c_float]),
    "kgo_index_new": (_P, []),
    "kgo_index_free": (None, [_P]),
    "kgo_index_error": (ctypes.c_char_p, [_P]),
    "kgo_set_threads": (_I, [_I]),
    "kgo_index_build": (_I, [_P, _P, _U, _U, _SZ, _I, ctypes.POINTER(kgo_index_params),
                             ctypes.POINTER(kgo_index_info), _P, _U]),
    "kgo_index_search": (_I, [_P, _P, _U, _U, _SZ, _I, _P, _U, _SZ, ctypes.POINTER(kgo_search_params),
                              _P, _P, _P, _P, _SZ, _P, _P, _P]),
    "kgo_index_search_trace": (_I, [_P, _P, _U, _U, _SZ, _I, _P, ctypes.POINTER(kgo_search_params), _P, _P, _P,
                                    _P, _P, _P, _U, _P, _P, _U, _P, _P, _U, _P]),
    "kgo_bruteforce": (_I, [_P, _U, _U, _SZ, _I, _P, _U, _SZ, _U, ctypes.c_float, _P, _P, _P]),
    "kgo_linear_knn": (_I, [_P, _U, _U, _SZ, _I, _P, _U, _U, _P, _P]),
    "kgo_index_size": (_U, [_P]),
    "kgo_index_edges": (_SZ, [_P]),
    "kgo_index_export": (None, [_P, _P, _P, _P, _P]),
    "kgo_index_import": (None, [_P, _U, _P, _P, _P, _P]),
    "kgo_index_save": (_I, [_P, ctypes.c_char_p, _I]),
    "kgo_index_load": (_I, [_P, ctypes.c_char_p]),
    "kgo_index_reverse": (_I, [_P, _I]),
    "kgo_index_prune": (_I, [_P, _U]),
}

_lib = None


def open_lib(path):
    """A copy of the oracle built elsewhere (tests build copies with planted
    mistakes to check that the walk pins reject them)."""
    l = ctypes.CDLL(path)
    for n, (r, a) in _SIG.items():
        f = getattr(l, n)
        f.restype = r
        f.argtypes = a
    return l


def lib():
    global _lib
    if _lib is None:
        p = lib_path()
        if not os.path.exists(p):
            build()
        _lib = open_lib(p)
    return _lib


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def set_threads(n=0):
    return lib().kgo_set_threads(n)


def dist(metric, a, b):
    a, b = f32(a), f32(b)
    return lib().kgo_dist(metric, _p(a), _p(b), a.shape[0])


def row_norms(metric, X):
    X = f32(X)
    out = np.zeros(X.shape[0], dtype=np.float32)
    lib().kgo_row_norms(metric, _p(X), X.shape[0], X.shape[1], X.shape[1], _p(out))
    return out


def mt19937(seed, n):
    out = np.zeros(n, dtype=np.uint32)
    lib().kgo_mt19937(seed, n, _p(out))
    return out


def genrandom(seed, trials, size, N):
    out = np.zeros(trials * size, dtype=np.uint32)
    lib().kgo_genrandom(seed, trials, size, N, _p(out))
    return out.reshape(trials, size)


def update_knn_list(ids, dists, K, nid, nd):
    ids = np.ascontiguousarray(ids, dtype=np.uint32).copy()
    dists = f32(dists).copy()
    r = lib().kgo_update_knn_list(_p(ids), _p(dists), K, nid, nd)
    return r, ids, dists


def bruteforce(X, Q, K, metric, epsilon=1e30):
    X, Q = f32(X), f32(Q)
    nq = Q.shape[0]
    ids = np.zeros((nq, K), dtype=np.uint32)
    d = np.zeros((nq, K), dtype=np.float32)
    c = np.zeros(nq, dtype=np.uint32)
    lib().kgo_bruteforce(_p(X), X.shape[0], X.shape[1], X.shape[1], metric, _p(Q), nq, Q.shape[1], K,
                         epsilon, _p(ids), _p(d), _p(c))
    return ids, d, c


def linear_knn(X, rows, K, metric):
    X = f32(X)
    rows = np.ascontiguousarray(rows, dtype=np.uint32)
    ids = np.zeros((len(rows), K), dtype=np.uint32)
    d = np.zeros((len(rows), K), dtype=np.float32)
    lib().kgo_linear_knn(_p(X), X.shape[0], X.shape[1], X.shape[1], metric, _p(rows), len(rows), K,
                         _p(ids), _p(d))
    return ids, d


class Index:
    """The oracle's KGraphImpl: build / search / save / load / reverse / prune."""

    def __init__(self, l=None):
        self._l = l if l is not None else lib()
        self.h = self._l.kgo_index_new()

    def __del__(self):
        h, self.h = getattr(self, "h", None), None
        if h:
            self._l.kgo_index_free(h)

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(self._l.kgo_index_error(self.h).decode())

    def build(self, X, metric, iterations=30, L=100, K=25, S=10, R=100, controls=100, seed=1998,
              delta=0.002, recall=0.99, prune=0, reverse=0):
        X = f32(X)
        p = kgo_index_params(iterations, L, K, S, R, controls, seed, delta, recall, prune, reverse)
        info = kgo_index_info()
        trace = np.zeros(4 * max(iterations, 1) + 4, dtype=np.float32)
        self._check(self._l.kgo_index_build(self.h, _p(X), X.shape[0], X.shape[1], X.shape[1], metric, p, info,
                                             _p(trace), len(trace)))
        self.info = info
        self.trace = trace[:4 * info.iterations].reshape(-1, 4)
        return info

    def search(self, X, Q, metric, K=25, P=100, M=0, S=10, T=1, epsilon=1e30, seed=1998, init=0,
               Ks=None, Ps=None, init_ids=None):
        X, Q = f32(X), f32(Q)
        nq = Q.shape[0]
        width = int(np.max(Ks)) if Ks is not None and nq else K
        ids = np.zeros((nq, width), dtype=np.uint32)
        if init_ids is not None:
            ids[:, :init_ids.shape[1]] = init_ids
        d = np.zeros((nq, width), dtype=np.float32)
        c = np.zeros(nq, dtype=np.uint32)
        comps = np.zeros(nq, dtype=np.uint32)
        scans = np.zeros(nq, dtype=np.uint32)
        sp = kgo_search_params(K, M, P, S, T, epsilon, seed, init)
        Ks_a = np.ascontiguousarray(Ks, dtype=np.uint32) if Ks is not None else None
        Ps_a = np.ascontiguousarray(Ps, dtype=np.uint32) if Ps is not None else None
        self._check(self._l.kgo_index_search(self.h, _p(X), X.shape[0], X.shape[1], X.shape[1], metric, _p(Q), nq,
                                             Q.shape[1], sp, _p(Ks_a), _p(Ps_a), _p(ids), _p(d), width, _p(c),
                                             _p(comps), _p(scans)))
        return ids, d, c, comps, scans

    def search_trace(self, X, q, metric, K=25, P=100, M=0, S=10, T=1, epsilon=1e30, seed=1998, init_ids=None):
        """One query's search and everything it did: returns (ids, dists,
        comps, scans, visited ids, their distances, expansions as rows of
        (node, beginM, endM), start of each trial's seeds in `visited`)."""
        X = f32(X)
        q = f32(q).ravel()
        init = 0 if init_ids is None else len(init_ids)
        ids = np.zeros(max(K, init, 1), dtype=np.uint32)
        if init:
            ids[:init] = init_ids
        d = np.zeros(max(K, 1), dtype=np.float32)
        comps, scans = ctypes.c_uint(), ctypes.c_uint()
        cap = X.shape[0] * T + 16
        vis = np.zeros(cap, dtype=np.uint32)
        vd = np.zeros(cap, dtype=np.float32)
        ecap = 4 * X.shape[0] * T + 16
        exps = np.zeros(3 * ecap, dtype=np.uint32)
        ts = np.zeros(T + 1, dtype=np.uint32)
        nv, ne, nt = ctypes.c_uint(), ctypes.c_uint(), ctypes.c_uint()
        sp = kgo_search_params(K, M, P, S, T, epsilon, seed, init)
        r = self._l.kgo_index_search_trace(self.h, _p(X), X.shape[0], X.shape[1], X.shape[1], metric, _p(q), sp,
                                           _p(ids), _p(d), ctypes.byref(comps), ctypes.byref(scans), _p(vis), _p(vd),
                                           cap, ctypes.byref(nv), _p(exps), ecap, ctypes.byref(ne), _p(ts), T + 1,
                                           ctypes.byref(nt))
        if r < 0:
            raise RuntimeError(self._l.kgo_index_error(self.h).decode())
        assert nv.value <= cap and ne.value <= ecap
        return (ids[:r], d[:r], comps.value, scans.value, vis[:nv.value], vd[:nv.value],
                exps[:3 * ne.value].reshape(-1, 3), ts[:nt.value])

    def export(self):
        n = self._l.kgo_index_size(self.h)
        e = self._l.kgo_index_edges(self.h)
        off = np.zeros(n + 1, dtype=np.uint64)
        ids = np.zeros(max(e, 1), dtype=np.uint32)
        d = np.zeros(max(e, 1), dtype=np.float32)
        M = np.zeros(max(n, 1), dtype=np.uint32)
        self._l.kgo_index_export(self.h, _p(off), _p(ids), _p(d), _p(M))
        return off, ids[:e], d[:e], M[:n]

    def import_(self, off, ids, dists, M):
        off = np.ascontiguousarray(off, dtype=np.uint64)
        ids = np.ascontiguousarray(ids, dtype=np.uint32)
        d = f32(dists) if dists is not None else None
        M = np.ascontiguousarray(M, dtype=np.uint32)
        self._l.kgo_index_import(self.h, len(M), _p(off), _p(ids), _p(d), _p(M))

    def save(self, path, fmt=0):
        self._check(self._l.kgo_index_save(self.h, path.encode(), fmt))

    def load(self, path):
        self._check(self._l.kgo_index_load(self.h, path.encode()))

    def reverse(self, rev):
        self._check(self._l.kgo_index_reverse(self.h, rev))

    def prune(self, level):
        self._check(self._l.kgo_index_prune(self.h, level))


def recall_at_k(result_dists, gt_dists, K):
    """AverageRecall (kgraph-data.h:333-364): distance-equality matches of the
    sorted per-row distances, so ties never count as misses."""
    found = 0
    for r, g in zip(result_dists, gt_dists):
        r = np.sort(np.asarray(r[:K], dtype=np.float32))
        g = np.sort(np.asarray(g[:K], dtype=np.float32))
        gi = ri = 0
        while gi < K and ri < K:
            if g[gi] < r[ri]:
                gi += 1
            elif g[gi] == r[ri]:
                found += 1
                gi += 1
                ri += 1
            else:
                raise RuntimeError("distance is unstable")
    return found / (K * len(result_dists))
