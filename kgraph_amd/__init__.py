"""kgraph_amd -- KGraph's NN-descent k-NN graph build and graph-walk search on an
AMD Instinct MI355X.

`KGraph(data, metric)` mirrors the reference's Python class
(python-api.cpp:331-450): same constructor, `build` / `search` keyword
arguments and defaults, `load` / `save` / `save_text`, uint32 ids and float32
distances.  All compute runs in libkgraph_amd.so (HIP kernels for gfx950)
through its C ABI (include/kgraph_amd.h).
"""
import numpy as np

from . import _lib
from ._lib import (FORMAT_DEFAULT, FORMAT_NO_DIST, FORMAT_TEXT, METRIC_ANGULAR, METRIC_EUCLIDEAN,
                   METRIC_L2SQR, PRECISION_BF16, PRECISION_FP32, PRECISION_FP32_MFMA, Handle, IndexInfo,
                   IndexParams, SearchParams, f32c, ptr, u32c)

# kgraph.h:14-37
default_iterations = 30
default_L = 100
default_K = 25
default_P = 100
default_M = 0
default_T = 1
default_S = 10
default_R = 100
default_controls = 100
default_seed = 1998
default_delta = 0.002
default_recall = 0.99
default_epsilon = 1e30
default_prune = 0
default_reverse = 0

git_commit = "kgraph_amd"
arch = {"name": "gfx950", "alignment": 16}

_METRICS = {"euclidean": METRIC_EUCLIDEAN, "angular": METRIC_ANGULAR}


def version():
    return _lib.lib().kg_version().decode()


class KGraph:
    """kgraph.KGraph (python-api.cpp:331-410).

    data: 2-D float32 or float64 array (float64 is converted to float32; the
    distance arithmetic is FP32).  metric: 'euclidean' (||x-y||^2 / 2, as
    n_x + n_y - dot) or 'angular' (-cos).  The rows are copied to the GPU here,
    as the reference copies them into an aligned matrix (python-api.cpp:255).
    """

    def __init__(self, data, metric, *, l2sqr=False):
        if metric not in _METRICS:
            raise RuntimeError("metric not supported")
        data = np.asarray(data)
        if data.ndim != 2:
            raise ValueError("bad array")
        if data.dtype not in (np.float32, np.float64):
            raise RuntimeError("data type not supported.")
        self._h = Handle()
        self._data = f32c(data)
        self._metric = METRIC_L2SQR if l2sqr else _METRICS[metric]
        self._hasIndex = False
        self._h.call("kg_set_data", ptr(self._data), self._data.shape[0], self._data.shape[1],
                     self._data.shape[1], self._metric)

    # -- index storage (python-api.cpp:237-246)
    def load(self, path):
        self._h.call("kg_load", path.encode())
        self._hasIndex = True

    def save(self, path):
        self._h.call("kg_save", path.encode(), FORMAT_DEFAULT)

    def save_text(self, path):
        self._h.call("kg_save", path.encode(), FORMAT_TEXT)

    # -- build (python-api.cpp:369-389, defaults :429-438)
    def build(self, iterations=default_iterations, L=default_L, K=default_K, S=default_S,
              controls=default_controls, delta=default_delta, recall=default_recall,
              prune=default_prune, reverse=default_reverse, *, R=default_R, seed=default_seed,
              precision=None, graphs=None, join=None, prerank=None, int8=None):
        """graphs: True (default) runs the iterations as replays of a captured
        HIP graph, False launches their kernels one by one.  join: 'auto'
        (default), 'screen' or 'tiles' -- the local-join form when the BF16
        join is a screen (precision BF16 on rows that are not BF16-exact).
        prerank: fraction of each node's join lists kept by the learned
        pre-rank (needs a projector, train_projector; 0 or 1 = off); see
        prerank_report().  int8: True (default) lets precision BF16 join rows
        of integers in [0, 255] on 8-bit MFMA operands (exact), False keeps
        the BF16 form."""
        p = IndexParams(iterations, L, K, S, R, controls, seed, delta, recall, prune, reverse)
        if precision is not None:
            if isinstance(precision, str):
                precision = {"fp32": _lib.PRECISION_FP32, "bf16": _lib.PRECISION_BF16,
                             "fp32-mfma": _lib.PRECISION_FP32_MFMA}[precision]
            self._h.call("kg_set_precision", precision)
        if graphs is not None:
            self._h.call("kg_set_build_graphs", 1 if graphs else 0)
        if join is not None:
            self._h.call("kg_set_join_form", {"auto": _lib.JOIN_AUTO, "screen": _lib.JOIN_SCREEN,
                                              "tiles": _lib.JOIN_TILES}[join])
        if prerank is not None:
            self._h.call("kg_set_join_prerank", float(prerank))
        if int8 is not None:
            self._h.call("kg_set_join_int8", 1 if int8 else 0)
        info = IndexInfo()
        self._h.call("kg_build", p, info)
        self._hasIndex = True
        self.info = info
        return info

    # -- search (python-api.cpp:266-328, 390-409, defaults :440-448)
    def set_entry_sample(self, E, seed=default_seed):
        """Sample E nodes as entry-point candidates for search(entries=m)."""
        self._h.call("kg_set_entry_sample", E, seed)

    def select_entries(self, query, m):
        """The m nearest sampled nodes of every query (uint32[Q, m])."""
        q = f32c(query)
        ids = np.zeros((q.shape[0], m), dtype=np.uint32)
        self._h.call("kg_select_entries", ptr(q), q.shape[0], q.shape[1], m, ptr(ids), m)
        return ids

    def search(self, query, K=default_K, P=default_P, M=default_M, S=default_S, T=default_T,
               threads=0, withDistance=False, blas=False, *, epsilon=default_epsilon,
               seed=default_seed, Ks=None, Ps=None, stats=False, entries=0):
        """Returns uint32[Q, K] ids, or (ids, float32[Q, K] dists) with
        withDistance.  `threads` and `blas` are accepted and ignored (the GPU
        path has no OpenMP pool, and `blas` selects an exact scan that the
        index-less path already provides).  Ks / Ps give per-query K / P for
        one batched launch; the output then has max(Ks) columns.  Rows past a
        short result are zero (the reference leaves them uninitialised,
        python-api.cpp:270-273).  entries=m > 0 starts each walk from the m
        nearest nodes of the entry sample (set_entry_sample), passed to the
        walk as the reference's SearchParams::init starting points."""
        q = f32c(query)
        if q.ndim != 2 or q.shape[1] != self._data.shape[1]:
            raise ValueError("bad array")
        nq = q.shape[0]
        Ks_a = u32c(Ks) if Ks is not None else None
        Ps_a = u32c(Ps) if Ps is not None else None
        width = int(Ks_a.max()) if Ks_a is not None and nq else K
        ids = np.zeros((nq, width), dtype=np.uint32)
        dists = np.zeros((nq, width), dtype=np.float32)
        counts = np.zeros(nq, dtype=np.uint32)
        comps = np.zeros(nq, dtype=np.uint32) if stats else None
        scans = np.zeros(nq, dtype=np.uint32) if stats else None
        if self._hasIndex:
            init = 0
            if entries:
                # per-query K / P: query q starts from min(entries, K_q + P_q - 1) of them
                self._h.call("kg_select_entries", ptr(q), nq, q.shape[1], entries, ptr(ids), width)
                init = entries
            sp = SearchParams(K, M, P, S, T, epsilon, seed, init)
            self._h.call("kg_search", ptr(q), nq, q.shape[1], sp, ptr(Ks_a), ptr(Ps_a), ptr(ids),
                         ptr(dists), width, ptr(counts), ptr(comps), ptr(scans))
        else:
            # no index: exact scan (python-api.cpp:303-314)
            if Ks_a is not None and nq and (Ks_a != Ks_a[0]).any():
                raise ValueError("brute-force search needs one K")
            self._h.call("kg_bruteforce", ptr(q), nq, q.shape[1], width, epsilon, ptr(ids), ptr(dists),
                         width, ptr(counts))
        self.last_counts = counts
        if stats:
            self.last_comps, self.last_scans = comps, scans
        if withDistance:
            return ids, dists
        return ids

    # -- extensions beyond the reference's Python class
    def get_nn(self, i):
        """KGraph::get_nn (kgraph.cpp:566-582): (ids, dists, M)."""
        M = ctypes_uint()
        L = ctypes_uint()
        self._h.call("kg_get_nn", i, None, None, _lib.ctypes.byref(M), _lib.ctypes.byref(L))
        ids = np.zeros(L.value, dtype=np.uint32)
        d = np.zeros(L.value, dtype=np.float32)
        self._h.call("kg_get_nn", i, ptr(ids), ptr(d), _lib.ctypes.byref(M), _lib.ctypes.byref(L))
        return ids, d, M.value

    def graph(self):
        """Whole index as CSR numpy arrays: (offsets, ids, dists, M)."""
        return export_graph(self._h)

    def set_graph(self, offsets, ids, dists, M):
        import_graph(self._h, offsets, ids, dists, M)
        self._hasIndex = True

    def reverse(self, rev):
        self._h.call("kg_reverse", rev)

    def prune(self, level):
        self._h.call("kg_prune", level)

    # -- learned projector (no counterpart in the reference)
    def train_projector(self, steps=500, batch=4096, hidden=256, out=32, lr=2e-3, margin=1.0, seed=1):
        """Train the D -> hidden -> out MLP on triplets from the current index;
        returns the per-step mean loss."""
        self._h.call("kg_mlp_init", hidden, out, seed)
        trace = np.zeros(max(steps, 1), dtype=np.float32)
        self._h.call("kg_mlp_train", steps, batch, lr, margin, seed, ptr(trace))
        self._proj_shape = (hidden, out)
        return trace[:steps]

    def set_entry_projection(self, shortlist):
        """shortlist > 0: score the entry sample in the projected space and
        re-rank the best `shortlist` exactly; 0: exact scoring."""
        self._h.call("kg_set_entry_projection", shortlist)
        self._shortlist = shortlist

    def validate_entry_projection(self, probe, K=10, P=20, entries=9, shortlist=16, tol=0.005):
        """Probe-set guard: compare recall@K of walks started from exactly and
        from projector-selected entry points on `probe` queries; keep the
        projector only if it loses at most `tol`, else fall back to exact
        scoring.  Returns {'recall_exact', 'recall_learned', 'enabled'}."""
        q = f32c(probe)
        nq = q.shape[0]
        gi = np.zeros((nq, K), np.uint32)
        gd = np.zeros((nq, K), np.float32)
        gc = np.zeros(nq, np.uint32)
        self._h.call("kg_bruteforce", ptr(q), nq, q.shape[1], K, default_epsilon, ptr(gi), ptr(gd), K, ptr(gc))

        def recall(d):
            hit = 0
            for r, g in zip(np.sort(d, 1), np.sort(gd, 1)):
                i = j = 0
                while i < K and j < K:
                    if g[j] < r[i]:
                        j += 1
                    elif g[j] == r[i]:
                        hit += 1
                        i += 1
                        j += 1
                    else:
                        raise RuntimeError("distance is unstable")
            return hit / float(K * nq)

        self.set_entry_projection(0)
        _, d0 = self.search(q, K=K, P=P, entries=entries, withDistance=True)
        self.set_entry_projection(shortlist)
        _, d1 = self.search(q, K=K, P=P, entries=entries, withDistance=True)
        r0, r1 = recall(d0), recall(d1)
        ok = r1 >= r0 - tol
        if not ok:
            self.set_entry_projection(0)
        return {"recall_exact": r0, "recall_learned": r1, "enabled": ok}

    def last_time_ms(self, which=0):
        v = _lib.ctypes.c_float()
        self._h.call("kg_last_time_ms", which, _lib.ctypes.byref(v))
        return v.value

    def autotune_search(self, query, K=default_K, P=default_P, S=default_S, *, Ks=None, Ps=None):
        """Times the search kernel on `query` for every launch shape it supports
        and keeps the fastest for later searches; returns (waves per block,
        blocks per CU)."""
        from . import devmem
        q = f32c(query)
        dq = devmem.DeviceArray.from_host(q)
        chosen = (_lib.ctypes.c_uint * 2)()
        sp = SearchParams(K, default_M, P, S, default_T, default_epsilon, default_seed, 0)
        self._h.call("kg_autotune_search", dq.ptr, q.shape[0], q.shape[1], sp, ptr(u32c(Ks) if Ks is not None else None),
                     ptr(u32c(Ps) if Ps is not None else None), chosen)
        devmem.synchronize()
        return chosen[0], chosen[1]

    def autotune_walk(self, query, K=default_K, P=default_P, S=default_S, *, Ks=None, Ps=None, entries=0):
        """Times the walk on `query` over the settings that leave its results
        unchanged (visited set, adjacency rows, walk layout, BF16 screen,
        queries per wave, launch shape), starting from the entry points of
        `entries` as search() would, and keeps the fastest for later searches.
        Returns the chosen settings."""
        from . import devmem
        q = f32c(query)
        nq = q.shape[0]
        Ks_a = u32c(Ks) if Ks is not None else None
        Ps_a = u32c(Ps) if Ps is not None else None
        width = int(Ks_a.max()) if Ks_a is not None and nq else K
        dq = devmem.DeviceArray.from_host(q)
        init = None
        if entries:
            ids = np.zeros((nq, max(width, entries)), dtype=np.uint32)
            self._h.call("kg_select_entries", ptr(q), nq, q.shape[1], entries, ptr(ids), ids.shape[1])
            init = devmem.DeviceArray.from_host(ids)
        chosen = (_lib.ctypes.c_uint * 6)()
        sp = SearchParams(K, default_M, P, S, default_T, default_epsilon, default_seed, entries)
        self._h.call("kg_autotune_walk", dq.ptr, nq, q.shape[1], sp, ptr(Ks_a), ptr(Ps_a),
                     init.ptr if init is not None else None, init.shape[1] if init is not None else 0, chosen)
        devmem.synchronize()
        order = _lib.ctypes.c_int()
        self._h.call("kg_search_order", _lib.ctypes.byref(order), None, None, None)
        return {"waves_per_block": chosen[0], "blocks_per_cu": chosen[1], "visited_entries": chosen[2],
                "adjacency": "csr" if chosen[3] else "padded", "screen": bool(chosen[4]),
                "queries_per_wave": chosen[5], "layout": "breadth-first" if order.value else "dataset"}

    def build_counters(self):
        """Last build: {'pairs', 'list_entries', 'pushes', 'iterations',
        'join_form', 'graph_replays', 'rescored'} with join_form 'fp32',
        'bf16-screen', 'bf16-exact', 'fp32-mfma-screen' or 'int8-exact', graph_replays the
        iterations run as replays of the captured iteration graph and rescored
        the pairs the screen sent to the exact FP32 re-score."""
        out = np.zeros(7, dtype=np.uint64)
        self._h.call("kg_build_counters", ptr(out), 7)
        forms = {0: "fp32", 1: "bf16-screen", 2: "bf16-exact", 3: "fp32-mfma-screen", 4: "int8-exact"}
        return {"pairs": int(out[0]), "list_entries": int(out[1]), "pushes": int(out[2]),
                "iterations": int(out[3]), "join_form": forms.get(int(out[4]), "fp32"),
                "graph_replays": int(out[5]), "rescored": int(out[6])}

    def prerank_report(self):
        """The last build's learned pre-rank: {'status': 'off' | 'on' |
        'probe-fallback' | 'recall-fallback', 'probe_accuracy', 'fallback_iteration'}."""
        out = np.zeros(3, dtype=np.float32)
        self._h.call("kg_build_prerank", ptr(out))
        return {"status": ["off", "on", "probe-fallback", "recall-fallback"][int(out[0])],
                "probe_accuracy": float(out[1]), "fallback_iteration": int(out[2])}

    def set_search_screen(self, on):
        """Screen walk candidates with BF16 dot products on MFMA before the
        exact FP32 re-score (walks over the FP32 rows; same results)."""
        self._h.call("kg_set_search_screen", 1 if on else 0)

    def set_search_queries_per_wave(self, q):
        """Walk 1 query per wavefront (default) or 2, each on half a wave
        (same results)."""
        self._h.call("kg_set_search_queries_per_wave", int(q))

    def search_screen_report(self):
        """The last walk launch: {'active': bool, 'rescored': candidates the
        screen passed to the exact FP32 re-score}."""
        act = _lib.ctypes.c_int()
        n = _lib.ctypes.c_ulonglong()
        self._h.call("kg_search_screen", _lib.ctypes.byref(act), _lib.ctypes.byref(n))
        return {"active": bool(act.value), "rescored": int(n.value)}

    def screen_dists(self, query, ids):
        """The walk screen's BF16 distance and error bound of node ids[i, j]
        against query i: (screened, bound), both nq x m."""
        q = f32c(query)
        ii = np.ascontiguousarray(ids, dtype=np.uint32)
        nq, m = ii.shape
        d = np.zeros((nq, m), np.float32)
        e = np.zeros((nq, m), np.float32)
        self._h.call("kg_search_screen_dists", ptr(q), nq, q.shape[1], ptr(ii), m, ptr(d), ptr(e))
        return d, e

    def build_iterations(self):
        """Last build, one row per iteration: pairs, pushes, list entries,
        join form (0 FP32 tiles, 1 BF16 screen, 2 exact BF16 tiles, 3 probe:
        screen and tiles on half of the nodes each), join kernel us (0 for
        graph replays), iteration us."""
        n = _lib.ctypes.c_uint()
        self._h.call("kg_build_iterations", None, 0, _lib.ctypes.byref(n))
        out = np.zeros(6 * max(n.value, 1), dtype=np.uint64)
        self._h.call("kg_build_iterations", ptr(out), n.value, _lib.ctypes.byref(n))
        return out[:6 * n.value].reshape(-1, 6)

    def build_trace(self):
        n = _lib.ctypes.c_uint()
        self._h.call("kg_build_trace", None, 0, _lib.ctypes.byref(n))
        out = np.zeros(4 * max(n.value, 1), dtype=np.float32)
        self._h.call("kg_build_trace", ptr(out), n.value, _lib.ctypes.byref(n))
        return out[:4 * n.value].reshape(-1, 4)


def ctypes_uint():
    return _lib.ctypes.c_uint()


def export_graph(h):
    n = _lib.ctypes.c_uint()
    e = _lib.ctypes.c_uint64()
    h.call("kg_index_size", _lib.ctypes.byref(n), _lib.ctypes.byref(e))
    off = np.zeros(n.value + 1, dtype=np.uint64)
    ids = np.zeros(max(e.value, 1), dtype=np.uint32)
    d = np.zeros(max(e.value, 1), dtype=np.float32)
    M = np.zeros(max(n.value, 1), dtype=np.uint32)
    h.call("kg_get_graph", ptr(off), ptr(ids), ptr(d), ptr(M))
    return off, ids[:e.value], d[:e.value], M[:n.value]


def import_graph(h, offsets, ids, dists, M):
    off = np.ascontiguousarray(offsets, dtype=np.uint64)
    ids = u32c(ids)
    M = u32c(M)
    d = f32c(dists) if dists is not None else None
    h.call("kg_set_graph", len(M), ptr(off), ptr(ids), ptr(d), ptr(M))


__all__ = ["KGraph", "IndexParams", "SearchParams", "IndexInfo", "version", "export_graph", "import_graph"]
