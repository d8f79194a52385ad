"""ctypes binding of libkgraph_amd.so (the C ABI declared in include/kgraph_amd.h).

The library is loaded from this package directory only; there is no fallback
implementation, so every entry point fails loudly when the HIP library has not
been built.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libkgraph_amd.so")

KG_OK = 0
KG_ERR_INVALID_ARGUMENT = -1
KG_ERR_RUNTIME = -2
KG_ERR_IO = -3
KG_ERR_DEVICE = -4

METRIC_L2SQR = 0
METRIC_EUCLIDEAN = 1
METRIC_ANGULAR = 2

FORMAT_DEFAULT = 0
FORMAT_NO_DIST = 1
FORMAT_TEXT = 128

PRECISION_FP32 = 0
PRECISION_BF16 = 1
PRECISION_FP32_MFMA = 2

JOIN_AUTO = 0
JOIN_SCREEN = 1
JOIN_TILES = 2


class IndexParams(ctypes.Structure):
    """KGraph::IndexParams (kgraph.h:94-110)."""
    _fields_ = [("iterations", ctypes.c_uint), ("L", ctypes.c_uint), ("K", ctypes.c_uint),
                ("S", ctypes.c_uint), ("R", ctypes.c_uint), ("controls", ctypes.c_uint),
                ("seed", ctypes.c_uint), ("delta", ctypes.c_float), ("recall", ctypes.c_float),
                ("prune", ctypes.c_uint), ("reverse", ctypes.c_int)]


class IndexInfo(ctypes.Structure):
    """KGraph::IndexInfo (kgraph.h:135-147)."""
    _fields_ = [("stop_condition", ctypes.c_int), ("iterations", ctypes.c_uint),
                ("cost", ctypes.c_float), ("recall", ctypes.c_float), ("accuracy", ctypes.c_float),
                ("delta", ctypes.c_float), ("M", ctypes.c_float)]


class SearchParams(ctypes.Structure):
    """KGraph::SearchParams (kgraph.h:113-126)."""
    _fields_ = [("K", ctypes.c_uint), ("M", ctypes.c_uint), ("P", ctypes.c_uint),
                ("S", ctypes.c_uint), ("T", ctypes.c_uint), ("epsilon", ctypes.c_float),
                ("seed", ctypes.c_uint), ("init", ctypes.c_uint)]


_P = ctypes.c_void_p
_U = ctypes.c_uint
_SZ = ctypes.c_size_t
_I = ctypes.c_int

# name -> (restype, argtypes)
SIGNATURES = {
    "kg_version": (ctypes.c_char_p, []),
    "kg_create": (_P, []),
    "kg_destroy": (None, [_P]),
    "kg_last_error": (ctypes.c_char_p, [_P]),
    "kg_index_params_default": (None, [ctypes.POINTER(IndexParams)]),
    "kg_search_params_default": (None, [ctypes.POINTER(SearchParams)]),
    "kg_set_verbosity": (None, [_U]),
    "kg_set_data": (_I, [_P, _P, _U, _U, _SZ, _I]),
    "kg_set_data_device": (_I, [_P, _P, _U, _U, _SZ, _I]),
    "kg_set_precision": (_I, [_P, _I]),
    "kg_set_row_storage": (_I, [_P, _I]),
    "kg_set_search_visited": (_I, [_P, _I]),
    "kg_search_visited_stats": (_I, [_P, _P, _P]),
    "kg_set_search_visited_size": (_I, [_P, _U]),
    "kg_search_visited_bits": (_I, [_P, _P]),
    "kg_set_search_adjacency": (_I, [_P, _I]),
    "kg_search_adjacency": (_I, [_P, _P, _P]),
    "kg_set_search_screen": (_I, [_P, _I]),
    "kg_set_search_queries_per_wave": (_I, [_P, _U]),
    "kg_autotune_walk": (_I, [_P, _P, _U, _SZ, ctypes.POINTER(SearchParams), _P, _P, _P, _SZ, _P]),
    "kg_search_screen": (_I, [_P, _P, _P]),
    "kg_search_screen_dists": (_I, [_P, _P, _U, _SZ, _P, _U, _P, _P]),
    "kg_set_build_graphs": (_I, [_P, _I]),
    "kg_screen_pairs": (_I, [_P, _P, _I, _P, _P]),
    "kg_build_iterations": (_I, [_P, _P, _I, _P]),
    "kg_set_join_form": (_I, [_P, _I]),
    "kg_row_storage": (_I, [_P, _P, _P]),
    "kg_build": (_I, [_P, ctypes.POINTER(IndexParams), ctypes.POINTER(IndexInfo)]),
    "kg_search": (_I, [_P, _P, _U, _SZ, ctypes.POINTER(SearchParams), _P, _P, _P, _P, _SZ, _P, _P, _P]),
    "kg_search_device": (_I, [_P, _P, _U, _SZ, ctypes.POINTER(SearchParams), _P, _P, _P, _P, _SZ, _P, _P, _P]),
    "kg_bruteforce": (_I, [_P, _P, _U, _SZ, _U, ctypes.c_float, _P, _P, _SZ, _P]),
    "kg_set_search_launch": (_I, [_P, _U, _U]),
    "kg_set_join_prerank": (_I, [_P, ctypes.c_float]),
    "kg_build_prerank": (_I, [_P, _P]),
    "kg_bruteforce_device_noalloc": (_I, [_P, _P, _U, _SZ, _U, ctypes.c_float, _P, _P, _SZ, _P]),
    "kg_save": (_I, [_P, ctypes.c_char_p, _I]),
    "kg_load": (_I, [_P, ctypes.c_char_p]),
    "kg_get_nn": (_I, [_P, _U, _P, _P, _P, _P]),
    "kg_reverse": (_I, [_P, _I]),
    "kg_prune": (_I, [_P, _U]),
    "kg_index_size": (_I, [_P, _P, _P]),
    "kg_set_graph": (_I, [_P, _U, _P, _P, _P, _P]),
    "kg_set_search_order": (_I, [_P, _I]),
    "kg_search_order": (_I, [_P, _P, _P, _P, _P]),
    "kg_prepare_search": (_I, [_P]),
    "kg_set_search_lean": (_I, [_P, _I]),
    "kg_set_join_int8": (_I, [_P, _I]),
    "kg_get_graph": (_I, [_P, _P, _P, _P, _P]),
    "kg_last_time_ms": (_I, [_P, _I, _P]),
    "kg_build_trace": (_I, [_P, _P, _U, _P]),
    "kg_build_counters": (_I, [_P, _P, _U]),
    "kg_set_entry_sample": (_I, [_P, _U, _U]),
    "kg_search_latencies": (_I, [_P, _P, _U]),
    "kg_mlp_init": (_I, [_P, _U, _U, _U]),
    "kg_mlp_train": (_I, [_P, _U, _U, ctypes.c_float, ctypes.c_float, _U, _P]),
    "kg_mlp_loss_grad": (_I, [_P, _P, _U, ctypes.c_float, _P, _P, _P, _P, _P]),
    "kg_mlp_get_weights": (_I, [_P, _P, _P, _P, _P]),
    "kg_mlp_get_normalization": (_I, [_P, _P, _P]),
    "kg_mlp_set_weights": (_I, [_P, _P, _P, _P, _P]),
    "kg_mlp_project": (_I, [_P, _P, _U, _SZ, _P]),
    "kg_set_entry_projection": (_I, [_P, _U]),
    "kg_autotune_search": (_I, [_P, _P, _U, _SZ, ctypes.POINTER(SearchParams), _P, _P, _P]),
    "kg_select_entries": (_I, [_P, _P, _U, _SZ, _U, _P, _SZ]),
    "kg_select_entries_device": (_I, [_P, _P, _U, _SZ, _U, _P, _SZ]),
}

_lib = None


class KGraphError(RuntimeError):
    pass


def lib():
    """Loads libkgraph_amd.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libkgraph_amd.so is not built: run `python -m kgraph_amd.build` "
                              "(or __graft_entry__.build())")
        l = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def ptr(a):
    """Address of a numpy array's data (None for None)."""
    if a is None:
        return None
    return ctypes.c_void_p(a.ctypes.data)


class Handle:
    """Owns one kg_handle; raises the reference's exception types on error."""

    def __init__(self):
        self._l = lib()
        self.h = self._l.kg_create()
        if not self.h:
            raise RuntimeError("error creating kgraph instance")

    def __del__(self):
        h, self.h = getattr(self, "h", None), None
        if h:
            self._l.kg_destroy(h)

    def check(self, rc):
        if rc == KG_OK:
            return
        msg = self._l.kg_last_error(self.h).decode()
        if rc == KG_ERR_INVALID_ARGUMENT:
            raise ValueError(msg)
        if rc == KG_ERR_IO:
            raise IOError(msg)
        raise RuntimeError(msg)

    def call(self, name, *args):
        self.check(getattr(self._l, name)(self.h, *args))


def f32c(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def u32c(a):
    return np.ascontiguousarray(a, dtype=np.uint32)
