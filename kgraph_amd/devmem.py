"""Minimal device-memory plumbing over the HIP runtime (hipMalloc / hipMemcpy /
hipFree), for callers that keep inputs resident in HBM and call the
kg_*_device entry points."""
import ctypes

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        h = ctypes.CDLL("libamdhip64.so")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
        h.hipDeviceSynchronize.argtypes = []
        h.hipGetErrorString.restype = ctypes.c_char_p
        _hip = h
    return _hip


def _check(rc, what):
    if rc != 0:
        raise RuntimeError("%s failed: %s" % (what, hip().hipGetErrorString(rc).decode()))


H2D, D2H = 1, 2


class DeviceArray:
    """A device allocation with a numpy dtype/shape for copies in and out."""

    def __init__(self, shape, dtype):
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = ctypes.c_void_p()
        _check(hip().hipMalloc(ctypes.byref(self.ptr), max(self.nbytes, 4)), "hipMalloc")

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        _check(hip().hipMemcpy(d.ptr, a.ctypes.data, d.nbytes, H2D), "hipMemcpy")
        return d

    def to_host(self):
        out = np.empty(self.shape, dtype=self.dtype)
        _check(hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H), "hipMemcpy")
        return out

    def zero(self):
        _check(hip().hipMemset(self.ptr, 0, self.nbytes), "hipMemset")

    def __del__(self):
        p, self.ptr = getattr(self, "ptr", None), None
        if p is not None and p.value:
            hip().hipFree(p)


def synchronize():
    _check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
