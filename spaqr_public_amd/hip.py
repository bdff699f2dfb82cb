"""Minimal HIP runtime access through ctypes (the same libamdhip64 that
libspaqr.so links), for device buffers and events without a framework."""
from __future__ import annotations

import ctypes as C

import numpy as np

_hip = None


def runtime():
    global _hip
    if _hip is None:
        from . import core
        core.load_library()  # resolves libamdhip64.so.7 through libspaqr's dependencies
        _hip = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
        _hip.hipGetErrorString.restype = C.c_char_p
        _hip.hipGetErrorString.argtypes = [C.c_int]
    return _hip


def check(rc):
    if rc != 0:
        raise RuntimeError("HIP error: " + runtime().hipGetErrorString(rc).decode())


H2D, D2H, D2D = 1, 2, 3


class DeviceArray:
    """A float64 device vector."""

    def __init__(self, n: int):
        self.n = int(n)
        p = C.c_void_p()
        check(runtime().hipMalloc(C.byref(p), max(8 * self.n, 8)))
        self.ptr = p.value

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        d = cls(a.size)
        check(runtime().hipMemcpy(C.c_void_p(d.ptr), a.ctypes.data_as(C.c_void_p), 8 * a.size, H2D))
        return d

    def to_host(self):
        out = np.zeros(self.n)
        check(runtime().hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), 8 * self.n, D2H))
        return out

    def copy_from(self, other: "DeviceArray"):
        check(runtime().hipMemcpy(C.c_void_p(self.ptr), C.c_void_p(other.ptr), 8 * self.n, D2D))

    def free(self):
        if self.ptr:
            runtime().hipFree(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    check(runtime().hipDeviceSynchronize())
