"""Device-resident sample buffers on the engine's own HIP runtime.

`DeviceArray(host_array)` copies a signal to HBM once so `acquire_` can run on
samples already resident in device memory (the engine reads them in place).
It calls hipMalloc/hipMemcpy/hipFree of the HIP runtime libfmdbzp.so is bound
to, so it never loads a second HIP runtime into the process.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native

_hip = None


def hip() -> ctypes.CDLL:
    global _hip
    if _hip is None:
        _native.lib()  # binds libamdhip64.so.7 first
        h = ctypes.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipMalloc.restype = ctypes.c_int
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipFree.restype = ctypes.c_int
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipMemcpy.restype = ctypes.c_int
        h.hipDeviceSynchronize.argtypes = []
        h.hipDeviceSynchronize.restype = ctypes.c_int
        _hip = h
    return _hip


_H2D, _D2H = 1, 2


class DeviceArray:
    """A host array's bytes in device memory, with its shape and dtype."""

    def __init__(self, host: np.ndarray):
        a = np.ascontiguousarray(host)
        self.shape = a.shape
        self.dtype = a.dtype
        self.nbytes = a.nbytes
        p = ctypes.c_void_p()
        rc = hip().hipMalloc(ctypes.byref(p), max(self.nbytes, 1))
        if rc != 0:
            raise _native.EngineError(f"hipMalloc({self.nbytes}) failed with {rc}")
        self.ptr = p.value
        rc = hip().hipMemcpy(ctypes.c_void_p(self.ptr), ctypes.c_void_p(a.ctypes.data), self.nbytes, _H2D)
        if rc != 0:
            self.free()
            raise _native.EngineError(f"hipMemcpy H2D failed with {rc}")

    @property
    def ndim(self):
        return len(self.shape)

    def to_host(self, count: int = None) -> np.ndarray:
        """First `count` rows (all when None) copied back to the host."""
        n = self.shape[0] if count is None else min(count, self.shape[0])
        out = np.empty((n,) + tuple(self.shape[1:]), dtype=self.dtype)
        if out.nbytes:
            rc = hip().hipMemcpy(ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(self.ptr), out.nbytes, _D2H)
            if rc != 0:
                raise _native.EngineError(f"hipMemcpy D2H failed with {rc}")
        return out

    def free(self):
        if getattr(self, "ptr", None):
            hip().hipFree(ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
