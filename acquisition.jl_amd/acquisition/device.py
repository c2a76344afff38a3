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
        self._alloc(a.shape, a.dtype)
        rc = hip().hipMemcpy(ctypes.c_void_p(self.ptr), ctypes.c_void_p(a.ctypes.data), self.nbytes, _H2D)
        if rc != 0:
            self.free()
            raise _native.EngineError(f"hipMemcpy H2D failed with {rc}")

    def _alloc(self, shape, dtype):
        self.ptr = None
        self.shape = tuple(int(d) for d in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        rc = hip().hipMalloc(ctypes.byref(p), max(self.nbytes, 1))
        if rc != 0:
            raise _native.EngineError(f"hipMalloc({self.nbytes}) failed with {rc}")
        self.ptr = p.value

    @classmethod
    def empty(cls, shape, dtype) -> "DeviceArray":
        """An uninitialised device buffer of `shape` and `dtype`."""
        self = cls.__new__(cls)
        self._alloc(shape, dtype)
        return self

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


def synthesize(n_samples: int, sampling_freq: float, complex_samples: bool, out_dtype, noise_scale: float,
               seed: int, noise_stream: int, emitters) -> DeviceArray:
    """Generate a seeded scene directly in device memory (`fmdbzp_synthesize`).

    `emitters` are `acquisition.signals.Emitter`s; the samples equal those of
    `acquisition.signals.render_host` for the same arguments, up to the last
    bit of the device's log / sqrt / cos / sin (one output unit in the last
    place, for rare samples). `out_dtype`: complex64 for complex scenes,
    int16 (rounded half to even, saturated) or float32 for real ones."""
    dt = np.dtype(out_dtype)
    elt = {np.dtype(np.complex64): "cf32", np.dtype(np.int16): "i16", np.dtype(np.float32): "f32"}.get(dt)
    if elt is None:
        raise ValueError(f"unsupported output dtype {dt}")
    keep = []
    arr = (_native.Emitter * max(len(emitters), 1))()
    for i, em in enumerate(emitters):
        e = arr[i]
        chips = np.ascontiguousarray(em.chips, dtype=np.int8)
        keep.append(chips)
        e.chips = chips.ctypes.data
        e.code_length = len(chips)
        e.boc = 1 if em.boc else 0
        if em.secondary is not None:
            sec = np.ascontiguousarray(em.secondary, dtype=np.int8)
            keep.append(sec)
            e.secondary = sec.ctypes.data
            e.secondary_length = len(sec)
            e.secondary_phase = int(em.secondary_phase)
        if em.bits is not None:
            bits = np.ascontiguousarray(em.bits, dtype=np.int8)
            keep.append(bits)
            e.bits = bits.ctypes.data
            e.num_bits = len(bits)
            e.bit_ms = int(em.bit_ms)
            e.bit_edge_ms = int(em.bit_edge_ms)
        e.amplitude = float(em.amplitude)
        e.code_freq_hz = float(em.code_freq)
        e.code_phase_chips = float(em.code_phase)
        e.carrier_hz = float(em.carrier_hz)
        e.carrier_phase = float(em.carrier_phase)
    sd = _native.SceneDesc(n_samples=int(n_samples), sampling_freq_hz=float(sampling_freq),
                           complex_samples=1 if complex_samples else 0, out_eltype=_native.ELTYPE[elt],
                           noise_scale=float(noise_scale), seed=int(seed) & 0xFFFFFFFFFFFFFFFF,
                           noise_stream=int(noise_stream) & 0xFFFFFFFF)
    out = DeviceArray.empty((int(n_samples),), dt)
    try:
        _native._check(_native.lib().fmdbzp_synthesize(ctypes.byref(sd), ctypes.cast(arr, ctypes.c_void_p),
                                                        len(emitters), ctypes.c_void_p(out.ptr)),
                       "fmdbzp_synthesize")
    except Exception:
        out.free()
        raise
    return out
