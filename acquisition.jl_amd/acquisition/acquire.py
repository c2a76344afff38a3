"""`acquire_` / `acquire`: run the engine and assemble `AcquisitionResults`.

Mirrors src/acquire.jl:54-882 and src/Acquisition.jl:81-142. The search
(downconversion through per-PRN peak, noise and neighbour cells) runs in
libfmdbzp on the GPU; this module validates arguments with the reference's
messages, and evaluates the reference's scalar result formulas with its exact
Float32/Float64 mix (peak-to-noise, CN0, code phase, parabolic interpolation,
the secondary-code phase gate and estimator).
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass
from typing import Any, Optional, Sequence, Union

import numpy as np

from . import _native
from .cfar import cfar_threshold
from .plan import AcquisitionPlan, ArgumentError, plan_acquire

f32 = np.float32


@dataclass(slots=True)
class AcquisitionResults:
    """Results of one PRN (src/Acquisition.jl:81-109). Frequencies in Hz.
    `engine_stats` exposes the engine's per-PRN statistics behind the result
    (peak, row, col, noise, code and Doppler neighbours, noise sum and cells)."""

    system: Any
    prn: int
    sampling_frequency: float
    carrier_doppler: float
    code_phase: float
    secondary_code_phase: Optional[int]
    CN0: float
    noise_power: np.float32
    peak_to_noise_ratio: np.float32
    num_noncoherent_integrations: int
    power_bins: Optional[np.ndarray]
    dopplers: np.ndarray
    num_blocks: int
    block_size: int
    num_secondary_rotations: int
    _stats: Optional[tuple] = None  # (peak, row, col, noise, nbr[4], noise_sum, noise_cells)

    @property
    def engine_stats(self) -> Optional[dict]:
        if self._stats is None:
            return None
        peak, row, col, noise, nbr, noise_sum, noise_cells = self._stats
        return dict(peak=peak, row=row, col=col, noise=noise, nbr=list(nbr), noise_sum=noise_sum,
                    noise_cells=noise_cells)

    def __repr__(self):
        return (f"AcquisitionResults: PRN {self.prn}, CN0 = {round(self.CN0, 2)} dB-Hz, "
                f"Doppler = {self.carrier_doppler} Hz, Code phase = {round(self.code_phase, 3)} chips")


def get_num_cells(result: AcquisitionResults) -> int:
    """Search cells: Doppler bins x samples per code x secondary rotations
    (src/Acquisition.jl:124-126)."""
    return len(result.dopplers) * result.num_blocks * result.block_size * result.num_secondary_rotations


def is_detected(result: AcquisitionResults, pfa: float = 0.01) -> bool:
    """peak_to_noise_ratio > cfar_threshold(pfa, get_num_cells) (src/Acquisition.jl:135-142)."""
    thr = cfar_threshold(pfa, get_num_cells(result),
                         num_noncoherent_integrations=result.num_noncoherent_integrations)
    return bool(float(result.peak_to_noise_ratio) > thr)


def fmdbzp_column_to_tau(scrambled_col: int, num_blocks: int, block_size: int) -> int:
    """Scrambled FM-DBZP column (0-based) -> delay in samples (src/plot.jl:16-20)."""
    block_row = scrambled_col // block_size
    within = scrambled_col % block_size
    return ((num_blocks - block_row) % num_blocks) * block_size + within


def jl_mod(x: float, y: float) -> float:
    """Julia `mod` for Float64: fmod, shifted into the sign of y."""
    r = math.fmod(x, y)
    if r == 0.0:
        return math.copysign(0.0, y)
    if (r > 0) != (y > 0):
        return r + y
    return r


def parabolic_interp(left, peak, right):
    """`_parabolic_interp` (src/acquire.jl:9-13) on Float32 operands: returns a
    float32 offset, or the Float64 0.0 the reference returns for a zero
    denominator."""
    left, peak, right = f32(left), f32(peak), f32(right)
    denom = f32(2) * f32(f32(f32(f32(2) * peak) - left) - right)
    if denom == 0:
        return 0.0
    return f32(f32(right - left) / denom)


# ---------------------------------------------------------------------------
# signal views
# ---------------------------------------------------------------------------
def _is_torch(x) -> bool:
    return type(x).__module__.startswith("torch") and hasattr(x, "data_ptr")


def _numpy_eltype(a: np.ndarray):
    if a.dtype == np.complex64 and a.ndim == 1:
        return "cf32", a.shape[0]
    if a.dtype == np.complex128 and a.ndim == 1:
        return "cf64", a.shape[0]
    if a.dtype == np.int16 and a.ndim == 1:
        return "i16", a.shape[0]
    if a.dtype == np.int16 and a.ndim == 2 and a.shape[1] == 2:
        return "ci16", a.shape[0]
    if a.dtype == np.int8 and a.ndim == 2 and a.shape[1] == 2:
        return "ci8", a.shape[0]
    if a.dtype == np.float32 and a.ndim == 1:
        return "f32", a.shape[0]
    if a.dtype == np.float64 and a.ndim == 1:
        return "f64", a.shape[0]
    raise ArgumentError(f"unsupported signal array dtype {a.dtype} with shape {a.shape}")


class Packed2BitSignal:
    """Packed 2-bit complex SDR samples (FMDBZP_C2BIT in include/fmdbzp.h): byte
    i holds samples 2i (low nibble) and 2i+1 (high nibble); a nibble is I in
    bits 0-1 and Q in bits 2-3, code c -> (c & 1 ? 3 : 1) * (c & 2 ? -1 : 1).
    `packed` is a host uint8 array or a `device.DeviceArray` of those bytes."""

    def __init__(self, packed, n: int):
        self.packed = packed
        self.n = int(n)

    @staticmethod
    def pack(values: np.ndarray) -> "Packed2BitSignal":
        """Pack an (n, 2) array of I/Q values in {+1, +3, -1, -3}."""
        v = np.asarray(values)
        code = ((np.abs(v) == 3).astype(np.uint8) | ((v < 0).astype(np.uint8) << 1))
        if not np.all(np.isin(np.abs(v), (1, 3))):
            raise ArgumentError("2-bit samples must be +1, +3, -1 or -3")
        nib = code[:, 0] | (code[:, 1] << 2)
        n = nib.shape[0]
        if n % 2:
            nib = np.concatenate([nib, np.zeros(1, dtype=np.uint8)])
        packed = (nib[0::2] | (nib[1::2] << 4)).astype(np.uint8)
        return Packed2BitSignal(packed, n)

    def unpack(self, count: Optional[int] = None) -> np.ndarray:
        """Complex64 values of the first `count` samples."""
        from .device import DeviceArray

        n = self.n if count is None else min(count, self.n)
        raw = self.packed.to_host((n + 1) // 2) if isinstance(self.packed, DeviceArray) else \
            np.asarray(self.packed[:(n + 1) // 2])
        nib = np.empty(2 * raw.shape[0], dtype=np.uint8)
        nib[0::2] = raw & 15
        nib[1::2] = raw >> 4
        nib = nib[:n]

        def val(c):
            return np.where(c & 1, 3.0, 1.0) * np.where(c & 2, -1.0, 1.0)

        return (val(nib & 3) + 1j * val((nib >> 2) & 3)).astype(np.complex64)


class _SignalView:
    """Pointer, sample format and length of a host or device signal."""

    def __init__(self, signal):
        from .device import DeviceArray

        self.obj = signal
        if isinstance(signal, Packed2BitSignal):
            self.kind, self.n = "c2bit", signal.n
            self.device = isinstance(signal.packed, DeviceArray)
            self.ptr = signal.packed.ptr if self.device else np.ascontiguousarray(signal.packed).ctypes.data
            if not self.device:
                self.obj = Packed2BitSignal(np.ascontiguousarray(signal.packed), signal.n)
                self.ptr = self.obj.packed.ctypes.data
        elif isinstance(signal, DeviceArray):
            self.kind, self.n = _numpy_eltype(np.empty((0,) + tuple(signal.shape[1:]), dtype=signal.dtype))
            self.n = signal.shape[0]
            self.device = True
            self.ptr = signal.ptr
        elif _is_torch(signal):
            import torch

            t = signal
            if not t.is_contiguous():
                raise ArgumentError("device signal must be contiguous")
            m = {torch.complex64: "cf32", torch.complex128: "cf64", torch.float32: "f32", torch.float64: "f64"}
            if t.dtype in m and t.dim() == 1:
                self.kind, self.n = m[t.dtype], t.shape[0]
            elif t.dtype == torch.int16 and t.dim() == 1:
                self.kind, self.n = "i16", t.shape[0]
            elif t.dtype == torch.int16 and t.dim() == 2 and t.shape[1] == 2:
                self.kind, self.n = "ci16", t.shape[0]
            elif t.dtype == torch.int8 and t.dim() == 2 and t.shape[1] == 2:
                self.kind, self.n = "ci8", t.shape[0]
            else:
                raise ArgumentError(f"unsupported signal tensor dtype {t.dtype} with shape {tuple(t.shape)}")
            self.device = t.is_cuda
            if self.device:
                torch.cuda.current_stream(t.device).synchronize()
            else:
                self.obj = np.ascontiguousarray(t.numpy())
            self.ptr = t.data_ptr() if self.device else self.obj.ctypes.data
        else:
            a = np.ascontiguousarray(signal)
            self.obj = a
            self.kind, self.n = _numpy_eltype(a)
            self.device = False
            self.ptr = a.ctypes.data


# ---------------------------------------------------------------------------
# acquire_
# ---------------------------------------------------------------------------
def _interpolate(plan: AcquisitionPlan, peak, noise, code_nbrs, dop_nbrs, delay: int, code_phase: float,
                 doppler: float):
    """Subsample interpolation of one peak (src/acquire.jl:593-619, :714-748): a parabola
    through the code (Doppler) neighbours when the larger one exceeds sqrt(noise)."""
    fs = plan.sampling_freq
    fc = plan.system.get_code_frequency()
    code_length = plan.system.get_code_length()
    pl, pr = f32(code_nbrs[0]), f32(code_nbrs[1])
    dl, dr = f32(dop_nbrs[0]), f32(dop_nbrs[1])
    sq = np.sqrt(noise)
    if max(pl, pr) > sq:
        frac = parabolic_interp(pl, peak, pr)
        if isinstance(frac, np.float32):
            di = f32(f32(delay) + frac)
            code_phase = jl_mod(float(-di) * fc / fs, code_length)
    if max(dl, dr) > sq:
        frac = parabolic_interp(dl, peak, dr)
        doppler = doppler + float(frac) * plan.doppler_spacing_hz
    return code_phase, doppler


def _get_result_buffer(plan: AcquisitionPlan, prn_idx: int) -> np.ndarray:
    """Plan-owned (ndop, spc_eff) Float32 surface, allocated on first use and
    reused (src/acquire.jl:221-228)."""
    ncols = getattr(plan, "power_bins_chip_bins", 0) or plan.samples_per_code_eff
    buf = plan.result_buffers[prn_idx]
    if buf is None or buf.shape[1] != ncols:
        buf = np.empty((plan.num_doppler_bins, ncols), dtype=np.float32, order="F")
        plan.result_buffers[prn_idx] = buf
    return buf


def acquire_(plan: AcquisitionPlan, signal, prns: Union[int, Sequence[int]], *, interm_freq: float = 0.0,
             subsample_interpolation: bool = False, store_power_bins: bool = False):
    """FM-DBZP acquisition with a pre-computed plan (`acquire!`, src/acquire.jl:54-89).

    `signal` is a 1-D numpy array (complex64, complex128, int16 real,
    float32/float64 real), an (n, 2) int8/int16 array of complex integer
    samples, the same in device memory (`device.DeviceArray`, read in place),
    or a torch tensor (a CUDA tensor is read in place; torch must then be
    imported before this package so one HIP runtime serves both).
    Returns one `AcquisitionResults` per PRN (a single result for an int).
    """
    if isinstance(prns, (int, np.integer)):
        return acquire_(plan, signal, [int(prns)], interm_freq=interm_freq,
                        subsample_interpolation=subsample_interpolation, store_power_bins=store_power_bins)[0]
    prns, sv, idx, flags, bufs, surfaces = _prepare(plan, signal, prns, interm_freq, store_power_bins)
    if not prns:
        plan.acq_results_buf = []
        return plan.acq_results_buf
    stats = plan.native.acquire(sv.ptr, _native.ELTYPE[sv.kind], sv.n, float(interm_freq), idx, flags, surfaces)
    results = _results_from_stats(plan, stats, prns, sv, interm_freq, subsample_interpolation, bufs)
    plan.acq_results_buf = results
    return results


def _prepare(plan: AcquisitionPlan, signal, prns, interm_freq, store_power_bins):
    """Validation shared by `acquire_` and `acquire_batch` (src/acquire.jl:62-72)."""
    prns = [int(p) for p in prns]
    if not all(p in plan.avail_prns for p in prns):
        raise ArgumentError(f"All requested PRNs must be in plan.avail_prns. Got: {prns}, "
                            f"available: {plan.avail_prns}")
    if plan.native is None:
        raise _native.EngineError("plan has no device plan (closed, or built without plan_acquire)")
    sv = _SignalView(signal)
    segment_length = plan.num_coherently_integrated_code_periods * plan.samples_per_code
    num_segments = sv.n // segment_length
    if num_segments < plan.num_noncoherent_accumulations:
        raise ArgumentError(
            f"Signal has {sv.n} samples → {num_segments} full segments of {segment_length}, "
            f"but plan.num_noncoherent_accumulations={plan.num_noncoherent_accumulations}. "
            "Provide a longer signal.")
    idx = [plan.avail_prns.index(p) for p in prns]
    flags = _native.FLAG_DEVICE_SIGNAL if sv.device else 0
    bufs = surfaces = None
    if store_power_bins and prns:
        flags |= _native.FLAG_STORE_SURFACES
        bufs = [_get_result_buffer(plan, i) for i in idx]
        surfaces = [b.ctypes.data for b in bufs]
    return prns, sv, idx, flags, bufs, surfaces


# numpy view of fmdbzp_prn_stats (the ctypes PrnStats layout), built once
_STATS_DTYPE = np.dtype({"names": [f[0] for f in _native.PrnStats._fields_],
                         "formats": [np.dtype(f[1]) for f in _native.PrnStats._fields_],
                         "offsets": [getattr(_native.PrnStats, f[0]).offset for f in _native.PrnStats._fields_],
                         "itemsize": ctypes.sizeof(_native.PrnStats)})


def _results_from_stats(plan, stats, prns, sv, interm_freq, subsample_interpolation, bufs):
    """Result extraction (src/acquire.jl:544-671 and :681-781) from the
    engine's per-PRN statistics, all PRNs of the call at once. The Float32
    steps (signal power, peak-to-noise ratio, the CN0 ratio) are numpy float32
    array operations, each rounded to Float32 as the reference's scalar code
    does; the code phase and Doppler follow in Float64 per PRN."""
    n = len(prns)
    a = np.frombuffer(stats, dtype=_STATS_DTYPE, count=n)
    system = plan.system
    fs = plan.sampling_freq
    fc = system.get_code_frequency()
    code_length = system.get_code_length()
    spc, nb, bsz = plan.samples_per_code, plan.num_blocks, plan.block_size
    peak = np.ascontiguousarray(a["peak_power"], dtype=np.float32)
    noise = np.ascontiguousarray(a["noise_power"], dtype=np.float32)
    with np.errstate(divide="ignore", invalid="ignore"):
        signal_power = peak - noise
        ptn = (signal_power + noise) / noise
        ratio = (signal_power / noise).tolist()
    code_period = code_length / fc
    tcoh = plan.num_coherently_integrated_code_periods * code_period
    rows, cols = a["peak_row"].tolist(), a["peak_col"].tolist()
    peaks, noises = peak.tolist(), noise.tolist()
    nbrs = np.stack([a["code_left"], a["code_right"], a["dop_left"], a["dop_right"]], axis=1).tolist()
    nsums, ncells = a["noise_sum"].tolist(), a["noise_cells"].tolist()
    # the secondary-code phase of a detected PRN (src/acquire.jl:646-652) is
    # estimated for all such PRNs in one device call below
    R = plan.num_secondary_rotations
    thr = None
    if R > 1:
        thr = cfar_threshold(0.01, plan.num_doppler_bins * spc * R,
                             num_noncoherent_integrations=plan.num_noncoherent_accumulations)
    nnc, dfreqs = plan.num_noncoherent_accumulations, plan.doppler_freqs
    results, pending = [], []
    for j, p in enumerate(prns):
        rj = ratio[j]
        cn0 = 10 * math.log10(rj / tcoh / 1.0) if rj > 0 else (-math.inf if rj == 0 else math.nan)
        row, col = rows[j], cols[j]
        delay = fmdbzp_column_to_tau(col % spc, nb, bsz)
        code_phase = jl_mod(-delay * fc / fs, code_length)
        doppler = float(plan.doppler_freqs[row])
        if subsample_interpolation:
            code_phase, doppler = _interpolate(plan, peak[j], noise[j], nbrs[j][0:2], nbrs[j][2:4], delay, code_phase,
                                               doppler)
        # positional: the field order of AcquisitionResults
        r = AcquisitionResults(system, p, fs, doppler, code_phase, None, cn0, noise[j], ptn[j], nnc,
                               bufs[j] if bufs is not None else None, dfreqs, nb, bsz, R,
                               (peaks[j], row, col, noises[j], nbrs[j], nsums[j], ncells[j]))
        if thr is not None and float(ptn[j]) > thr:
            pending.append(r)
        results.append(r)
    if pending:
        # `_estimate_secondary_code_phase` (src/acquire.jl:503-538) on the device
        sysm = plan.system
        items = [(sysm.gen_code(spc, r.prn, fs, sysm.get_code_frequency(), r.code_phase),
                  np.array([sysm.secondary_value(r.prn, k) for k in range(R)], dtype=np.int8), r.carrier_doppler)
                 for r in pending]
        flags = _native.FLAG_DEVICE_SIGNAL if sv.device else 0
        phases = plan.native.secondary_phase(sv.ptr, _native.ELTYPE[sv.kind], sv.n, float(interm_freq), flags, items)
        for r, ph in zip(pending, phases):
            r.secondary_code_phase = ph
    return results


def acquire_batch(requests):
    """Run several acquisitions, possibly on different plans, as one submission
    (`fmdbzp_acquire_batch`): requests on different plans overlap on the GPU.
    Each request is (plan, signal, prns) or (plan, signal, prns, kwargs) with
    the keywords of `acquire_`; returns one result list per request."""
    prepared, natives, keep = [], [], []
    for req in requests:
        plan, signal, prns = req[0], req[1], req[2]
        kw = dict(req[3]) if len(req) > 3 else {}
        interm_freq = float(kw.get("interm_freq", 0.0))
        interp = bool(kw.get("subsample_interpolation", False))
        store = bool(kw.get("store_power_bins", False))
        if isinstance(prns, (int, np.integer)):
            prns = [int(prns)]
        prns, sv, idx, flags, bufs, surfaces = _prepare(plan, signal, prns, interm_freq, store)
        q, held = plan.native.request(sv.ptr, _native.ELTYPE[sv.kind], sv.n, interm_freq, idx, flags, surfaces)
        natives.append(q)
        keep.append((held, sv))
        prepared.append((plan, prns, sv, interm_freq, interp, bufs, held[1]))
    _native.acquire_batch(natives)
    out = []
    for plan, prns, sv, interm_freq, interp, bufs, stats in prepared:
        res = _results_from_stats(plan, stats, prns, sv, interm_freq, interp, bufs)
        plan.acq_results_buf = res
        out.append(res)
    return out


def acquire(system, signal, sampling_freq: float, prns, *, interm_freq: float = 0.0,
            min_doppler_coverage: float = 7000.0, num_coherently_integrated_code_periods: int = 1,
            bit_edge_search_steps: int = 1, num_noncoherent_accumulations: int = 1,
            use_secondary_code: bool = True, max_secondary_code_rotations: int = 32, fft_flag=None,
            subsample_interpolation: bool = False, store_power_bins: bool = False, **kw):
    """`plan_acquire` then `acquire_` (src/acquire.jl:838-882)."""
    single = isinstance(prns, (int, np.integer))
    plist = [int(prns)] if single else [int(p) for p in prns]
    plan = plan_acquire(system, sampling_freq, plist, min_doppler_coverage=min_doppler_coverage,
                        num_coherently_integrated_code_periods=num_coherently_integrated_code_periods,
                        bit_edge_search_steps=bit_edge_search_steps,
                        num_noncoherent_accumulations=num_noncoherent_accumulations,
                        use_secondary_code=use_secondary_code,
                        max_secondary_code_rotations=max_secondary_code_rotations, **kw)
    try:
        res = acquire_(plan, signal, plist, interm_freq=interm_freq,
                       subsample_interpolation=subsample_interpolation, store_power_bins=store_power_bins)
    finally:
        plan.close()
    return res[0] if single else res
