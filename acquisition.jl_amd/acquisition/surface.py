"""Power-surface layouts for plotting and analysis.

The reference's plot recipe (src/plot.jl:35-57) sorts the FM-DBZP columns of
`power_bins` into ascending code phase. `fmdbzp_sort_by_chip` restates it for
a stored surface. `set_power_bins_layout` makes the engine produce the
chip-ordered surface on the device instead (`fmdbzp_plan_set_surface_layout`),
optionally max-pooled into fewer code-phase bins, so only that smaller
surface crosses the bus: the 25 MHz E1-like surface is 20 GB per PRN, and its
two-bins-per-chip form is 65 MB.
"""
from __future__ import annotations

import numpy as np


def julia_mod(x, y: float) -> np.ndarray:
    """Julia's floored `mod` for floats: the result has the sign of y, and a
    zero result is copysign(0, y)."""
    x = np.asarray(x, dtype=np.float64)
    r = np.fmod(x, y)
    r = np.where((r != 0) & ((r < 0) != (y < 0)), r + y, r)
    return np.where(r == 0, np.copysign(0.0, y), r)


def chip_phases(num_blocks: int, block_size: int, code_freq_hz: float, sampling_freq_hz: float) -> np.ndarray:
    """Code phase in chips of every scrambled column, `mod(-tau f_code / fs,
    spc f_code / fs)` with tau the column's delay (src/plot.jl:45-52)."""
    spc = num_blocks * block_size
    code_length = spc * code_freq_hz / sampling_freq_hz
    c = np.arange(spc, dtype=np.int64)
    tau = (((num_blocks - c // block_size) % num_blocks) * block_size + c % block_size).astype(np.float64)
    return julia_mod(-tau * code_freq_hz / sampling_freq_hz, code_length)


def chip_order(num_blocks: int, block_size: int, code_freq_hz: float, sampling_freq_hz: float) -> np.ndarray:
    """The columns in ascending code phase (a stable `sortperm`)."""
    return np.argsort(chip_phases(num_blocks, block_size, code_freq_hz, sampling_freq_hz), kind="stable")


def fmdbzp_sort_by_chip(power_bins: np.ndarray, num_blocks: int, block_size: int, code_freq_hz: float,
                        sampling_freq_hz: float):
    """`_fmdbzp_sort_by_chip` (src/plot.jl:35-57): (chip_axis, chip-sorted
    power_bins) of an (ndop, spc) surface."""
    ph = chip_phases(num_blocks, block_size, code_freq_hz, sampling_freq_hz)
    order = np.argsort(ph, kind="stable")
    return ph[order], power_bins[:, order]


def set_power_bins_layout(plan, chip_bins: int = 0) -> None:
    """Surfaces that `acquire_(..., store_power_bins=True)` returns from now on:
    chip_bins = 0 restores the reference's (ndop, spc_eff) scrambled layout;
    chip_bins in [1, spc] gives (ndop, chip_bins) surfaces in ascending code
    phase, bin b the maximum over the chip-sorted columns
    [b spc / chip_bins, (b + 1) spc / chip_bins) and over the rotation slices.
    `plan.power_bins_chip_axis` then holds each bin's lowest code phase."""
    chip_bins = int(chip_bins)
    if chip_bins == 0:
        plan.native.set_surface_layout(0)
        plan.power_bins_chip_bins = 0
        plan.power_bins_chip_axis = None
        return
    code_freq = plan.system.get_code_frequency()
    ph = chip_phases(plan.num_blocks, plan.block_size, code_freq, plan.sampling_freq)
    order = np.argsort(ph, kind="stable")
    plan.native.set_surface_layout(chip_bins, order.astype(np.int32))
    spc = plan.samples_per_code
    starts = (np.arange(chip_bins, dtype=np.int64) * spc) // chip_bins
    plan.power_bins_chip_bins = chip_bins
    plan.power_bins_chip_axis = ph[order][starts]


def pool_chip_sorted(power_bins: np.ndarray, plan, chip_bins: int) -> np.ndarray:
    """Host restatement of the device's chip-ordered surface from a full
    (ndop, spc_eff) surface of the same plan: chip sort, then the maximum
    over each bin's columns and over the rotation slices."""
    spc, R = plan.samples_per_code, plan.num_secondary_rotations
    code_freq = plan.system.get_code_frequency()
    order = chip_order(plan.num_blocks, plan.block_size, code_freq, plan.sampling_freq)
    per_slice = np.stack([power_bins[:, s * spc:(s + 1) * spc][:, order] for s in range(R)])
    cols = per_slice.max(axis=0)
    out = np.empty((power_bins.shape[0], chip_bins), dtype=power_bins.dtype)
    for b in range(chip_bins):
        i0, i1 = (b * spc) // chip_bins, ((b + 1) * spc) // chip_bins
        out[:, b] = cols[:, i0:i1].max(axis=1)
    return out
