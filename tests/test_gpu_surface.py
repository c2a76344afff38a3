"""Chip-ordered surface export (fmdbzp_plan_set_surface_layout) on the GPU.

The device reduces the accumulator to (ndop, nbins) in ascending code phase;
the host restatement (`acquisition.surface.pool_chip_sorted`) does the same
from the full surface of the same plan, so the two must be bit-identical
(both take maxima of the same FP32 values). Covered: a row-major accumulator
(GPS L1 C/A 16.368 MHz int16, N_nc 2) and a column-major one with 25
rotation slices (E1-like, ndop 2000), at nbins = spc (the reference's
chip-sorted surface itself) and at coarser binnings.
"""
import numpy as np
import pytest

from acquisition import E1Like, acquire_, plan_acquire
from acquisition.signals import e1like_scene, l1ca_if_int16_scene
from acquisition.surface import fmdbzp_sort_by_chip, pool_chip_sorted, set_power_bins_layout

pytestmark = pytest.mark.gpu


def _check(plan, signal, prns, if_hz, bins_list):
    full = [np.array(r.power_bins, copy=True)
            for r in acquire_(plan, signal, prns, interm_freq=if_hz, store_power_bins=True)]
    for nb in bins_list:
        set_power_bins_layout(plan, nb)
        res = acquire_(plan, signal, prns, interm_freq=if_hz, store_power_bins=True)
        for r, f in zip(res, full):
            assert r.power_bins.shape == (plan.num_doppler_bins, nb)
            np.testing.assert_array_equal(r.power_bins, pool_chip_sorted(f, plan, nb))
        assert len(plan.power_bins_chip_axis) == nb and np.all(np.diff(plan.power_bins_chip_axis) > 0)
        print(f"nbins {nb}: equal to the host restatement")
    set_power_bins_layout(plan, 0)
    back = acquire_(plan, signal, prns, interm_freq=if_hz, store_power_bins=True)
    for r, f in zip(back, full):
        np.testing.assert_array_equal(r.power_bins, f)
    return full


def test_chip_surface_row_major_accumulator():
    sc = l1ca_if_int16_scene(n_ms=20, npresent=2, cn0_range=(44.0, 45.0))
    prns = sc.present[:1] + [p for p in range(1, 33) if p not in sc.present][:1]
    plan = plan_acquire(sc.system, sc.sampling_freq, prns, num_coherently_integrated_code_periods=10,
                        num_noncoherent_accumulations=2)
    try:
        full = _check(plan, sc.signal, prns, sc.interm_freq, [plan.samples_per_code, 2046, 97])
        # nbins = spc with one slice is the reference's chip-sorted surface
        _, sorted_ref = fmdbzp_sort_by_chip(full[0], plan.num_blocks, plan.block_size,
                                            sc.system.get_code_frequency(), sc.sampling_freq)
        np.testing.assert_array_equal(pool_chip_sorted(full[0], plan, plan.samples_per_code), sorted_ref)
    finally:
        plan.close()


def test_chip_surface_rotation_slices_column_major():
    sysm = E1Like()
    sc = e1like_scene(n_periods=25, nnc=2, fs=1e6, nprn=4, npresent=2, cn0_range=(46.0, 48.0), system=sysm)
    prns = sc.present[:1]
    plan = plan_acquire(sysm, sc.sampling_freq, prns, min_doppler_coverage=9000.0,
                        num_coherently_integrated_code_periods=25, num_noncoherent_accumulations=2)
    try:
        assert plan.num_secondary_rotations == 25 and plan.num_doppler_bins == 2000
        _check(plan, sc.signal, prns, 0.0, [plan.samples_per_code, 500])
    finally:
        plan.close()
