"""Chip-ordered power surfaces (CPU): the restatement of the reference's
`_fmdbzp_sort_by_chip` (src/plot.jl:35-57) against the properties its own
tests check (test/plot.jl:44-75) and against the column-to-delay definition,
and the bin pooling that the device export reproduces."""
import math

import numpy as np

from acquisition import GPSL1CA, E1Like
from acquisition.acquire import fmdbzp_column_to_tau
from acquisition.plan import _plan_host
from acquisition.surface import chip_order, chip_phases, fmdbzp_sort_by_chip, julia_mod, pool_chip_sorted


def test_julia_mod_floats():
    assert julia_mod(-0.0, 3.0) == 0.0 and math.copysign(1.0, float(julia_mod(-0.0, 3.0))) == 1.0
    assert julia_mod(-1.0, 3.0) == 2.0
    assert julia_mod(5.5, 3.0) == 2.5
    assert julia_mod(-3.0, 3.0) == 0.0


def test_sort_by_chip_reference_properties():
    """test/plot.jl:44-75: L1 C/A at 2.048 MHz, +-5 kHz, N_coh 1."""
    sysm = GPSL1CA()
    fs = 2.048e6
    h = _plan_host(sysm, fs, [1], min_doppler_coverage=5000.0, num_coherently_integrated_code_periods=1)
    dummy = np.ones((h.num_doppler_bins, h.samples_per_code), dtype=np.float32)
    axis, ordered = fmdbzp_sort_by_chip(dummy, h.num_blocks, h.block_size, sysm.get_code_frequency(), fs)
    assert len(axis) == h.samples_per_code and ordered.shape == dummy.shape
    assert np.all(np.diff(axis) >= 0)
    code_length = h.samples_per_code * sysm.get_code_frequency() / fs
    assert axis.min() >= 0.0 and axis.max() < code_length
    assert len(np.unique(axis)) == h.samples_per_code


def test_chip_phase_is_the_delay_definition():
    """Each column's phase is mod(-tau f_code / fs, L) with tau from the
    column-to-delay map, and ascending phase means descending delay."""
    nb, bs, fc, fs = 16, 1023, 1.023e6, 16.368e6
    ph = chip_phases(nb, bs, fc, fs)
    L = nb * bs * fc / fs
    for c in (0, 1, 1022, 1023, 5000, nb * bs - 1):
        tau = fmdbzp_column_to_tau(c, nb, bs)
        assert ph[c] == (0.0 if tau == 0 else L - tau * fc / fs)
    order = chip_order(nb, bs, fc, fs)
    taus = np.array([fmdbzp_column_to_tau(int(c), nb, bs) for c in order])
    assert taus[0] == 0 and np.all(np.diff(taus[1:]) < 0)


class _P:
    pass


def test_pooling_matches_sort_and_max():
    rng = np.random.default_rng(3)
    p = _P()
    p.system = E1Like()
    p.num_blocks, p.block_size, p.sampling_freq = 4, 25, 0.5e6
    p.samples_per_code, p.num_secondary_rotations = 100, 3
    surf = rng.random((8, 300)).astype(np.float32)
    full = pool_chip_sorted(surf, p, 100)
    order = chip_order(4, 25, p.system.get_code_frequency(), 0.5e6)
    want = np.max(np.stack([surf[:, s * 100:(s + 1) * 100][:, order] for s in range(3)]), axis=0)
    np.testing.assert_array_equal(full, want)
    pooled = pool_chip_sorted(surf, p, 7)
    for b in range(7):
        i0, i1 = (b * 100) // 7, ((b + 1) * 100) // 7
        np.testing.assert_array_equal(pooled[:, b], want[:, i0:i1].max(axis=1))
