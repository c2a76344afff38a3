"""Host logic (no GPU): plan geometry and validation against the reference's
test/plan.jl values, the Doppler grid, sign patterns, drift shifts, CFAR
argument checks and the scalar result formulas."""
import math

import numpy as np
import pytest

from acquisition import GPSL1CA, GPSL5I, ArgumentError, cfar_threshold
from acquisition.acquire import (AcquisitionResults, fmdbzp_column_to_tau, get_num_cells, is_detected, jl_mod,
                                 parabolic_interp)
from acquisition.plan import _plan_host, julia_float_range, nextprod, sign_patterns
from acquisition.systems import E1Like, GNSSSystem, L5Like, samples_per_code


def test_geometry_l1ca_2048khz():
    """test/plan.jl:4-23."""
    p = _plan_host(GPSL1CA(), 2.048e6, [1], min_doppler_coverage=10000.0)
    assert (p.samples_per_code, p.num_blocks, p.block_size) == (2048, 32, 64)
    assert p.num_data_bits == 1 and p.bit_edge_search_steps == 1 and p.num_noncoherent_accumulations == 1
    assert len(p.doppler_freqs) == 32
    assert p.doppler_spacing_hz == pytest.approx(1000.0)
    assert p.doppler_freqs[0] == pytest.approx(-16000.0)
    assert p.doppler_freqs[-1] >= 10000.0 and p.doppler_freqs[0] <= -10000.0


def test_geometry_multi_ms():
    """test/plan.jl:25-40."""
    p = _plan_host(GPSL1CA(), 2.048e6, [1], min_doppler_coverage=10000.0,
                   num_coherently_integrated_code_periods=40, bit_edge_search_steps=4,
                   num_noncoherent_accumulations=5)
    assert (p.samples_per_code, p.num_blocks, p.num_data_bits) == (2048, 32, 2)
    assert len(p.doppler_freqs) == 1280 and p.doppler_spacing_hz == pytest.approx(25.0)


@pytest.mark.parametrize("fs,nb,bs", [(2.048e6, 16, 128), (5e6, 20, 250), (10e6, 16, 625), (16.368e6, 11, 1488),
                                      (25e6, 20, 1250)])
def test_smallest_divisor_num_blocks(fs, nb, bs):
    """test/plan.jl:194-220."""
    p = _plan_host(GPSL1CA(), fs, [1], min_doppler_coverage=5000.0, num_coherently_integrated_code_periods=5)
    assert (p.num_blocks, p.block_size) == (nb, bs)
    assert p.num_blocks * p.block_size == samples_per_code(GPSL1CA(), fs)


def test_secondary_code_configuration():
    """test/plan.jl:42-68."""
    p = _plan_host(GPSL1CA(), 2.048e6, [1], min_doppler_coverage=10000.0)
    assert p.use_secondary_code and p.max_secondary_code_rotations == 32 and p.num_secondary_rotations == 1
    p = _plan_host(GPSL5I(), 10.24e6, [1], min_doppler_coverage=5000.0, num_coherently_integrated_code_periods=10)
    assert p.num_secondary_rotations == 10 and p.sign_patterns_by_prn[1].shape == (10, 10)
    p = _plan_host(GPSL5I(), 10.24e6, [1], min_doppler_coverage=5000.0, num_coherently_integrated_code_periods=10,
                   use_secondary_code=False)
    assert p.num_secondary_rotations == 1 and not p.sign_patterns_by_prn and p.path == 0


def test_validation_errors():
    """test/plan.jl:70-168 (messages as the reference's)."""
    s = GPSL1CA()
    with pytest.raises(ArgumentError):
        _plan_host(s, 2.048e6, [1], num_coherently_integrated_code_periods=30)
    with pytest.raises(ArgumentError):
        _plan_host(s, 2.048e6, [1], num_coherently_integrated_code_periods=20, bit_edge_search_steps=3)
    with pytest.raises(ArgumentError):
        _plan_host(s, 2.048e6, [1], num_coherently_integrated_code_periods=1, bit_edge_search_steps=2)
    with pytest.raises(ArgumentError, match="num_noncoherent_accumulations must be >= 1, got 0"):
        _plan_host(s, 2.048e6, [1], num_noncoherent_accumulations=0)
    with pytest.raises(ArgumentError):
        _plan_host(s, 2.048e6, [1], num_coherently_integrated_code_periods=0)
    with pytest.raises(ArgumentError):
        _plan_host(s, 2.048e6, [1], bit_edge_search_steps=0)
    with pytest.raises(ArgumentError, match="No valid num_blocks found"):
        _plan_host(s, 2.048e6, [1], min_doppler_coverage=1.1e6)
    with pytest.raises(ArgumentError) as e:
        _plan_host(GPSL5I(), 10.24e6, [1], min_doppler_coverage=1000.0, num_coherently_integrated_code_periods=15)
    assert "secondary" in str(e.value).lower() and "15" in str(e.value) and "10" in str(e.value)
    with pytest.raises(ArgumentError) as e:
        _plan_host(GPSL5I(), 10.24e6, [1], min_doppler_coverage=1000.0, num_coherently_integrated_code_periods=5)
    m = str(e.value)
    for frag in ("secondary", "68", "num_coherently_integrated_code_periods", "num_noncoherent_accumulations",
                 "use_secondary_code"):
        assert frag in m
    assert _plan_host(GPSL5I(), 10.24e6, [1], min_doppler_coverage=1000.0,
                      num_coherently_integrated_code_periods=10).num_secondary_rotations == 10
    assert _plan_host(GPSL5I(), 10.24e6, [1], min_doppler_coverage=1000.0, num_coherently_integrated_code_periods=5,
                      use_secondary_code=False).num_secondary_rotations == 1
    l1cp = GNSSSystem("GPSL1C_P", 10230, 1.023e6, 1575.42e6, secondary=np.ones(1800, dtype=np.int8))
    with pytest.raises(ArgumentError) as e:
        _plan_host(l1cp, 25e6, [1], min_doppler_coverage=1000.0, num_coherently_integrated_code_periods=2)
    m = str(e.value)
    for frag in ("1800", "32", "=2", "GPSL1C_P", "num_coherently_integrated_code_periods",
                 "max_secondary_code_rotations", "use_secondary_code"):
        assert frag in m


def test_derived_geometry_table():
    """Geometry of the measured workloads (reference arithmetic restated)."""
    c1 = _plan_host(GPSL1CA(), 5e6, range(1, 33))
    assert (c1.samples_per_code, c1.num_blocks, c1.block_size, c1.fft_size, c1.num_doppler_bins) == \
        (5000, 20, 250, 500, 20)
    assert c1.doppler_freqs[0] == -10000.0 and c1.doppler_freqs[-1] == 9000.0
    c2 = _plan_host(GPSL1CA(), 16.368e6, range(1, 33), num_coherently_integrated_code_periods=10,
                    bit_edge_search_steps=10, num_noncoherent_accumulations=8)
    assert (c2.samples_per_code, c2.num_blocks, c2.block_size, c2.fft_size, c2.num_doppler_bins) == \
        (16368, 16, 1023, 2048, 160)
    assert c2.doppler_freqs[0] == -8000.0 and c2.doppler_freqs[-1] == 7900.0 and c2.path == 1
    c4 = _plan_host(E1Like(), 25e6, range(1, 37), num_coherently_integrated_code_periods=25,
                    num_noncoherent_accumulations=4)
    assert (c4.samples_per_code, c4.num_blocks, c4.block_size, c4.fft_size, c4.num_doppler_bins) == \
        (100000, 80, 1250, 2500, 2000)
    assert c4.num_secondary_rotations == 25 and c4.samples_per_code_eff == 2500000 and c4.path == 2
    c5 = _plan_host(GPSL1CA(), 5e6, [1], min_doppler_coverage=50000.0)
    assert (c5.num_blocks, c5.num_doppler_bins) == (125, 125)
    l5 = _plan_host(L5Like(), 30.69e6, [1], min_doppler_coverage=5000.0, num_coherently_integrated_code_periods=10)
    assert (l5.samples_per_code, l5.num_blocks, l5.fft_size, l5.num_doppler_bins) == (30690, 11, 5600, 110)


def test_julia_range_and_nextprod():
    assert list(julia_float_range(-8000.0, 100.0, 3)) == [-8000.0, -7900.0, -7800.0]
    r = julia_float_range(-0.5, 0.1, 11)
    assert r[10] == 0.5 and r[3] == pytest.approx(-0.2, abs=0)  # rational evaluation, no accumulation error
    assert nextprod((2, 3, 5, 7), 2046) == 2048 and nextprod((2, 3, 5, 7), 2976) == 3000
    assert nextprod((2, 3, 5, 7), 2728) == 2744 and nextprod((2, 3, 5, 7), 5580) == 5600


def test_sign_patterns_host():
    nh10 = GPSL5I()
    p = sign_patterns(nh10, 1, 1, 10, 10, True)
    assert list(p[:, 0]) == [1, 1, 1, 1, -1, -1, 1, -1, 1, -1]
    p = sign_patterns(None, 1, 2, 1, 20, False)
    assert p.shape == (20, 2) and np.all(p[10:, 1] == -1)
    p = sign_patterns(nh10, 1, 2, 10, 20, True)
    assert p.shape == (20, 20)  # rotation-major, data-minor
    assert np.all(p[:10, 1] == p[:10, 0]) and np.all(p[10:, 1] == -p[10:, 0])


def test_drift_shifts():
    """test/noncoherent_integration.jl:166-225."""
    pa = _plan_host(GPSL1CA(), 5e6, [1], min_doppler_coverage=2000.0, num_noncoherent_accumulations=8)
    assert np.all(pa.code_drift_shifts == 0)
    pb = _plan_host(GPSL1CA(), 2.048e6, [1], min_doppler_coverage=12000.0,
                    num_coherently_integrated_code_periods=20, num_noncoherent_accumulations=8)
    assert np.all(pb.code_drift_shifts[0] == 0) and np.any(pb.code_drift_shifts[7] != 0)
    assert np.all((pb.code_drift_shifts >= 0) & (pb.code_drift_shifts < pb.samples_per_code))


def test_cfar_threshold_host():
    assert cfar_threshold(0.01, 1) == pytest.approx(-math.log(0.01), rel=1e-12)
    with pytest.raises(ArgumentError):
        cfar_threshold(0.0, 100)
    with pytest.raises(ArgumentError):
        cfar_threshold(1.0, 100)
    with pytest.raises(ArgumentError):
        cfar_threshold(-0.1, 100)
    with pytest.raises(ArgumentError):
        cfar_threshold(0.01, 0)
    import oracle

    for M in (1, 3, 8):
        assert cfar_threshold(0.01, 2.6e6, num_noncoherent_integrations=M) == pytest.approx(
            oracle.cfar_threshold(0.01, 2.6e6, M), rel=1e-10)


def test_scalar_formulas():
    assert [fmdbzp_column_to_tau(c, 4, 8) for c in (0, 1, 8, 16, 24)] == [0, 1, 24, 16, 8]
    assert jl_mod(-5.5, 3.0) == 0.5 and jl_mod(0.0, 1023.0) == 0.0 and jl_mod(-0.0, 1023.0) == 0.0
    assert parabolic_interp(1.0, 2.0, 1.0) == 0.0
    assert isinstance(parabolic_interp(1.0, 1.0, 1.0), float)  # zero denominator -> Float64 0.0
    assert parabolic_interp(1.0, 3.0, 2.0) == np.float32(1.0) / np.float32(6.0)


def test_num_cells_and_detection():
    r = AcquisitionResults(system=GPSL5I(), prn=1, sampling_frequency=12e6, carrier_doppler=0.0, code_phase=0.0,
                           secondary_code_phase=None, CN0=0.0, noise_power=np.float32(1),
                           peak_to_noise_ratio=np.float32(30), num_noncoherent_integrations=1, power_bins=None,
                           dopplers=np.zeros(150), num_blocks=15, block_size=800, num_secondary_rotations=10)
    assert get_num_cells(r) == 150 * 12000 * 10
    thr = cfar_threshold(0.01, get_num_cells(r))
    assert is_detected(r) == (30 > thr)
