"""Secondary-code rotation search and bit-edge search behaviour on the
device, restating the reference's end-to-end checks with the same signals
and bounds (test/secondary_code_search.jl, test/noncoherent_integration.jl):
rotation discovery at every NH10 starting phase with the exact secondary
phase, the coherent gain over the opt-out, the worst-case code phase, the
CFAR cell count with the rotation expansion, the rotation search's Doppler
axis, two data symbols under the rotation search, data-bit transitions, the
4-step bit-edge search on a strong signal, and the rotation search against one
10 ms transform of the combined 102300-chip code.
"""
import math

import numpy as np
import pytest

from acquisition import GPSL1CA, GPSL5I, acquire, acquire_, cfar_threshold, get_num_cells, is_detected, plan_acquire
from acquisition.signals import generate_test_signal
from acquisition.systems import GNSSSystem

pytestmark = pytest.mark.gpu


def _l5i(fs, r0=0, nper=10, prn=1):
    """Clean L5I pilot: period k carries NH10 chip (k + r0) mod 10 times the
    primary code (code phase 0)."""
    sysm = GPSL5I()
    spc = math.ceil(sysm.code_length / sysm.code_frequency * fs)
    primary = sysm.gen_code(spc, prn, fs, sysm.code_frequency, 0.0).astype(np.float32)
    return np.concatenate([np.float32(sysm.secondary_value(prn, k + r0)) * primary for k in range(nper)]).astype(
        np.complex64)


def test_rotation_discovery_at_every_nh10_phase():
    """test/secondary_code_search.jl:32-87: for every starting phase r the
    rotation search beats the opt-out peak by more than 5x and reports
    secondary_code_phase == r; with the opt-out, and for L1 C/A, the phase is
    None."""
    print("secondary: rotation discovery", flush=True)
    fs = 10.24e6
    kw = dict(min_doppler_coverage=1000.0, num_coherently_integrated_code_periods=10, num_noncoherent_accumulations=1)
    plan = plan_acquire(GPSL5I(), fs, [1], **kw)
    off = plan_acquire(GPSL5I(), fs, [1], use_secondary_code=False, **kw)
    try:
        ref = _l5i(fs)
        peak_off = float(np.max(acquire_(off, ref, [1], store_power_bins=True)[0].power_bins))
        for r in range(10):
            res = acquire_(plan, _l5i(fs, r0=r), [1], store_power_bins=True)[0]
            peak_on = float(np.max(res.power_bins))
            print(f"r {r}: on/off {peak_on / peak_off:.2f}, secondary phase {res.secondary_code_phase}")
            assert peak_on > 5 * peak_off
            assert res.secondary_code_phase == r
        assert acquire_(off, ref, [1])[0].secondary_code_phase is None
    finally:
        plan.close()
        off.close()
    l1 = plan_acquire(GPSL1CA(), 5e6, [1], num_coherently_integrated_code_periods=1)
    try:
        rng = np.random.default_rng(1)
        n = l1.samples_per_code
        sig = ((rng.standard_normal(n) + 1j * rng.standard_normal(n)) / math.sqrt(2)).astype(np.complex64)
        assert acquire_(l1, sig, [1])[0].secondary_code_phase is None
    finally:
        l1.close()


def test_rotation_gain_over_optout():
    """test/secondary_code_search.jl:133-175: on a clean 10 ms NH10 signal the
    rotation search's peak exceeds the opt-out peak by more than 8x."""
    print("secondary: gain over opt-out", flush=True)
    fs = 10.24e6
    kw = dict(min_doppler_coverage=5000.0, num_coherently_integrated_code_periods=10, num_noncoherent_accumulations=1)
    on = plan_acquire(GPSL5I(), fs, [1], **kw)
    off = plan_acquire(GPSL5I(), fs, [1], use_secondary_code=False, **kw)
    try:
        sig = _l5i(fs)
        p_on = float(np.max(acquire_(on, sig, [1], store_power_bins=True)[0].power_bins))
        p_off = float(np.max(acquire_(off, sig, [1], store_power_bins=True)[0].power_bins))
    finally:
        on.close()
        off.close()
    print(f"L5I 10 ms peak ratio (on/off) {p_on / p_off:.2f}")
    assert p_on / p_off > 8


def test_rotation_worst_case_code_phase():
    """test/secondary_code_search.jl:330-386: code phase 5115 chips puts the
    primary-code wrap mid-block; the search still detects at the right code
    phase and Doppler bin."""
    print("secondary: worst-case code phase", flush=True)
    fs, cp, dop = 12e6, 5115.0, 1000.0
    g = generate_test_signal(GPSL5I(), 1, num_samples=10 * 12000, doppler=dop, code_phase=cp, sampling_freq=fs,
                             interm_freq=0.0, CN0=45)
    plan = plan_acquire(GPSL5I(), fs, [1], num_coherently_integrated_code_periods=10, num_noncoherent_accumulations=1)
    try:
        assert plan.num_secondary_rotations > 1
        r = acquire_(plan, g["signal"], 1, interm_freq=0.0)
        assert is_detected(r)
        assert abs(r.code_phase - cp) <= 1.0
        assert abs(r.carrier_doppler - dop) < plan.doppler_step
    finally:
        plan.close()


def test_cell_count_includes_rotation_expansion():
    """test/secondary_code_search.jl:473-530: get_num_cells counts
    ndop x spc x R on the rotation path (the stored surface's size) and
    ndop x spc elsewhere; is_detected flips at that count's threshold."""
    print("secondary: cell count", flush=True)
    fs = 12e6
    plan = plan_acquire(GPSL5I(), fs, [1], num_coherently_integrated_code_periods=10, num_noncoherent_accumulations=1)
    try:
        assert plan.num_secondary_rotations == 10
        g = generate_test_signal(GPSL5I(), 1, num_samples=10 * 12000, doppler=1000.0, code_phase=1234.5,
                                 sampling_freq=fs, interm_freq=0.0, CN0=45)
        r = acquire_(plan, g["signal"], 1, interm_freq=0.0, store_power_bins=True)
        ndop = len(r.dopplers)
        assert r.num_secondary_rotations == plan.num_secondary_rotations
        assert get_num_cells(r) == ndop * plan.samples_per_code * plan.num_secondary_rotations
        assert get_num_cells(r) == r.power_bins.size
        assert get_num_cells(r) == 10 * ndop * plan.num_blocks * plan.block_size
        thr = cfar_threshold(0.01, get_num_cells(r), num_noncoherent_integrations=r.num_noncoherent_integrations)
        assert is_detected(r, 0.01) == (float(r.peak_to_noise_ratio) > thr)
        assert is_detected(r, 0.01) and r.secondary_code_phase is not None
    finally:
        plan.close()
    g = generate_test_signal(GPSL1CA(), 1)
    r = acquire(GPSL1CA(), g["signal"], g["sampling_freq"], [1], interm_freq=g["interm_freq"])[0]
    assert r.num_secondary_rotations == 1
    assert get_num_cells(r) == len(r.dopplers) * r.num_blocks * r.block_size


def test_bit_edge_search_keeps_strong_signal():
    """test/noncoherent_integration.jl:458-491: N_coh 40 at 60 dB-Hz, 3050 Hz;
    with 1 and with 4 bit-edge steps the Doppler is an alias of the truth
    within one bin (the 1 kHz alias period of 1 ms blocks)."""
    print("secondary: bit-edge steps", flush=True)
    fs, dop = 2.048e6, 3050.0
    kw = dict(min_doppler_coverage=10_000.0, num_coherently_integrated_code_periods=40, num_noncoherent_accumulations=1)
    g = generate_test_signal(GPSL1CA(), 1, num_samples=40 * 2048, doppler=dop, code_phase=100.0, sampling_freq=fs,
                             interm_freq=0.0, CN0=60)
    sig = g["signal"].astype(np.complex64)
    for steps in (1, 4):
        plan = plan_acquire(GPSL1CA(), fs, [1], bit_edge_search_steps=steps, **kw)
        try:
            alias = plan.sampling_freq / plan.samples_per_code
            r = acquire_(plan, sig, [1], interm_freq=0.0)[0]
            diff = (r.carrier_doppler - dop) % alias
            print(f"bit-edge steps {steps}: Doppler {r.carrier_doppler}, alias distance {min(diff, alias - diff):.1f}")
            assert min(diff, alias - diff) < plan.doppler_step
        finally:
            plan.close()


def test_rotation_doppler_axis():
    """test/secondary_code_search.jl:177-206: the rotation search reports the
    absolute Doppler within one bin (not half the band away)."""
    print("secondary: rotation Doppler", flush=True)
    fs, dop, cp = 12e6, 1000.0, 1234.5
    g = generate_test_signal(GPSL5I(), 1, num_samples=10 * 12000, doppler=dop, code_phase=cp, sampling_freq=fs,
                             interm_freq=0.0, CN0=45)
    plan = plan_acquire(GPSL5I(), fs, [1], num_coherently_integrated_code_periods=10, num_noncoherent_accumulations=1)
    try:
        assert plan.num_secondary_rotations > 1
        r = acquire_(plan, g["signal"], 1, interm_freq=0.0)
        assert is_detected(r)
        assert abs(r.code_phase - cp) <= 1.0
        assert abs(r.carrier_doppler - dop) < plan.doppler_step
    finally:
        plan.close()


def test_rotation_with_two_data_symbols():
    """test/secondary_code_search.jl:208-258: N_coh 20 on L5I (two 10 ms
    symbols, a data-bit flip between them) runs the data-bit search and the
    rotation search together; detection, code phase, Doppler and secondary
    phase 0 are recovered."""
    print("secondary: N_coh 20, data flip", flush=True)
    fs, dop, cp, N, L = 12e6, 1000.0, 1234.5, 20, 10
    sysm = GPSL5I()
    spc = math.ceil(sysm.code_length / sysm.code_frequency * fs)
    primary = sysm.gen_code(spc, 1, fs, sysm.code_frequency, cp).astype(np.float64)
    sig = np.concatenate([sysm.secondary_value(1, k % L) * (1.0 if k < L else -1.0) * primary for k in range(N)])
    sig = (sig * np.exp(1j * 2 * np.pi * dop / fs * np.arange(sig.size))).astype(np.complex64)
    plan = plan_acquire(sysm, fs, [1], num_coherently_integrated_code_periods=N, num_noncoherent_accumulations=1)
    try:
        assert plan.num_secondary_rotations == L and plan.num_data_bits == N // L
        r = acquire_(plan, sig, 1, interm_freq=0.0)
        assert is_detected(r)
        assert abs(r.code_phase - cp) <= 1.0
        assert abs(r.carrier_doppler - dop) < plan.doppler_step
        assert r.secondary_code_phase == 0
    finally:
        plan.close()


@pytest.mark.parametrize("ncoh,flip", [(40, (20, 40)), (60, (20, 40))])
def test_data_bit_transitions_recovered(ncoh, flip):
    """test/noncoherent_integration.jl:602-702: at 0 Hz and 60 dB-Hz, a signal
    with data-bit transitions (the bit after 20 ms negated, of 2 or 3 bits)
    reaches the peak of the same signal without transitions within 10 %."""
    print(f"secondary: bit transitions, N_coh {ncoh}", flush=True)
    fs = 2.048e6
    g = generate_test_signal(GPSL1CA(), 1, num_samples=ncoh * 2048, doppler=0.0, code_phase=0.0, sampling_freq=fs,
                             interm_freq=0.0, CN0=60, seed=42)
    clean = g["signal"].astype(np.complex64)
    flipped = clean.copy()
    flipped[flip[0] * 2048:flip[1] * 2048] *= -1
    plan = plan_acquire(GPSL1CA(), fs, [1], min_doppler_coverage=10_000.0, num_coherently_integrated_code_periods=ncoh,
                        bit_edge_search_steps=1, num_noncoherent_accumulations=1)
    try:
        assert plan.num_data_bits == ncoh // 20
        p0 = float(np.max(acquire_(plan, clean, [1], interm_freq=0.0, store_power_bins=True)[0].power_bins))
        p1 = float(np.max(acquire_(plan, flipped, [1], interm_freq=0.0, store_power_bins=True)[0].power_bins))
    finally:
        plan.close()
    print(f"N_coh {ncoh}: peak with transitions / without {p1 / p0:.4f}")
    assert abs(p1 / p0 - 1) <= 0.1


class _LongL5I(GNSSSystem):
    """GPS L5I re-presented as one 102300-chip primary code (10 primary
    periods times NH10) with no secondary code, so that an N_coh 1 search runs
    one 10 ms column transform (test/secondary_code_search.jl:1-29)."""

    def __init__(self):
        l5 = GPSL5I()
        super().__init__("LongL5I", 10 * l5.code_length, l5.get_code_frequency(), l5.get_center_frequency())
        self._l5 = l5

    def _make_chips(self, prn):
        l5 = self._l5
        return np.concatenate([l5.primary_chips(prn) * l5.secondary_value(prn, k) for k in range(10)]).astype(np.int8)


def test_rotation_matches_full_10ms_transform():
    """test/secondary_code_search.jl:388-471: on a noiseless NH10 signal the
    rotation search's peak reaches more than 0.95 of the one-long-code 10 ms
    search's peak, on the 1 kHz sub-block grid (1000 Hz) and off it (100 Hz,
    500 Hz)."""
    print("secondary: rotation vs one 10 ms transform", flush=True)
    fs, ncoh, spc = 12e6, 10, 12000
    l5, long = GPSL5I(), _LongL5I()
    code = l5.gen_code(ncoh * spc, 1, fs, l5.get_code_frequency(), 0.0).astype(np.float64)
    rot = plan_acquire(l5, fs, [1], num_coherently_integrated_code_periods=ncoh, num_noncoherent_accumulations=1)
    lng = plan_acquire(long, fs, [1], num_coherently_integrated_code_periods=1, num_noncoherent_accumulations=1)
    try:
        assert rot.num_doppler_bins == lng.num_doppler_bins == 150
        assert rot.doppler_step == pytest.approx(lng.doppler_step)
        for dop in (1000.0, 100.0, 500.0):
            sig = (np.exp(1j * 2 * np.pi * dop / fs * np.arange(code.size)) * code).astype(np.complex64)
            pr = float(np.max(acquire_(rot, sig, 1, interm_freq=0.0, store_power_bins=True).power_bins))
            pl = float(np.max(acquire_(lng, sig, 1, interm_freq=0.0, store_power_bins=True).power_bins))
            print(f"{dop:.0f} Hz: rotation / long {pr / pl:.4f}")
            assert pr / pl > 0.95
    finally:
        rot.close()
        lng.close()
