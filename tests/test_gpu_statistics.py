"""Statistical behaviour of the whole search on the device.

These restate the reference's end-to-end statistical tests with the same
trial counts and bounds:
- the CFAR false-alarm rate on noise-only input with noncoherent
  integration (test/cfar.jl:54-89);
- the detection probability rising with noncoherent integration at CN0 30
  dB-Hz while the false-alarm rate stays controlled
  (test/noncoherent_integration.jl:532-600);
- the CN0 estimate's invariance to the coherent integration length
  (test/acquire.jl:462-520).
The noise and signals come from this package's seeded generators, so
individual trials differ from the reference's Julia RNG streams. The
assertions are on rates and averages, which do not depend on the stream.
"""
import math

import numpy as np
import pytest

from acquisition import GPSL1CA, acquire_, cfar_threshold, get_num_cells, is_detected, plan_acquire
from acquisition.signals import generate_test_signal

pytestmark = pytest.mark.gpu


def _noise(rng, n):
    return (rng.standard_normal(n) + 1j * rng.standard_normal(n)) / math.sqrt(2.0)


@pytest.mark.parametrize("M,ncoh", [(2, 1), (4, 1), (8, 1), (4, 4)])
def test_cfar_false_alarm_rate_with_noncoherent_integration(M, ncoh):
    """50 noise-only trials x 32 PRNs at 2.048 MHz: false alarms at most 3x
    the nominal 0.01 per PRN search."""
    print(f"statistics: false alarms M {M} N_coh {ncoh}", flush=True)
    system, fs, spc, pfa, n_trials = GPSL1CA(), 2.048e6, 2048, 0.01, 50
    rng = np.random.default_rng(0xC0FFEE + M * 1000)
    prns = list(range(1, 33))
    plan = plan_acquire(system, fs, prns, num_coherently_integrated_code_periods=ncoh,
                        num_noncoherent_accumulations=M)
    try:
        num_cells = plan.num_doppler_bins * plan.num_blocks * plan.block_size
        thr = cfar_threshold(pfa, num_cells, num_noncoherent_integrations=M)
        false_alarms = 0
        for _ in range(n_trials):
            res = acquire_(plan, _noise(rng, ncoh * spc * M), prns, interm_freq=0.0)
            false_alarms += sum(float(r.peak_to_noise_ratio) > thr for r in res)
    finally:
        plan.close()
    expected = n_trials * 32 * pfa
    print(f"M {M} N_coh {ncoh}: {false_alarms} false alarms, {expected:.0f} expected, threshold {thr:.4f}")
    assert false_alarms <= math.ceil(3 * expected)


def test_detection_probability_rises_with_noncoherent_integration():
    """CN0 30 dB-Hz, N_coh 10, 100 trials: M 1 detects rarely, M 8 almost
    always; the M 1 threshold applied to M 8 results detects less often than
    the M 8 threshold; false alarms stay below 3x pfa for both."""
    print("statistics: detection probability", flush=True)
    system, fs, prn, pfa, n_trials, cn0, ncoh, spc = GPSL1CA(), 2.048e6, 1, 0.01, 100, 30.0, 10, 2048
    rng = np.random.default_rng(20260423)
    kw = dict(min_doppler_coverage=10_000.0, num_coherently_integrated_code_periods=ncoh)
    p1 = plan_acquire(system, fs, [prn], num_noncoherent_accumulations=1, **kw)
    p8 = plan_acquire(system, fs, [prn], num_noncoherent_accumulations=8, **kw)
    det1 = det8 = det8_wrong = fa1 = fa8 = 0
    try:
        for seed in range(1, n_trials + 1):
            s1 = generate_test_signal(system, prn, num_samples=ncoh * spc, doppler=1500.0, code_phase=100.0,
                                      sampling_freq=fs, interm_freq=0.0, CN0=cn0, seed=seed)["signal"]
            s8 = generate_test_signal(system, prn, num_samples=8 * ncoh * spc, doppler=1500.0, code_phase=100.0,
                                      sampling_freq=fs, interm_freq=0.0, CN0=cn0, seed=seed)["signal"]
            r1 = acquire_(p1, s1, [prn], interm_freq=0.0)[0]
            r8 = acquire_(p8, s8, [prn], interm_freq=0.0)[0]
            det1 += is_detected(r1, pfa)
            det8 += is_detected(r8, pfa)
            thr1 = cfar_threshold(pfa, get_num_cells(r8), num_noncoherent_integrations=1)
            det8_wrong += float(r8.peak_to_noise_ratio) > thr1
            fa1 += is_detected(acquire_(p1, _noise(rng, ncoh * spc), [prn], interm_freq=0.0)[0], pfa)
            fa8 += is_detected(acquire_(p8, _noise(rng, 8 * ncoh * spc), [prn], interm_freq=0.0)[0], pfa)
    finally:
        p1.close()
        p8.close()
    pd1, pd8, pd8w = det1 / n_trials, det8 / n_trials, det8_wrong / n_trials
    pfa1, pfa8 = fa1 / n_trials, fa8 / n_trials
    print(f"Pd M1 {pd1:.2f}, M8 {pd8:.2f}, M8 with the M1 threshold {pd8w:.2f}; Pfa M1 {pfa1:.2f}, M8 {pfa8:.2f}")
    assert pd8 > pd1
    assert pd1 < 0.3
    assert pd8 > 0.7
    assert pfa1 < 3 * pfa and pfa8 < 3 * pfa
    assert pd8w < pd8


def test_cn0_is_independent_of_coherent_length():
    """PRN 1 at 45 dB-Hz, 4 MHz, +-25 kHz: the mean CN0 over 8 code phases
    (detected trials only, at least 5) lands within 6 dB of the truth for
    every N_coh in (1, 2, 4, 5, 10), the spread is below 4 dB, and N_coh 10
    exceeds N_coh 1 by less than 2 dB."""
    print("statistics: CN0 invariance", flush=True)
    system, prn, fs, true_cn0 = GPSL1CA(), 1, 4e6, 45.0
    reported = []
    for n in (1, 2, 4, 5, 10):
        est = []
        for k in range(1, 9):
            sig = generate_test_signal(system, prn, seed=700 + k, num_samples=4000 * n + 100, doppler=1234.0,
                                       code_phase=(k - 0.5) * 1023 / 8 + 0.137, sampling_freq=fs, interm_freq=0.0,
                                       CN0=true_cn0)["signal"]
            plan = plan_acquire(system, fs, [prn], num_coherently_integrated_code_periods=n,
                                min_doppler_coverage=25_000.0)
            try:
                r = acquire_(plan, sig, [prn], interm_freq=0.0, subsample_interpolation=True)[0]
            finally:
                plan.close()
            if is_detected(r):
                est.append(r.CN0)
        assert len(est) >= 5, (n, est)
        reported.append(sum(est) / len(est))
    print("mean CN0 per N_coh (1, 2, 4, 5, 10):", [round(c, 2) for c in reported])
    for c in reported:
        assert abs(c - true_cn0) <= 6.0
    assert max(reported) - min(reported) < 4.0
    assert reported[-1] - reported[0] < 2.0
