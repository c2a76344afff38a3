"""Device scene synthesis (fmdbzp_synthesize) against the host generator.

Both evaluate the same FP64 expressions in the same order; only log / sqrt /
cos / sin may differ in their last bit between the device math library and
the host's. After rounding to the output type that leaves: int16 samples
equal, or 1 LSB apart where the FP64 value sits within an ulp of a rounding
boundary (allowed in at most 1 of 10^4 samples); float samples equal, or one
float32 ulp apart (at most 1 in 10^3). Sizes are the benchmark scenes:
GPS L1 C/A 16.368 MHz real int16 80 ms with 8 PRNs, and the E1-like 25 MHz
complex 4 x 100 ms scene with 6 PRNs.
"""
import numpy as np
import pytest

from acquisition import GPSL1CA, acquire_, is_detected, plan_acquire
from acquisition.signals import Emitter, e1like_scene, l1ca_if_int16_scene, render_host

pytestmark = pytest.mark.gpu


def _assert_int16_close(dev, host):
    assert dev.dtype == host.dtype == np.int16 and dev.shape == host.shape
    d = np.abs(dev.astype(np.int32) - host.astype(np.int32))
    assert d.max() <= 1, d.max()
    assert np.count_nonzero(d) <= max(2, host.size // 10000), np.count_nonzero(d)


def _assert_f32_close(dev, host):
    assert dev.shape == host.shape
    d = np.abs(dev.astype(np.float64) - host.astype(np.float64))
    assert (d <= np.spacing(np.abs(host).astype(np.float32)).astype(np.float64)).all()
    assert np.count_nonzero(d) <= max(2, host.size // 1000), np.count_nonzero(d)


def test_l1ca_int16_scene_device_equals_host():
    print("synth: l1ca int16 80 ms", flush=True)
    host = l1ca_if_int16_scene(seed=0x5EED0002)
    dev = l1ca_if_int16_scene(seed=0x5EED0002, device=True)
    try:
        assert dev.present == host.present and dev.truth == host.truth
        _assert_int16_close(dev.signal.to_host(), host.signal)
    finally:
        dev.signal.free()


def test_e1like_scene_device_equals_host():
    print("synth: e1like complex 4 x 100 ms", flush=True)
    kw = dict(n_periods=25, nnc=4, fs=25e6, nprn=36, npresent=6)
    host = e1like_scene(**kw)
    dev = e1like_scene(device=True, **kw)
    try:
        got = dev.signal.to_host()
        assert got.dtype == np.complex64
        _assert_f32_close(got.real, host.signal.real)
        _assert_f32_close(got.imag, host.signal.imag)
    finally:
        dev.signal.free()


@pytest.mark.parametrize("n", [1, 2, 1001])
def test_real_float_odd_lengths(n):
    """Real scenes take one counter block per sample pair; an odd length ends
    on half a pair. Float output, negative secondary phase, data bits."""
    from acquisition.device import synthesize

    sysm = GPSL1CA()
    ems = [Emitter(sysm.primary_chips(5), 0.7, 1.023e6 + 0.3, 17.25, 4.092e6 - 1234.5, 2.5,
                   bits=np.array([1, -1, 1], np.int8), bit_ms=20, bit_edge_ms=3),
           Emitter(sysm.primary_chips(6), 0.4, 1.023e6 - 0.2, 700.5, 4.092e6 + 999.0, 0.1, boc=True,
                   secondary=np.array([1, -1, -1], np.int8), secondary_phase=-4)]
    args = (n, 16.368e6, False, np.float32, 2.0, 99, 4, ems)
    host = render_host(*args)
    dev = synthesize(*args)
    try:
        _assert_f32_close(dev.to_host(), host)
    finally:
        dev.free()


def test_device_scene_acquires_like_host_scene():
    """The engine reads the device-generated samples in place; the search on
    them matches the search on the host-generated ones."""
    print("synth: acquire on device scene", flush=True)
    host = l1ca_if_int16_scene(seed=77, n_ms=20, npresent=3, cn0_range=(42.0, 44.0))
    dev = l1ca_if_int16_scene(seed=77, n_ms=20, npresent=3, cn0_range=(42.0, 44.0), device=True)
    prns = sorted(host.present + [p for p in host.prns if p not in host.present][:2])
    plan = plan_acquire(GPSL1CA(), host.sampling_freq, prns, min_doppler_coverage=7000.0,
                        num_coherently_integrated_code_periods=10)
    try:
        rh = acquire_(plan, host.signal, prns, interm_freq=host.interm_freq)
        rd = acquire_(plan, dev.signal, prns, interm_freq=host.interm_freq)
    finally:
        plan.close()
        dev.signal.free()
    for a, b in zip(rh, rd):
        assert a.prn == b.prn
        assert is_detected(a) == is_detected(b) == (a.prn in host.present)
        if a.prn in host.present:
            assert abs(a.code_phase - b.code_phase) < 1e-3
            assert a.carrier_doppler == b.carrier_doppler
            assert abs(a.peak_to_noise_ratio - b.peak_to_noise_ratio) <= 1e-3 * a.peak_to_noise_ratio
