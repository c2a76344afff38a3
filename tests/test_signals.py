"""Seeded scene generator on the host (CPU only).

The device generator (fmdbzp_synthesize) is checked against `render_host`
in test_gpu_synth.py, so `render_host` is pinned here against what fixes it
independently: the published Philox4x32-10 known answers (Random123
kat_vectors), the systems' own code generators, and closed forms.
"""
import math

import numpy as np
import pytest

from acquisition import GPSL5I, E1Like
from acquisition.rng import PhiloxStream, philox4x32_10
from acquisition.signals import Emitter, render_host

# (counter, key, output) from Random123's kat_vectors for philox4x32 with 10 rounds
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,want", PHILOX_KAT)
def test_philox_known_answers(ctr, key, want):
    got = philox4x32_10(*[np.array([c], dtype=np.uint32) for c in ctr], *key)
    assert tuple(int(v[0]) for v in got) == want


def test_normals_moments_and_pairing():
    g = PhiloxStream(7, 3).normal(400000)
    assert abs(g.mean()) < 0.01 and abs(g.var() - 1.0) < 0.01
    # one counter block -> (r cos t, r sin t): the pair's squared norm is -2 log(1 - a)
    x0, x1, _, _ = philox4x32_10(np.array([0], np.uint32), np.array([0], np.uint32), np.array([3], np.uint32),
                                 np.array([0], np.uint32), 7, 0)
    a = (float(x0[0] >> 5) * 67108864.0 + float(x1[0] >> 6)) / 9007199254740992.0
    assert math.isclose(g[0] ** 2 + g[1] ** 2, -2.0 * math.log(1.0 - a), rel_tol=1e-13)


def test_noise_only_scenes():
    n = 1001
    real = render_host(n, 1e6, False, np.float32, 3.0, 11, 5, [])
    np.testing.assert_array_equal(real, (PhiloxStream(11, 5).normal(n) * 3.0).astype(np.float32))
    cplx = render_host(n, 1e6, True, np.complex64, 0.5, 11, 5, [])
    g = PhiloxStream(11, 5).normal(2 * n) * 0.5
    np.testing.assert_array_equal(cplx.real, g[0::2].astype(np.float32))
    np.testing.assert_array_equal(cplx.imag, g[1::2].astype(np.float32))


def test_code_matches_system_generators():
    """Zero carrier, no noise: the samples are amplitude x the code that the
    systems' gen_code produces (BOC for E1-like, NH10 secondary baked into
    GPS L5I), with the same code Doppler and start phase."""
    fs, n = 5e6, 60000
    e1 = E1Like()
    cf = 1.023e6 * (1 + 1e-6)
    x = render_host(n, fs, False, np.float32, 0.0, 1, 1,
                    [Emitter(e1.primary_chips(3), 2.0, cf, 1234.6, 0.0, 0.0, boc=True)])
    np.testing.assert_array_equal(x, 2.0 * e1.gen_code(n, 3, fs, cf, 1234.6).astype(np.float32))
    l5 = GPSL5I()
    fs, n = 30.69e6, 200000
    cf = 10.23e6 - 3.7
    x = render_host(n, fs, False, np.float32, 0.0, 1, 1,
                    [Emitter(l5.primary_chips(9), 1.0, cf, 9876.25, 0.0, 0.0, secondary=l5.secondary_code(9))])
    np.testing.assert_array_equal(x, l5.gen_code(n, 9, fs, cf, 9876.25).astype(np.float32))


def test_secondary_phase_shifts_periods():
    """Secondary phase p reads secondary chip (k + p) in primary period k."""
    L, sec = 4, np.array([1, -1, -1, 1, -1], dtype=np.int8)
    n = 40
    x = render_host(n, 1.0, False, np.float32, 0.0, 1, 1,
                    [Emitter(np.ones(L, np.int8), 1.0, 1.0, 0.0, 0.0, 0.0, secondary=sec, secondary_phase=3)])
    k = np.arange(n) // L
    np.testing.assert_array_equal(x, sec[(k + 3) % len(sec)].astype(np.float32))


def test_data_bits_edges_and_clamp():
    """At 1 kHz sampling sample n is millisecond n: bit (n + edge) // bit_ms,
    held at the last bit beyond the table."""
    bits = np.array([1, -1, 1, -1, -1], dtype=np.int8)
    n = 130
    x = render_host(n, 1000.0, False, np.float32, 0.0, 1, 1,
                    [Emitter(np.ones(1, np.int8), 1.0, 0.0, 0.0, 0.0, 0.0, bits=bits, bit_ms=20, bit_edge_ms=7)])
    idx = np.minimum((np.arange(n) + 7) // 20, len(bits) - 1)
    np.testing.assert_array_equal(x, bits[idx].astype(np.float32))


def test_carrier_quarter_rate_and_phase():
    n = 64
    x = render_host(n, 4.0, True, np.complex64, 0.0, 1, 1,
                    [Emitter(np.ones(1, np.int8), 3.0, 0.0, 0.0, 1.0, math.pi / 2)])
    want = 3.0 * np.exp(1j * (np.pi / 2 * np.arange(n) + np.pi / 2))
    np.testing.assert_allclose(x, want, atol=2e-6)


def test_int16_rounding_and_saturation():
    one = np.ones(1, np.int8)
    for amp, want in [(2.5, 2), (3.5, 4), (-2.5, -2), (40000.0, 32767), (-40000.0, -32768)]:
        x = render_host(4, 1.0, False, np.int16, 0.0, 1, 1, [Emitter(one, amp, 0.0, 0.0, 0.0, 0.0)])
        assert x.dtype == np.int16 and (x == want).all(), (amp, x)
