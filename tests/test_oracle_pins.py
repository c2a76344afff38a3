"""Pins of the CPU oracle against the reference's own known answers, closed
forms and brute-force restatements (no GPU).

What each check pins:
  - L5I opt-out KAT: values the reference's test suite pins
    (test/secondary_code_search.jl:98-130) -> code generation, geometry,
    signal/code FFTs, tile build, column FFT, fold, reductions.
  - brute force on tiny geometries (time-domain correlation + direct DFT in
    numpy FP64, written from the reference's definitions) for the simple,
    sign (data bits + bit edges) and rotation paths and for multistep drift.
  - FFT against numpy and a direct DFT for smooth and prime sizes.
  - est_signal_noise_power matrices (test/est_signal_noise_power.jl),
    column->tau (test/plot.jl), sign patterns (test/sign_patterns.jl),
    CFAR at M = 1 (test/cfar.jl) and the integer-M closed form vs scipy.
  - peak-column closed form (test/coherent_integration.jl:38-49, :67-77).
"""
import math

import numpy as np
import pytest

import oracle
from acquisition.plan import _plan_host
from acquisition.systems import GNSSSystem, GPSL1CA, GPSL5I
from acquisition.signals import generate_test_signal


# ---------------------------------------------------------------- FFT ------
@pytest.mark.parametrize("n", [1, 2, 3, 8, 20, 40, 55, 110, 125, 160, 257, 500, 2048, 2500, 5600])
def test_fft_against_numpy_and_direct_dft(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = np.fft.fft(x)
    f64 = oracle.fft(x, bits=64)
    assert np.max(np.abs(f64 - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))
    inv = oracle.fft(x, inverse=True, bits=64)
    assert np.max(np.abs(inv - np.fft.ifft(x) * n)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))
    if n <= 257:
        k = np.arange(n)
        direct = np.exp(-2j * np.pi * np.outer(k, k) / n) @ x
        assert np.max(np.abs(f64 - direct)) <= 1e-10 * max(1.0, np.max(np.abs(ref)))
    f32 = oracle.fft(x.astype(np.complex64), bits=32)
    assert np.max(np.abs(f32 - ref)) <= 2e-6 * math.sqrt(math.log2(max(n, 2))) * np.sqrt(n) * np.max(np.abs(x))


# ---------------------------------------------------- reference KATs -------
def l5i_signal(fs, prn=1, nper=10, r0=0, cp=0.0):
    sysm = GPSL5I()
    spc = math.ceil(sysm.code_length / sysm.code_frequency * fs)
    primary = sysm.gen_code(spc, prn, fs, sysm.code_frequency, cp).astype(np.float32)
    return np.concatenate([np.float32(sysm.secondary_value(prn, k + r0)) * primary for k in range(nper)]).astype(
        np.complex64)


@pytest.mark.parametrize("bits", [32, 64])
def test_l5i_optout_known_answers(bits):
    """test/secondary_code_search.jl:113-130 (rtol 1e-4 there)."""
    plan = _plan_host(GPSL5I(), 10.24e6, [1], min_doppler_coverage=1000.0,
                      num_coherently_integrated_code_periods=10, use_secondary_code=False)
    res, surf = oracle.acquire(plan, l5i_signal(10.24e6), [0], surfaces=True, bits=bits)
    nim = surf[0]
    assert nim.shape == (40, 10240)
    assert nim.max() == pytest.approx(1.146687232e9, rel=1e-6)
    assert nim.sum() == pytest.approx(5.2553412608e10, rel=1e-6)
    assert nim[0, 99] == pytest.approx(8.2944015625e4, rel=1e-6)
    assert nim[-1, -1] == pytest.approx(2.318815625e4, rel=2e-6)
    assert (res[0]["row"] + 1, res[0]["col"] + 1) in ((17, 1), (25, 1))
    assert nim[16, 0] == pytest.approx(nim[24, 0], rel=1e-6)


def test_est_signal_noise_power_matrices():
    """test/est_signal_noise_power.jl:4-97."""
    base = np.abs(1 / np.sqrt(2) * (np.ones((29, 1023)) + 1j * np.ones((29, 1023)))) ** 2
    for (r, c) in [(11, 489), (1, 1), (29, 1023)]:
        for est in (0, 1):
            pb = base.copy()
            pb[r - 1, c - 1] += 10 ** 1.5
            s, n, ci, di = oracle.est_signal_noise_power(pb, 4e6, 1e6, est, bits=64)
            assert n == pytest.approx(1.0)
            assert s == pytest.approx(10 ** 1.5)
            assert (ci, di) == (c, r)
    pb = np.ones((29, 1023))
    pb[10, :] = 5.0
    pb[10, 488] = 5.0 + 10 ** 1.5
    s, n, ci, di = oracle.est_signal_noise_power(pb, 4e6, 1e6, 0, bits=64)
    assert n == pytest.approx(1.0) and s == pytest.approx(5.0 + 10 ** 1.5 - 1)
    assert (ci, di) == (489, 11)


def test_column_to_tau():
    """test/plot.jl:28-49."""
    assert [oracle.column_to_tau(c, 4, 8) for c in (0, 1, 8, 16, 24)] == [0, 1, 24, 16, 8]
    assert oracle.column_to_tau(0, 32, 64) == 0
    assert sorted(oracle.column_to_tau(c, 4, 8) for c in range(32)) == list(range(32))


def test_sign_patterns_known_answers():
    """test/sign_patterns.jl:2-54."""
    assert np.array_equal(oracle.sign_patterns(None, 1, 1, 4, False), np.ones((4, 1)))
    p = oracle.sign_patterns(None, 2, 1, 20, False)
    assert p.shape == (20, 2)
    assert np.all(p[:, 0] == 1) and np.all(p[:10, 1] == 1) and np.all(p[10:, 1] == -1)
    nh10 = GPSL5I().secondary_code(1)
    p = oracle.sign_patterns(nh10, 1, 10, 10, True)
    assert p.shape == (10, 10)
    assert list(p[:, 0]) == [1, 1, 1, 1, -1, -1, 1, -1, 1, -1]
    for r in range(10):
        assert list(p[:, r]) == [nh10[(k + r) % 10] for k in range(10)]
    assert np.array_equal(oracle.sign_patterns(nh10, 1, 10, 10, False), np.ones((10, 1)))


def test_cfar_closed_form():
    """test/cfar.jl:5-26; integer M against scipy's inverse incomplete gamma."""
    from scipy.special import gammainccinv

    assert oracle.cfar_threshold(0.01, 1) == pytest.approx(-math.log(0.01), rel=1e-13)
    assert oracle.cfar_threshold(0.001, 1) == pytest.approx(-math.log(0.001), rel=1e-13)
    assert oracle.cfar_threshold(0.01, 10000) > oracle.cfar_threshold(0.01, 100)
    assert oracle.cfar_threshold(0.001, 1000) > oracle.cfar_threshold(0.1, 1000)
    assert oracle.cfar_threshold(0.01, 1000, 3) < oracle.cfar_threshold(0.01, 1000, 1)
    for M in (1, 2, 4, 8, 25):
        for cells in (1, 100, 2.6e6, 5e9):
            p = 1 - (1 - 0.01) ** (1 / cells)
            assert oracle.cfar_threshold(0.01, cells, M) == pytest.approx(gammainccinv(M, p) / M, rel=1e-10)


# ----------------------------------------------- closed-form peaks ---------
@pytest.mark.parametrize("code_phase", [991.0, 923.097656])
def test_peak_column_closed_form(code_phase):
    """test/coherent_integration.jl:4-78: every row peaks at
    mod(nb - tau / bs, nb) bs + tau % bs."""
    sysm = GPSL1CA()
    fs = 2.048e6
    plan = _plan_host(sysm, fs, [1], min_doppler_coverage=5000.0)
    assert (plan.num_blocks, plan.block_size) == (16, 128)
    tau = round((1023 - code_phase) * plan.samples_per_code / 1023)
    g = generate_test_signal(sysm, 1, num_samples=plan.samples_per_code, doppler=0.0, code_phase=code_phase,
                             sampling_freq=fs, interm_freq=0.0, CN0=100)
    code = sysm.gen_code(plan.samples_per_code, 1, fs, sysm.code_frequency, 0.0)
    cim = oracle.coherent_matrix(plan, code, g["signal"].astype(np.complex64))
    exp = ((plan.num_blocks - tau // plan.block_size) % plan.num_blocks) * plan.block_size + tau % plan.block_size
    assert int(np.argmax(np.abs(cim[0, :]))) == exp


def test_doppler_alias_and_code_delay():
    """test/noncoherent_integration.jl:279-341 (N_coh 20, 1050 Hz)."""
    sysm = GPSL1CA()
    fs = 2.048e6
    plan = _plan_host(sysm, fs, [1], min_doppler_coverage=10000.0, num_coherently_integrated_code_periods=20)
    g = generate_test_signal(sysm, 1, num_samples=20 * 2048, doppler=1050.0, code_phase=973.0, sampling_freq=fs,
                             interm_freq=0.0, CN0=45, seed=21)
    res, surf = oracle.acquire(plan, g["signal"].astype(np.complex64), [0], surfaces=True, bits=64)
    tau = round((1023 - 973.0) * 2048 / 1023)
    exp_col = ((plan.num_blocks - tau // plan.block_size) % plan.num_blocks) * plan.block_size + tau % plan.block_size
    assert res[0]["col"] == exp_col
    det = plan.doppler_freqs[res[0]["row"]]
    d = (det - 1050.0) % 1000.0
    assert min(d, 1000.0 - d) < plan.doppler_spacing_hz


def test_data_bit_transition_recovery():
    """test/noncoherent_integration.jl:401-456: a polarity flip at 20 ms is
    recombined by the N_db = 2 sign search; the detected Doppler is an alias."""
    sysm = GPSL1CA()
    fs = 2.048e6
    g = generate_test_signal(sysm, 1, num_samples=40 * 2048, doppler=1050.0, code_phase=100.0, sampling_freq=fs,
                             interm_freq=0.0, CN0=60, seed=22)
    sig = g["signal"].astype(np.complex64)
    sig[20 * 2048:] *= -1
    plan_s = _plan_host(sysm, fs, [1], min_doppler_coverage=10000.0, num_coherently_integrated_code_periods=40)
    assert plan_s.num_data_bits == 2 and plan_s.path == 1
    rs, ss = oracle.acquire(plan_s, sig, [0], surfaces=True)
    # pilot (no sign search): the same coherent window summed without patterns
    plan_p = _plan_host(GNSSSystem("pilot", 1023, 1.023e6, 1575.42e6), fs, [1], min_doppler_coverage=10000.0,
                        num_coherently_integrated_code_periods=40)
    code = sysm.gen_code(2048, 1, fs, sysm.code_frequency, 0.0)[None, :]
    rp, sp = oracle.acquire(plan_p, sig, [0], surfaces=True, codes=code)
    assert ss[0].max() > sp[0].max()
    det = plan_s.doppler_freqs[rs[0]["row"]]
    d = (det - 1050.0) % 1000.0
    assert min(d, 1000.0 - d) < plan_s.doppler_spacing_hz


# ------------------------------------------------ brute force --------------
class TinySystem(GNSSSystem):
    """32-chip seeded code at 1 Mchip/s (spc = 32 at fs = 1 MHz)."""

    def __init__(self, data_frequency=0.0, secondary=None, center=1575.42e6, seed=5):
        super().__init__("tiny", 32, 1e6, center, data_frequency=data_frequency,
                         secondary=None if secondary is None else np.asarray(secondary, dtype=np.int8))
        self.seed = seed

    def _make_chips(self, prn):
        return np.random.default_rng(self.seed + prn).choice([-1, 1], 32).astype(np.int8)


def _tile_direct(seg, code, nb, bs, ndop):
    """Linear correlation of every double block with its code sub-block, all
    column blocks: out[r][k][lag] (complex128), from the definitions of
    `_precompute_signal_block_ffts!` and `_build_coherent_tile!`."""
    seglen = len(seg)
    out = np.zeros((nb, ndop, bs), dtype=np.complex128)
    for r in range(nb):
        for k in range(ndop):
            db = np.array([seg[(k * bs + i) % seglen] for i in range(2 * bs)])
            sub = code[((k % nb + r) % nb) * bs:((k % nb + r) % nb) * bs + bs].astype(np.float64)
            for lag in range(bs):
                out[r, k, lag] = np.dot(sub, db[lag:lag + bs])
    return out


def _fold_simple(X):
    ndop = X.shape[0]
    P = np.abs(X) ** 2
    out = np.zeros_like(P)
    for w in range(ndop):
        out[(w + ndop // 2) % ndop] = P[w]
    return out


def _fold_sign(P, stage, use_max):
    ndop = P.shape[0]
    half = ndop // 2
    for w in range(half):
        stage[w + half] = max(stage[w + half], P[w]) if use_max else P[w]
    for w in range(half, ndop):
        stage[w - half] = max(stage[w - half], P[w]) if use_max else P[w]


def test_brute_force_simple_path():
    sysm = TinySystem()
    fs = 1e6
    plan = _plan_host(sysm, fs, [1, 2], min_doppler_coverage=31250.0, num_coherently_integrated_code_periods=2)
    nb, bs, ndop, spc = plan.num_blocks, plan.block_size, plan.num_doppler_bins, plan.samples_per_code
    assert (nb, bs, ndop) == (4, 8, 8)
    rng = np.random.default_rng(1)
    sig = (rng.standard_normal(2 * spc) + 1j * rng.standard_normal(2 * spc)).astype(np.complex64)
    res, surf = oracle.acquire(plan, sig, [0, 1], surfaces=True, bits=64)
    for j, prn in enumerate((1, 2)):
        code = sysm.gen_code(spc, prn, fs, 1e6, 0.0)
        tiles = _tile_direct(sig.astype(np.complex128), code, nb, bs, ndop)
        ref = np.zeros((ndop, spc))
        for r in range(nb):
            X = np.fft.fft(tiles[r], axis=0)
            ref[:, r * bs:(r + 1) * bs] = _fold_simple(X)
        assert np.max(np.abs(surf[j] - ref)) <= 1e-11 * ref.max()


def test_brute_force_sign_path_data_bits_and_edges():
    sysm = TinySystem(data_frequency=15625.0)
    fs = 1e6
    plan = _plan_host(sysm, fs, [1], min_doppler_coverage=31250.0, num_coherently_integrated_code_periods=4,
                      bit_edge_search_steps=2)
    nb, bs, ndop, spc = plan.num_blocks, plan.block_size, plan.num_doppler_bins, plan.samples_per_code
    assert plan.num_data_bits == 2 and plan.path == 1 and ndop == 16
    rng = np.random.default_rng(2)
    sig = (rng.standard_normal(4 * spc) + 1j * rng.standard_normal(4 * spc)).astype(np.complex64)
    res, surf = oracle.acquire(plan, sig, [0], surfaces=True, bits=64)
    code = sysm.gen_code(spc, 1, fs, 1e6, 0.0)
    tiles = _tile_direct(sig.astype(np.complex128), code, nb, bs, ndop)
    pats = plan.sign_patterns_by_prn[1]
    ndb, E = 2, 2
    sub_len = ndop // ndb
    cpb = 4 // ndb
    edge_step = (4 // ndb) * nb // E
    ref = np.zeros((ndop, spc))
    for r in range(nb):
        for lag in range(bs):
            x = tiles[r, :, lag]
            stage = np.zeros(ndop)
            for e in range(E):
                ro = e * edge_step
                subs = []
                for b in range(ndb):
                    y = np.zeros(ndop, dtype=np.complex128)
                    for j in range(sub_len):
                        y[j] = x[(ro + b * sub_len + j) % ndop]
                    subs.append(np.fft.fft(y))
                for d in range(pats.shape[1]):
                    comb = sum(pats[b * cpb, d] * subs[b] for b in range(ndb))
                    _fold_sign(np.abs(comb) ** 2, stage, True)
            ref[:, r * bs + lag] = stage
    assert np.max(np.abs(surf[0] - ref)) <= 1e-11 * ref.max()


def test_brute_force_rotation_path_and_fft_identity():
    sec = [1, -1, -1, 1]
    sysm = TinySystem(secondary=sec)
    fs = 1e6
    plan = _plan_host(sysm, fs, [1], min_doppler_coverage=31250.0, num_coherently_integrated_code_periods=4)
    nb, bs, ndop, spc = plan.num_blocks, plan.block_size, plan.num_doppler_bins, plan.samples_per_code
    R, ncoh = plan.num_secondary_rotations, 4
    assert plan.path == 2 and R == 4
    rng = np.random.default_rng(3)
    sig = (rng.standard_normal(4 * spc) + 1j * rng.standard_normal(4 * spc)).astype(np.complex64)
    res, surf = oracle.acquire(plan, sig, [0], surfaces=True, bits=64)
    code = sysm.gen_code(spc, 1, fs, 1e6, 0.0)
    tiles = _tile_direct(sig.astype(np.complex128), code, nb, bs, ndop)
    pats = plan.sign_patterns_by_prn[1]
    ref = np.zeros((ndop, spc * R))
    alt = np.zeros((ndop, spc * R))
    for r in range(nb):
        for lag in range(bs):
            x = tiles[r, :, lag]
            pc = r * bs + lag
            wrap = ((nb - pc // bs) % nb) * bs + pc % bs
            eff = ((wrap + (bs >> 1)) // bs) % nb
            subs = []
            for k in range(ncoh):
                y = np.zeros(ndop, dtype=np.complex128)
                for j in range(nb):
                    y[j] = x[(eff + k * nb + j) % ndop]
                subs.append(np.fft.fft(y))
            w = np.arange(ndop)
            for q in range(R):
                comb = sum(pats[p, q] * np.exp(-2j * np.pi * p * (w % ncoh) / ncoh) * subs[p] for p in range(ncoh))
                stage = np.zeros(ndop)
                _fold_sign(np.abs(comb) ** 2, stage, False)
                ref[:, q * spc + pc] = stage
                # FFT of the +-1-modulated, offset column (the GPU's form)
                z = np.array([pats[n // nb, q] * x[(eff + n) % ndop] for n in range(ndop)])
                st2 = np.zeros(ndop)
                _fold_sign(np.abs(np.fft.fft(z)) ** 2, st2, False)
                alt[:, q * spc + pc] = st2
    assert np.max(np.abs(surf[0] - ref)) <= 1e-11 * ref.max()
    assert np.max(np.abs(alt - ref)) <= 1e-11 * ref.max()


def test_brute_force_multistep_drift():
    sysm = TinySystem(center=1e6)
    fs = 1e6
    plan = _plan_host(sysm, fs, [1], min_doppler_coverage=31250.0, num_coherently_integrated_code_periods=2,
                      num_noncoherent_accumulations=3)
    nb, bs, ndop, spc = plan.num_blocks, plan.block_size, plan.num_doppler_bins, plan.samples_per_code
    assert np.any(plan.code_drift_shifts != 0)
    rng = np.random.default_rng(4)
    seglen = 2 * spc
    sig = (rng.standard_normal(3 * seglen) + 1j * rng.standard_normal(3 * seglen)).astype(np.complex64)
    res, surf = oracle.acquire(plan, sig, [0], surfaces=True, bits=64)
    code = sysm.gen_code(spc, 1, fs, 1e6, 0.0)
    acc = np.zeros((ndop, spc))
    for m in range(3):
        tiles = _tile_direct(sig[m * seglen:(m + 1) * seglen].astype(np.complex128), code, nb, bs, ndop)
        for r in range(nb):
            P = _fold_simple(np.fft.fft(tiles[r], axis=0))
            for row in range(ndop):
                d = round(plan.doppler_freqs[row] * m * (2 * spc / fs) * fs / sysm.center_frequency)
                for lag in range(bs):
                    acc[row, (r * bs + lag + d) % spc] += P[row, lag]
    assert np.max(np.abs(surf[0] - acc)) <= 1e-11 * acc.max()


def test_code_drift_shift_36_samples():
    """test/noncoherent_integration.jl:493-530: 7 kHz row, step 100, N_coh 40 at
    2.048 MHz shifts by round(7000*100*0.04*2.048e6/1575.42e6) = 36."""
    plan = _plan_host(GPSL1CA(), 2.048e6, [1], min_doppler_coverage=10000.0,
                      num_coherently_integrated_code_periods=40, num_noncoherent_accumulations=200)
    row = int(np.argmin(np.abs(plan.doppler_freqs - 7000.0)))
    assert plan.doppler_freqs[row] == 7000.0
    assert plan.code_drift_shifts[100, row] == 36
    assert np.all(plan.code_drift_shifts[0] == 0)


def test_downconvert_float32_phase():
    """The IF phase argument is Float32 (src/acquire.jl:96-101): the oracle's
    downconversion equals a numpy Float32 evaluation of the same formula, and
    differs from an FP64 phase by a growing amount over a long segment."""
    n = 200000
    rng = np.random.default_rng(5)
    x = rng.integers(-2000, 2000, n).astype(np.int16)
    fs, fif = 16.368e6, 4.092e6
    got = oracle.downconvert(x, 50000, 100000, fif, fs)
    step = np.float32((-6.283185307179586 * fif) / fs)
    off = np.float32(np.float32(50000) * step)
    k = np.arange(100000, dtype=np.float32)
    ph = (off + (k * step).astype(np.float32)).astype(np.float32)
    c = np.cos(ph.astype(np.float64)).astype(np.float32)
    s = np.sin(ph.astype(np.float64)).astype(np.float32)
    xs = x[50000:150000].astype(np.float32)
    exp = (xs * c).astype(np.float64) + 1j * (xs * s).astype(np.float64)
    assert np.array_equal(got, exp)
    ph64 = -2 * np.pi * fif / fs * (50000 + np.arange(100000))
    dphi = np.angle(np.exp(1j * (ph.astype(np.float64) - ph64)))
    assert np.max(np.abs(dphi)) > 1e-3


def test_secondary_phase_estimator_oracle():
    """`_estimate_secondary_code_phase`: the planted NH10 rotation is recovered
    by the C restatement and by an FP64 evaluation of the same sums with exact
    per-sample phasors."""
    sysm = GPSL5I()
    fs = 12e6
    for r0, cp, dop in [(0, 0.0, 0.0), (3, 2000.0, 0.0), (7, 5115.0, 1300.0), (9, 9000.0, 1300.0)]:
        sig = l5i_signal(fs, r0=r0, cp=cp)
        n = np.arange(sig.shape[0])
        sig = (sig * np.exp(2j * np.pi * dop * n / fs)).astype(np.complex64)
        code = sysm.gen_code(12000, 1, fs, sysm.code_frequency, cp)
        sec = sysm.secondary_code(1)
        assert oracle.secondary_phase(sig, 12000, 10, fs, 0.0, dop, code, sec) == r0
        carr = np.exp(-2j * np.pi * dop * n / fs)
        z = (sig.astype(np.complex128) * carr).reshape(10, 12000) @ code.astype(np.float64)
        mags = [abs(sum(float(sec[(k + r) % 10]) * z[k] for k in range(10))) ** 2 for r in range(10)]
        assert int(np.argmax(mags)) == r0
