"""Oracle parity over seeded sweeps of small geometries: the simple path on GPS
L1 C/A, then the sign search (L1 C/A) and the rotation search (GPS L5I).

24 geometries drawn with a fixed seed:
- sampling rates 2.048, 4, 5, 8.184 and 10 MHz;
- N_coh 1, 2, 5 or 10 and N_nc 1, 2 or 3;
- +-5, 7 or 10 kHz coverage, so F from 128 to 1250 and ndop from 16 to 320;
- 1, 3 or 8 PRNs.

Each is run through the engine and the oracle with the parity bar of
tests/parity.py, with subsample interpolation on every other case. The
sweep reaches the runtime-radix and compile-time correlation variants, the
split correlation grid, the ndop-160 column kernel at several block sizes and
the generic column kernel.
"""
import random

import numpy as np
import pytest

from acquisition import GPSL1CA, GPSL5I
from acquisition.signals import generate_test_signal
from test_gpu_parity import run_case

pytestmark = pytest.mark.gpu


def _cases(seed=7, n=24):
    rng = random.Random(seed)
    out = []
    for case in range(n):
        fs = rng.choice([2.048e6, 4e6, 5e6, 8.184e6, 10e6])
        ncoh = rng.choice([1, 2, 5, 10])
        nnc = rng.choice([1, 2, 3])
        cov = rng.choice([5000.0, 7000.0, 10000.0])
        nprn = rng.choice([1, 3, 8])
        prns = sorted(rng.sample(range(1, 33), nprn))
        dop, cp = rng.uniform(-3000, 3000), rng.uniform(0, 1023)
        out.append((case, fs, ncoh, nnc, cov, prns, dop, cp))
    return out


@pytest.mark.parametrize("case,fs,ncoh,nnc,cov,prns,dop,cp", _cases())
def test_geometry_sweep(case, fs, ncoh, nnc, cov, prns, dop, cp):
    label = f"sweep {case}: fs {fs / 1e6} MHz N_coh {ncoh} N_nc {nnc} +-{cov / 1e3:g} kHz PRNs {prns}"
    print(label, flush=True)
    spc = int(np.ceil(fs / 1000))
    g = generate_test_signal(GPSL1CA(), prns[0], num_samples=ncoh * nnc * spc + 50, sampling_freq=fs,
                             interm_freq=0.0, CN0=48, seed=case + 11, doppler=dop, code_phase=cp)
    run_case(GPSL1CA(), fs, prns, g["signal"].astype(np.complex64), interp=bool(case % 2), label=label,
             min_doppler_coverage=cov, num_coherently_integrated_code_periods=ncoh,
             num_noncoherent_accumulations=nnc)


def _sign_rotation_cases(seed=5, n=12):
    """Alternating GPS L1 C/A sign-search and GPS L5I rotation-search
    geometries. L1 C/A: 2.048 or 4 MHz, N_coh 20 or 40 (one or two data
    bits), 1, 2 or 4 bit-edge steps. L5I: 2.046 or 4.092 MHz, N_coh 10 or 20
    (R 10, one or two data bits), PRNs 1-10. Both: N_nc 1 or 2."""
    rng = random.Random(seed)
    out = []
    for case in range(n):
        if case % 2 == 0:
            sysm, fs = GPSL1CA(), rng.choice([2.048e6, 4e6])
            ncoh, kw = rng.choice([20, 40]), dict(bit_edge_search_steps=rng.choice([1, 2, 4]))
            nnc, cov = rng.choice([1, 2]), rng.choice([5000.0, 7000.0])
        else:
            sysm, fs = GPSL5I(), rng.choice([2.046e6, 4.092e6])
            ncoh, kw = rng.choice([10, 20]), {}
            nnc, cov = rng.choice([1, 2]), rng.choice([1000.0, 3000.0, 5000.0])
        prns = sorted(rng.sample(range(1, 33) if case % 2 == 0 else range(1, 11), rng.choice([1, 2])))
        dop, cp = rng.uniform(-800, 800), rng.uniform(0, sysm.code_length)
        out.append((case, sysm, fs, ncoh, nnc, cov, kw, prns, dop, cp))
    return out


@pytest.mark.parametrize("case,sysm,fs,ncoh,nnc,cov,kw,prns,dop,cp", _sign_rotation_cases(),
                         ids=lambda v: v.name if hasattr(v, "name") else None)
def test_sign_and_rotation_sweep(case, sysm, fs, ncoh, nnc, cov, kw, prns, dop, cp):
    label = f"sweep {case}: {sysm.name} fs {fs / 1e6} MHz N_coh {ncoh} N_nc {nnc} {kw} PRNs {prns}"
    print(label, flush=True)
    spc = int(np.ceil(sysm.code_length / sysm.code_frequency * fs))
    g = generate_test_signal(sysm, prns[0], num_samples=ncoh * nnc * spc + 50, sampling_freq=fs, interm_freq=0.0,
                             CN0=50, seed=case + 31, doppler=dop, code_phase=cp)
    run_case(sysm, fs, prns, g["signal"].astype(np.complex64), interp=bool(case % 3 == 0), label=label,
             min_doppler_coverage=cov, num_coherently_integrated_code_periods=ncoh,
             num_noncoherent_accumulations=nnc, **kw)
