"""Oracle parity over a seeded sweep of small GPS L1 C/A geometries.

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

from acquisition import GPSL1CA
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
