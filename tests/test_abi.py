"""C-ABI boundary (no GPU compute): the library loads, exports every function
include/fmdbzp.h declares, its struct layouts match the ctypes mirrors, and
descriptor validation rejects bad plans before touching a device."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from acquisition import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "fmdbzp.h")


def declared_functions():
    text = open(HDR).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(fmdbzp_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = _native.load_library()
    names = declared_functions()
    assert len(names) >= 7
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_native.EXPORTED_SYMBOLS)


def test_struct_layouts_match_header(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text("""
#include <stdio.h>
#include <stddef.h>
#include "fmdbzp.h"
int main(void) {
  printf("%zu %zu %zu %zu\\n", sizeof(fmdbzp_plan_desc), offsetof(fmdbzp_plan_desc, sampling_freq_hz),
         sizeof(fmdbzp_prn_stats), offsetof(fmdbzp_prn_stats, noise_cells));
  return 0;
}
""")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    assert int(out[0]) == ctypes.sizeof(_native.PlanDesc)
    assert int(out[1]) == _native.PlanDesc.sampling_freq_hz.offset
    assert int(out[2]) == ctypes.sizeof(_native.PrnStats)
    assert int(out[3]) == _native.PrnStats.noise_cells.offset


def test_plan_create_rejects_bad_descriptor():
    lib = _native.load_library()
    d = _native.make_desc(dict(samples_per_code=100, num_blocks=3, block_size=33, fft_size=66, num_coh=1,
                               num_data_bits=1, bit_edge_search_steps=1, num_noncoherent=1,
                               num_secondary_rotations=1, num_doppler_bins=3, samples_per_code_eff=100,
                               num_patterns=1, path=0, noise_estimator=0, samples_per_chip=1, precision=0,
                               sampling_freq_hz=1e6))
    codes = np.ones(100, dtype=np.int8)
    shifts = np.zeros(3, dtype=np.int32)
    h = ctypes.c_void_p()
    rc = lib.fmdbzp_plan_create(ctypes.byref(d), 1, codes.ctypes.data, 0, None, None, shifts.ctypes.data,
                                ctypes.byref(h))
    assert rc == -1
    assert b"num_blocks * block_size" in lib.fmdbzp_last_error()
    d.abi_version = 99
    rc = lib.fmdbzp_plan_create(ctypes.byref(d), 1, codes.ctypes.data, 0, None, None, shifts.ctypes.data,
                                ctypes.byref(h))
    assert rc == -1 and b"abi_version" in lib.fmdbzp_last_error()
    assert lib.fmdbzp_plan_destroy(None) == 0
    assert lib.fmdbzp_acquire(None, None, 0, 0, 0.0, None, 0, 0, None, None) == -5
