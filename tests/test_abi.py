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
  printf("%zu %zu %zu %zu %zu %zu\\n", sizeof(fmdbzp_emitter), offsetof(fmdbzp_emitter, bits),
         offsetof(fmdbzp_emitter, carrier_phase), sizeof(fmdbzp_scene), offsetof(fmdbzp_scene, seed),
         offsetof(fmdbzp_scene, noise_stream));
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
    E, S = _native.Emitter, _native.SceneDesc
    assert [int(v) for v in out[4:]] == [ctypes.sizeof(E), E.bits.offset, E.carrier_phase.offset,
                                         ctypes.sizeof(S), S.seed.offset, S.noise_stream.offset]


def test_synthesize_rejects_bad_scenes():
    """Validation happens before any device work, so it runs without a GPU."""
    lib = _native.load_library()
    chips = np.ones(8, dtype=np.int8)
    em = _native.Emitter(chips=chips.ctypes.data, code_length=8, amplitude=1.0, code_freq_hz=1e3)
    sc = _native.SceneDesc(n_samples=16, sampling_freq_hz=1e4, complex_samples=0, out_eltype=_native.ELTYPE["i16"])
    out = ctypes.c_void_p(1)
    assert lib.fmdbzp_synthesize(None, None, 0, out) == -1
    sc.out_eltype = _native.ELTYPE["cf32"]
    assert lib.fmdbzp_synthesize(ctypes.byref(sc), ctypes.byref(em), 1, out) == -1
    assert b"out_eltype" in lib.fmdbzp_last_error()
    sc.out_eltype = _native.ELTYPE["i16"]
    chips[3] = 0
    assert lib.fmdbzp_synthesize(ctypes.byref(sc), ctypes.byref(em), 1, out) == -1
    assert b"+-1" in lib.fmdbzp_last_error()
    chips[3] = 1
    bits = np.ones(4, dtype=np.int8)
    em.bits, em.num_bits, em.bit_ms, em.bit_edge_ms = bits.ctypes.data, 4, 20, -1
    assert lib.fmdbzp_synthesize(ctypes.byref(sc), ctypes.byref(em), 1, out) == -1
    assert b"bit_edge_ms" in lib.fmdbzp_last_error()
    sc.n_samples = 0
    assert lib.fmdbzp_synthesize(ctypes.byref(sc), ctypes.byref(em), 1, out) == -1


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
