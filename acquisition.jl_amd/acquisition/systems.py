"""GNSS signal descriptors and spreading-code generation.

The reference takes these from GNSSSignals.jl (compat "3, 4", not vendored):
`get_code_length`, `get_code_frequency`, `get_center_frequency`,
`get_data_frequency`, `get_secondary_code(_length)`, `secondary_value` and
`gen_code` (call sites src/plan.jl:241, :335, :386, :389, :511, :646-649;
src/acquire.jl:137-139, :508-509, :529). The engine receives code samples as
data through the C ABI, so these generators are host-side inputs only.

`gen_code(n, prn, fs, f_code, phase)` returns sample i as the chip at code
position phase + i f_code / fs (floor), for the primary code; a system whose
`secondary_baked` is set multiplies the secondary chip of period
floor(position / code_length) into the sample, as GNSSSignals does for GPS L5I
(test/secondary_code_search.jl:8-10, :428-432). A BOC(1,1) system multiplies
the square subcarrier (+1 in the first half chip, -1 in the second).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from .rng import PhiloxStream

# ---------------------------------------------------------------------------
# GPS L1 C/A Gold codes (IS-GPS-200): G1 = 1 + x^3 + x^10,
# G2 = 1 + x^2 + x^3 + x^6 + x^8 + x^9 + x^10, G2 output taps per PRN.
# ---------------------------------------------------------------------------
_L1CA_G2_TAPS = {
    1: (2, 6), 2: (3, 7), 3: (4, 8), 4: (5, 9), 5: (1, 9), 6: (2, 10), 7: (1, 8), 8: (2, 9),
    9: (3, 10), 10: (2, 3), 11: (3, 4), 12: (5, 6), 13: (6, 7), 14: (7, 8), 15: (8, 9),
    16: (9, 10), 17: (1, 4), 18: (2, 5), 19: (3, 6), 20: (4, 7), 21: (5, 8), 22: (6, 9),
    23: (1, 3), 24: (4, 6), 25: (5, 7), 26: (6, 8), 27: (7, 9), 28: (8, 10), 29: (1, 6),
    30: (2, 7), 31: (3, 8), 32: (4, 9),
}


def l1ca_bits(prn: int) -> np.ndarray:
    """1023 code bits (0/1) of GPS L1 C/A PRN `prn`."""
    t1, t2 = _L1CA_G2_TAPS[prn]
    g1 = [1] * 10
    g2 = [1] * 10
    out = np.empty(1023, dtype=np.int8)
    for i in range(1023):
        out[i] = g1[9] ^ g2[t1 - 1] ^ g2[t2 - 1]
        f1 = g1[2] ^ g1[9]
        f2 = g2[1] ^ g2[2] ^ g2[5] ^ g2[7] ^ g2[8] ^ g2[9]
        g1 = [f1] + g1[:9]
        g2 = [f2] + g2[:9]
    return out


# ---------------------------------------------------------------------------
# GPS L5 I5 codes (IS-GPS-705): XA = 1 + x^9 + x^10 + x^12 + x^13 short-cycled
# to 8190 chips, XB = 1 + x + x^3 + x^4 + x^6 + x^7 + x^8 + x^12 + x^13 with a
# per-PRN initial state; code = XA xor XB over 10230 chips. NH10 secondary.
# ---------------------------------------------------------------------------
_L5I_XB_INIT = {
    1: "0101011100100", 2: "1100000110101", 3: "0100000001000", 4: "1011000100110",
    5: "1110111010111", 6: "0110011111010", 7: "1010010011111", 8: "1011110100100",
    9: "1111100101011", 10: "0111111011110",
}
_NH10_BITS = (0, 0, 0, 0, 1, 1, 0, 1, 0, 1)


def _lfsr13(state, taps, nchips, reset_at=None):
    s = list(state)
    out = np.empty(nchips, dtype=np.int8)
    init = list(state)
    for i in range(nchips):
        if reset_at is not None and i > 0 and i % reset_at == 0:
            s = list(init)
        out[i] = s[12]
        fb = 0
        for t in taps:
            fb ^= s[t - 1]
        s = [fb] + s[:12]
    return out


def l5i_bits(prn: int) -> np.ndarray:
    xa = _lfsr13([1] * 13, (9, 10, 12, 13), 10230, reset_at=8190)
    xb = _lfsr13([int(c) for c in _L5I_XB_INIT[prn]], (1, 3, 4, 6, 7, 8, 12, 13), 10230)
    return (xa ^ xb).astype(np.int8)


def l5_xb_state_after(advance: int) -> str:
    """XB register state (stages 1..13) after `advance` clocks from all ones."""
    s = [1] * 13
    for _ in range(advance):
        fb = 0
        for t in (1, 3, 4, 6, 7, 8, 12, 13):
            fb ^= s[t - 1]
        s = [fb] + s[:12]
    return "".join(str(b) for b in s)


def bits_to_chips(bits: np.ndarray) -> np.ndarray:
    """Logic 0 -> +1, logic 1 -> -1."""
    return (1 - 2 * bits.astype(np.int16)).astype(np.int8)


@dataclass
class GNSSSystem:
    """A GNSS signal: code geometry, carrier, data and secondary code."""

    name: str
    code_length: int
    code_frequency: float
    center_frequency: float
    data_frequency: float = 0.0
    secondary: Optional[np.ndarray] = None    # shared secondary code (+1/-1), or None
    secondary_by_prn: Optional[dict] = None   # per-PRN secondary codes
    secondary_baked: bool = False
    boc: bool = False
    _chips: dict = field(default_factory=dict, repr=False)

    # GNSSSignals-style getters -------------------------------------------
    def get_code_length(self) -> int:
        return self.code_length

    def get_code_frequency(self) -> float:
        return self.code_frequency

    def get_center_frequency(self) -> float:
        return self.center_frequency

    def get_data_frequency(self) -> float:
        return self.data_frequency

    def get_code_center_frequency_ratio(self) -> float:
        return self.code_frequency / self.center_frequency

    def get_secondary_code_length(self) -> int:
        if self.secondary_by_prn:
            return len(next(iter(self.secondary_by_prn.values())))
        return 1 if self.secondary is None else len(self.secondary)

    def is_per_prn_secondary(self) -> bool:
        return bool(self.secondary_by_prn)

    def secondary_code(self, prn: int) -> np.ndarray:
        if self.secondary_by_prn:
            return np.asarray(self.secondary_by_prn[prn], dtype=np.int8)
        if self.secondary is None:
            return np.ones(1, dtype=np.int8)
        return np.asarray(self.secondary, dtype=np.int8)

    def secondary_value(self, prn: int, i: int) -> int:
        sec = self.secondary_code(prn)
        return int(sec[i % len(sec)])

    # code generation -------------------------------------------------------
    def _make_chips(self, prn: int) -> np.ndarray:
        raise NotImplementedError

    def primary_chips(self, prn: int) -> np.ndarray:
        if prn not in self._chips:
            self._chips[prn] = self._make_chips(prn)
        return self._chips[prn]

    def gen_code(self, num_samples: int, prn: int, sampling_freq: float, code_frequency: float,
                 start_phase: float = 0.0) -> np.ndarray:
        chips = self.primary_chips(prn)
        n = np.arange(num_samples, dtype=np.float64)
        pos = start_phase + (n * code_frequency) / sampling_freq
        c = np.floor(pos).astype(np.int64)
        out = chips[np.mod(c, self.code_length)].astype(np.int8)
        if self.secondary_baked:
            sec = self.secondary_code(prn)
            per = np.floor_divide(c, self.code_length)
            out = (out * sec[np.mod(per, len(sec))]).astype(np.int8)
        if self.boc:
            frac = pos - np.floor(pos)
            out = np.where(frac < 0.5, out, -out).astype(np.int8)
        return out


class GPSL1CA(GNSSSystem):
    def __init__(self):
        super().__init__("GPSL1CA", 1023, 1.023e6, 1575.42e6, data_frequency=50.0)

    def _make_chips(self, prn):
        return bits_to_chips(l1ca_bits(prn))


class GPSL5I(GNSSSystem):
    def __init__(self):
        super().__init__("GPSL5I", 10230, 10.23e6, 1176.45e6, data_frequency=100.0,
                         secondary=bits_to_chips(np.array(_NH10_BITS)), secondary_baked=True)

    def _make_chips(self, prn):
        return bits_to_chips(l5i_bits(prn))


class SeededCodeSystem(GNSSSystem):
    """Synthetic system with seeded random +-1 primary codes (one Philox stream
    per PRN) and a seeded shared secondary code."""

    def __init__(self, name, code_length, code_frequency, center_frequency, secondary_length,
                 seed, boc=False, data_frequency=0.0):
        sec = None
        if secondary_length > 1:
            sec = PhiloxStream(seed, 0xFFFF).signs(secondary_length)
        super().__init__(name, code_length, code_frequency, center_frequency,
                         data_frequency=data_frequency, secondary=sec, boc=boc)
        self.seed = seed

    def _make_chips(self, prn):
        return PhiloxStream(self.seed, prn).signs(self.code_length)


def E1Like(seed: int = 0x5EED0004) -> SeededCodeSystem:
    """Galileo-E1-like pilot: 4092-chip seeded codes at 1.023 Mchip/s, BOC(1,1),
    seeded 25-chip secondary code, 1575.42 MHz, no data."""
    return SeededCodeSystem("E1Like", 4092, 1.023e6, 1575.42e6, 25, seed, boc=True)


def L5Like(seed: int = 0x5EED0005) -> SeededCodeSystem:
    """L5-like pilot: 10230-chip seeded codes at 10.23 Mchip/s, seeded 10-chip
    secondary code, 1176.45 MHz, no data."""
    return SeededCodeSystem("L5Like", 10230, 10.23e6, 1176.45e6, 10, seed)


def samples_per_code(system: GNSSSystem, sampling_freq: float) -> int:
    """`ceil(Int, get_code_length(system) / get_code_frequency(system) * fs)` (src/plan.jl:386)."""
    return math.ceil(system.get_code_length() / system.get_code_frequency() * sampling_freq)
