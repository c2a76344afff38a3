"""Counter-based random numbers for seeded synthetic inputs.

Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
as 1, 2, 3", SC'11). A counter-based generator gives the same stream for a
given (seed, stream id, counter) on any device, so the same sample arrays can
be produced on the CPU and, later, on the GPU. Uniform doubles use 53 bits of
two 32-bit words; normals use Box-Muller in FP64.
"""
from __future__ import annotations

import numpy as np

_M0 = np.uint64(0xD2511F53)
_M1 = np.uint64(0xCD9E8D57)
_W0 = np.uint32(0x9E3779B9)
_W1 = np.uint32(0xBB67AE85)
_MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Vectorised Philox4x32-10 on uint32 arrays; returns four uint32 arrays."""
    x0 = np.asarray(c0, dtype=np.uint32).copy()
    x1 = np.asarray(c1, dtype=np.uint32).copy()
    x2 = np.asarray(c2, dtype=np.uint32).copy()
    x3 = np.asarray(c3, dtype=np.uint32).copy()
    k0 = np.uint32(k0)
    k1 = np.uint32(k1)
    with np.errstate(over="ignore"):
        for rnd in range(10):
            if rnd:
                k0 = np.uint32((int(k0) + int(_W0)) & 0xFFFFFFFF)
                k1 = np.uint32((int(k1) + int(_W1)) & 0xFFFFFFFF)
            p0 = x0.astype(np.uint64) * _M0
            p1 = x2.astype(np.uint64) * _M1
            hi0 = (p0 >> np.uint64(32)).astype(np.uint32)
            lo0 = (p0 & _MASK).astype(np.uint32)
            hi1 = (p1 >> np.uint64(32)).astype(np.uint32)
            lo1 = (p1 & _MASK).astype(np.uint32)
            x0, x1, x2, x3 = hi1 ^ x1 ^ k0, lo1, hi0 ^ x3 ^ k1, lo0
    return x0, x1, x2, x3


class PhiloxStream:
    """Sequential view of one Philox stream: key = seed, counter = (i, stream)."""

    def __init__(self, seed: int, stream: int = 0):
        self.k0 = seed & 0xFFFFFFFF
        self.k1 = (seed >> 32) & 0xFFFFFFFF
        self.stream = stream & 0xFFFFFFFF
        self.pos = 0  # counter blocks consumed

    def _blocks(self, nblocks: int):
        i = np.arange(self.pos, self.pos + nblocks, dtype=np.uint64)
        self.pos += nblocks
        lo = (i & _MASK).astype(np.uint32)
        hi = (i >> np.uint64(32)).astype(np.uint32)
        s = np.full(nblocks, self.stream, dtype=np.uint32)
        z = np.zeros(nblocks, dtype=np.uint32)
        return philox4x32_10(lo, hi, s, z, self.k0, self.k1)

    def uniform(self, n: int) -> np.ndarray:
        """n doubles in [0, 1)."""
        nb = (n + 1) // 2
        x0, x1, x2, x3 = self._blocks(nb)
        a = ((x0 >> np.uint32(5)).astype(np.float64) * 67108864.0 + (x1 >> np.uint32(6)).astype(np.float64))
        b = ((x2 >> np.uint32(5)).astype(np.float64) * 67108864.0 + (x3 >> np.uint32(6)).astype(np.float64))
        out = np.empty(2 * nb, dtype=np.float64)
        out[0::2] = a
        out[1::2] = b
        return out[:n] * (1.0 / 9007199254740992.0)

    def normal(self, n: int) -> np.ndarray:
        """n standard normal doubles (Box-Muller)."""
        m = (n + 1) // 2
        u = self.uniform(2 * m)
        u1 = 1.0 - u[0::2]  # (0, 1]
        u2 = u[1::2]
        r = np.sqrt(-2.0 * np.log(u1))
        t = 2.0 * np.pi * u2
        out = np.empty(2 * m, dtype=np.float64)
        out[0::2] = r * np.cos(t)
        out[1::2] = r * np.sin(t)
        return out[:n]

    def signs(self, n: int) -> np.ndarray:
        """n values of +1/-1 (int8)."""
        u = self.uniform(n)
        return np.where(u < 0.5, 1, -1).astype(np.int8)

    def integers(self, n: int, lo: int, hi: int) -> np.ndarray:
        """n integers uniform in [lo, hi)."""
        u = self.uniform(n)
        return (lo + np.floor(u * (hi - lo))).astype(np.int64)
