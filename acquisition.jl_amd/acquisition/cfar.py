"""CFAR detection threshold (src/cfar.jl:64-85)."""
from __future__ import annotations

from scipy.special import gammainccinv

from .plan import ArgumentError


def cfar_threshold(pfa, num_cells, *, num_noncoherent_integrations=1) -> float:
    """Threshold on `peak_to_noise_ratio` such that any of `num_cells` noise-only
    cells exceeds it with probability `pfa` (Bonferroni-like per-cell pfa).

    Under H0 M * peak_to_noise_ratio ~ Gamma(M, 1) with M the number of
    noncoherent integrations, so the threshold is the upper-tail quantile of
    Gamma(M, 1) at the per-cell pfa, divided by M. The reference calls
    SpecialFunctions.gamma_inc_inv(M, 1 - p, p); scipy's gammainccinv(M, p)
    is the same inverse of the regularised upper incomplete gamma.
    """
    if not (0 < pfa < 1):
        raise ArgumentError(f"pfa must be in (0, 1), got {pfa}")
    if not (num_cells > 0):
        raise ArgumentError(f"num_cells must be positive, got {num_cells}")
    if not (num_noncoherent_integrations > 0):
        raise ArgumentError("num_noncoherent_integrations must be positive")
    pfa_per_cell = 1 - (1 - pfa) ** (1 / num_cells)
    a = float(num_noncoherent_integrations)
    x = gammainccinv(a, pfa_per_cell)
    return float(x / a)
