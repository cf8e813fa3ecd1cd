"""Chebyshev-filtered subspace iteration driving the HIP kernels.

Mirrors ``src/cdft/eigen/chefsi.py`` of the reference (``chebyshev_coefficients`` 116-136,
``ChebyshevFilteredSubspace.solve`` 442-585): the recurrence scalars are built on the host in the
same float64 operations and uploaded as a ``(degree, 3)`` device table read by the fused step
kernel, so a filter is ``degree`` launches with no host synchronisation.
"""

from __future__ import annotations


def chebyshev_coefficients(degree: int, lower: float, upper: float, spectrum_min: float) -> list[tuple[float, float, float]]:
    """Rows ``(c, a_k, b_k)``: step 1 ``y = (A x - c x) a_1``; step k ``y_k = (A y - c y) a_k - b_k y_{k-1}``."""
    if upper <= lower:
        raise ValueError(f"filter interval is empty: lower={lower}, upper={upper}")
    e = 0.5 * (upper - lower)
    c = 0.5 * (upper + lower)
    sigma = e / (spectrum_min - c)
    sigma1 = sigma
    rows = [(c, sigma1 / e, 0.0)]
    for _ in range(2, degree + 1):
        sigma2 = 1.0 / (2.0 / sigma1 - sigma)
        rows.append((c, 2.0 * sigma2 / e, sigma * sigma2))
        sigma = sigma2
    return rows
