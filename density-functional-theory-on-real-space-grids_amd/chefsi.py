"""The scalars of the Chebyshev filter, tabulated for the fused step kernel.

The filter damps the unwanted part of the spectrum, ``[lower, upper]``, with the Chebyshev polynomial
of that interval normalised at ``spectrum_min`` (Zhou and Saad, "Self-consistent-field calculations
using Chebyshev-filtered subspace iteration", J. Comput. Phys. 219 (2006), Algorithm 4.1). Each step
of its three-term recurrence needs three numbers; they are built here once per filter on the host, in
the reference's float64 operation order (``src/cdft/eigen/chefsi.py:116-136``, so that device and CPU
filters agree bitwise), and uploaded as a ``(degree, 3)`` device table that the fused kernel reads by
step index. A whole filter is then ``degree`` launches with no host synchronisation.
"""

from __future__ import annotations


def scaled_chebyshev_table(degree: int, lower: float, upper: float, spectrum_min: float) -> list[tuple[float, float, float]]:
    """Rows ``(centre, gain, damping)`` of the ``degree`` filter steps.

    Step 0 forms ``y_1 = (A x - centre x) gain_0``; step ``j > 0`` forms
    ``y_{j+1} = (A y_j - centre y_j) gain_j - damping_j y_{j-1}``. With ``half`` the half-width of
    ``[lower, upper]`` and ``r_0 = half / (spectrum_min - centre)``, the ratios follow
    ``r_j = 1 / (2 / r_0 - r_{j-1})``, ``gain_j = 2 r_j / half`` and ``damping_j = r_{j-1} r_j``
    (``gain_0 = r_0 / half``)."""
    if not upper > lower:
        raise ValueError(f"Chebyshev filter needs a damped interval with upper > lower, got [{lower}, {upper}]")
    half = 0.5 * (upper - lower)
    centre = 0.5 * (upper + lower)
    ratio = half / (spectrum_min - centre)
    first = ratio
    table = [(centre, first / half, 0.0)]
    for _ in range(1, degree):
        following = 1.0 / (2.0 / first - ratio)
        table.append((centre, 2.0 * following / half, ratio * following))
        ratio = following
    return table


#: The reference's name for the same rows (``eigen/chefsi.py:116``), kept for callers written against it.
chebyshev_coefficients = scaled_chebyshev_table
