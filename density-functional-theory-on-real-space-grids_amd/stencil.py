"""Exact-rational finite-difference weights for the staggered divergence form and central gradients.

Same definitions as the reference (``src/cdft/operators/divergence.py:30-74`` and
``src/cdft/operators/laplacian.py:35-75``): the weights solve a Vandermonde system in rational
arithmetic and are rounded once to float64.
"""

from __future__ import annotations

import math
from fractions import Fraction
from functools import lru_cache


def _gauss_jordan(rows: list[list[Fraction]], rhs: list[Fraction]) -> list[Fraction]:
    n = len(rhs)
    for col in range(n):
        piv = next(r for r in range(col, n) if rows[r][col] != 0)
        rows[col], rows[piv] = rows[piv], rows[col]
        rhs[col], rhs[piv] = rhs[piv], rhs[col]
        inv = rows[col][col]
        rows[col] = [v / inv for v in rows[col]]
        rhs[col] /= inv
        for r in range(n):
            if r != col and rows[r][col] != 0:
                f = rows[r][col]
                rows[r] = [a - f * b for a, b in zip(rows[r], rows[col])]
                rhs[r] -= f * rhs[col]
    return rhs


@lru_cache(maxsize=None)
def staggered_coefficients(order: int) -> tuple[float, ...]:
    """Weights ``c_k``, ``k = -p .. p-1``, of the face derivative ``f'(x + h/2) ~ h^-1 sum c_k f(x_{1+k})``."""
    if order % 2 or order < 2:
        raise ValueError(f"order must be a positive even integer, got {order}")
    p = order // 2
    nodes = [Fraction(2 * k + 1, 2) for k in range(-p, p)]
    rows = [[x**m for x in nodes] for m in range(order)]
    rhs = [Fraction(int(m == 1)) for m in range(order)]
    return tuple(float(v) for v in _gauss_jordan(rows, rhs))


@lru_cache(maxsize=None)
def central_coefficients(derivative: int, order: int) -> tuple[float, ...]:
    """Central weights for ``d^n/dx^n`` at offsets ``-hw .. hw``, ``hw = order/2 + (n-1)/2``."""
    hw = order // 2 + (derivative - 1) // 2
    offsets = list(range(-hw, hw + 1))
    rows = [[Fraction(o) ** m for o in offsets] for m in range(len(offsets))]
    rhs = [Fraction(math.factorial(derivative) if m == derivative else 0) for m in range(len(offsets))]
    return tuple(float(v) for v in _gauss_jordan(rows, rhs))
