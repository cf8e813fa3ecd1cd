"""Oracle of the analytic model systems -- TEST INFRASTRUCTURE ONLY (the checker, never the product).

The plain operator ``H = -1/2 lap + V`` (reference ``src/cdft/operators/hamiltonian.py:92-125``) with
the grid's order-``p`` central Laplacian (``src/cdft/operators/laplacian.py:85-105``: per axis 0, 1, 2,
offsets -R..R, each term scaled by ``c_j * (1/h^2)``) over a halo filled per boundary mode
(``src/cdft/grid.py:714-733,808-827``): zero, or odd reflection about walls one spacing outside the
outermost points. The uniform measure ``h^3`` (``src/cdft/eigen/measure.py``). Solved by this oracle's
CheFSI.

Pins (``tests/test_model.py``): the Laplacian is exact on a quadratic away from the zero halo; the
sampled sine ``sin(pi (i + 1) / (n + 1))`` is an eigenvector of the odd-reflection operator to round-off
with the closed-form eigenvalue ``-1/2 * 3 * sum_j c_j cos(j theta) / h^2``, ``theta = pi / (n + 1)``,
which also reproduces the reference's recorded ``box_L10`` eigenvalue; the harmonic well's lowest
eigenvalue on a coarse grid lies within the discretisation error of the analytic 3/2.
"""

from __future__ import annotations

import numpy as np

from .core import Grid, Measure, fd_coefficients


def plain_laplacian(field: np.ndarray, shape, spacing: float, order: int, boundary: str) -> np.ndarray:
    """The order-``order`` central Laplacian of ``(..., N)`` fields, the halo per ``boundary``
    (``"zero"`` or ``"odd"``)."""
    coeffs = fd_coefficients(2, order)
    reach = order // 2
    scale = 1.0 / (spacing * spacing)
    lead = field.shape[:-1]
    box = np.asarray(field, dtype=np.float64).reshape(*lead, *shape)
    pad = [(0, 0)] * len(lead) + [(reach, reach)] * 3
    padded = np.pad(box, pad)
    if boundary == "odd":
        for axis in range(3):
            ax = len(lead) + axis
            n = shape[axis]
            for m in range(1, reach):
                # low side: interior index -1 - m mirrors m - 1; high side: n + m mirrors n - m
                lo_dst = [slice(None)] * padded.ndim
                lo_src = [slice(None)] * padded.ndim
                lo_dst[ax], lo_src[ax] = reach - 1 - m, reach + m - 1
                padded[tuple(lo_dst)] = -padded[tuple(lo_src)]
                hi_dst = [slice(None)] * padded.ndim
                hi_src = [slice(None)] * padded.ndim
                hi_dst[ax], hi_src[ax] = reach + n + m, reach + n - m
                padded[tuple(hi_dst)] = -padded[tuple(hi_src)]
    elif boundary != "zero":
        raise ValueError(f"boundary {boundary!r}: 'zero' or 'odd'")
    out = None
    for axis in range(3):
        ax = len(lead) + axis
        for j, c in enumerate(coeffs):
            sl = [slice(None)] * len(lead) + [slice(reach, reach + n) for n in shape]
            sl[ax] = slice(j, j + shape[axis])
            term = padded[tuple(sl)] * (c * scale)
            out = term if out is None else out + term
    return out.reshape(*lead, -1)


class PlainHamiltonian:
    """``-1/2 lap + V`` with the uniform measure (reference ``LocalHamiltonian``)."""

    uses_divergence_form = False

    def __init__(self, grid: Grid, potential, boundary: str, order: int = 8):
        self.grid = grid
        self.potential = np.asarray(potential, dtype=np.float64)
        self.boundary = boundary
        self.order = order
        self.measure = Measure(grid)
        self.apply_count = 0

    def kinetic_apply(self, phi):
        return plain_laplacian(phi, self.grid.shape, self.grid.spacing, self.order, self.boundary) * -0.5

    def apply(self, phi):
        self.apply_count += 1
        return self.kinetic_apply(phi) + np.asarray(phi) * self.potential

    def spectral_bounds(self, steps=12, safety=1.02, generator=None):
        from .eigen import lanczos_bounds, seeded_randn

        start = seeded_randn(self.grid.n_points, generator)
        return lanczos_bounds(self.apply, start, self.measure.inner, self.measure.norm, steps, safety)


def harmonic_potential(grid: Grid, omega: float) -> np.ndarray:
    """``1/2 omega^2 r^2`` at the grid points (reference ``operators/external.py:51-53``)."""
    axes = [np.arange(n) * grid.spacing + o for n, o in zip(grid.shape, grid.origin)]
    x, y, z = np.meshgrid(*axes, indexing="ij")
    return (0.5 * omega**2 * ((x * x + y * y) + z * z)).reshape(-1)


def well_eigenvalue(n: int, spacing: float, order: int) -> float:
    """The lowest eigenvalue of the odd-reflection operator on ``n^3`` points: the sampled sine
    ``sin(theta (i + 1))``, ``theta = pi / (n + 1)``, meets the walls' odd halo exactly, so the stencil
    acts as on the infinite lattice: ``-1/2 * 3 * sum_j c_j cos(j theta) / h^2``."""
    coeffs = fd_coefficients(2, order)
    reach = order // 2
    theta = np.pi / (n + 1)
    symbol = sum(c * np.cos((j - reach) * theta) for j, c in enumerate(coeffs))
    return -0.5 * 3.0 * symbol / (spacing * spacing)


def solve_plain(grid: Grid, potential, boundary: str, config, n_states: int = 1, seed: int = 0, order: int = 8):
    """The oracle's CheFSI on the plain operator (reference ``scf/noninteracting.py:118-190``)."""
    from .eigen import chefsi_solve, make_generator

    ham = PlainHamiltonian(grid, potential, boundary, order)
    return chefsi_solve(ham, grid, config, n_states, generator=make_generator(seed)), ham
