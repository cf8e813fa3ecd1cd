"""Oracle of the model systems' plain operator (CPU), pinned by closed forms and the reference's
recorded values.

Pins: the order-8 central Laplacian is exact on a cubic polynomial at points at least four spacings
from the zero halo; the sampled sine ``sin(pi (i + 1) / (n + 1))`` is an eigenvector of the odd-reflection
operator (walls one spacing outside the outermost points) with eigenvalue
``-1/2 * 3 * sum_j c_j cos(j theta) / h^2``, ``theta = pi / (n + 1)`` (a mirror off by one index, or a halo
entry at the wall that is not zero, breaks both); that closed form at n = 49, h = 0.2 is the reference's
recorded ``box_L10`` eigenvalue (``reports/fingerprint_cpu_reference.json``) to 1e-13 Ha; the oracle's
CheFSI on a coarse infinite well converges to the closed form, and on a coarse harmonic well to
within the discretisation error of the analytic 3/2 Ha.
"""

from __future__ import annotations

import json
import pathlib

import numpy as np
import pytest

from oracle.core import Grid, fd_coefficients
from oracle.eigen import EigenConfig
from oracle.model import harmonic_potential, plain_laplacian, solve_plain, well_eigenvalue

FINGERPRINT = json.loads((pathlib.Path(__file__).parent / "golden" / "fingerprint_cpu_reference.json").read_text())
MODEL_EIGEN = EigenConfig(n_extra_states=6, chebyshev_degree=16, max_iterations=200, residual_tol=1.0e-8)


def test_laplacian_is_exact_on_a_cubic_away_from_the_halo():
    shape, h = (14, 15, 16), 0.3
    axes = [np.arange(n) * h - 1.0 for n in shape]
    x, y, z = np.meshgrid(*axes, indexing="ij")
    f = (x**3 - 2 * x * y * y + 0.5 * z**3 + x * z).reshape(-1)
    want = (6 * x - 4 * x + 3 * z).reshape(-1)
    got = plain_laplacian(f[None], shape, h, 8, "zero")[0].reshape(shape)
    inner = (slice(4, -4),) * 3
    np.testing.assert_allclose(got[inner], want.reshape(shape)[inner], rtol=0, atol=1e-10)
    # next to the zero halo the stencil sees zeros: not the polynomial's Laplacian
    assert abs(got[0, 7, 7] - want.reshape(shape)[0, 7, 7]) > 1.0


@pytest.mark.parametrize("order", [2, 4, 8, 12])
def test_sine_is_an_eigenvector_of_the_odd_reflection_operator(order):
    shape, h = (13, 14, 17), 0.25
    s = [np.sin(np.pi * (np.arange(n) + 1) / (n + 1)) for n in shape]
    f = np.einsum("i,j,k->ijk", *s).reshape(-1)
    got = plain_laplacian(f[None], shape, h, order, "odd")[0]
    coeffs = fd_coefficients(2, order)
    r = order // 2
    symbol = sum(sum(c * np.cos((j - r) * np.pi / (n + 1)) for j, c in enumerate(coeffs)) for n in shape) / h**2
    np.testing.assert_allclose(got, symbol * f, rtol=0, atol=1e-12 * abs(symbol))
    # with the zero halo the sine is not an eigenvector (the walls are what makes it one)
    zero = plain_laplacian(f[None], shape, h, order, "zero")[0]
    assert order == 2 or np.abs(zero - symbol * f).max() > 1e-6


def test_well_closed_form_is_the_reference_box_value():
    ref = float(FINGERPRINT["box_L10"]["eigenvalues"][0])
    assert abs(well_eigenvalue(49, 0.2, 8) - ref) < 1e-13
    # and tends to the continuum 3 pi^2 / (2 L^2) of the 10-bohr well
    assert abs(well_eigenvalue(49, 0.2, 8) - 3 * np.pi**2 / 200) < 1e-6


def test_oracle_solves_of_coarse_model_systems():
    grid = Grid((15, 15, 15), 0.5, (-3.5, -3.5, -3.5))
    result, _ = solve_plain(grid, np.zeros(grid.n_points), "odd", MODEL_EIGEN)
    assert result.converged
    assert abs(result.eigenvalues[0] - well_eigenvalue(15, 0.5, 8)) < 1e-12
    grid = Grid((25, 25, 25), 0.5, (-6.0, -6.0, -6.0))
    result, _ = solve_plain(grid, harmonic_potential(grid, 1.0), "zero", MODEL_EIGEN)
    assert result.converged and abs(result.eigenvalues[0] - 1.5) < 2e-3
