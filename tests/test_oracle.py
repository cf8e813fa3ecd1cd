"""The oracle pinned against the reference's own golden values, closed forms and brute force.

Every value below is copied from the reference's data files (values only): ``tests/golden/values.json``,
``docs/figures/convergence_ladders.json`` and ``reports/fingerprint_cpu_reference.json``. Each check
is chosen so that a plausible slip (a wrong tap index, a dropped cell average, a transposed weight)
moves it far outside its tolerance.
"""

from __future__ import annotations

import json
import math

import numpy as np
import pytest

from conftest import ROOT
from oracle.core import (
    CuspFactor, CuspHamiltonian, Grid, Measure, fd_coefficients, staggered_coefficients, weighted_divergence,
)
from oracle.eigen import chebyshev_coefficients, chebyshev_filter

GOLDEN = json.loads((ROOT / "tests" / "golden" / "reference_values.json").read_text())


def _small_operator(charges, spacing=0.5, length=6.0):
    grid = Grid.from_box((length, length, length), spacing)
    positions = [[0.0, 0.0, 0.0] for _ in charges]
    if len(charges) == 2:
        positions = [[0.0, 0.0, -1.0], [0.0, 0.0, 1.0]]
    factor = CuspFactor(grid, charges, positions)
    return grid, CuspHamiltonian(grid, factor)


def test_staggered_weights_closed_forms():
    assert staggered_coefficients(2) == (-1.0, 1.0)
    assert tuple(round(c, 10) for c in staggered_coefficients(4)) == (0.0416666667, -1.125, 1.125, -0.0416666667)
    assert fd_coefficients(2, 2) == (1.0, -2.0, 1.0)
    assert fd_coefficients(1, 2) == (-0.5, 0.0, 0.5)
    for order in (2, 4, 6, 8, 12):
        c = np.array(staggered_coefficients(order))
        k = np.arange(-order // 2, order // 2) + 0.5
        for m in range(order):
            # exact for x^m, m < order; the bound scales with the terms (5.5^11 ~ 1e8 at order 12)
            scale = max(1.0, float(np.abs(c * k**m).sum()))
            assert abs((c * k**m).sum() - (1.0 if m == 1 else 0.0)) < 1e-14 * scale
        # and not for x^(order+1) (x^order holds by the antisymmetry of c): the order is exactly this
        assert abs((c * k ** (order + 1)).sum()) > 1e-6


@pytest.mark.parametrize("order", [2, 4, 6, 8, 12])
def test_transpose_is_exact_and_composite_is_spd(order):
    """Assemble ``D^T w D`` column by column on a 1-D line: symmetric to round-off, PSD, and for
    ``order = 2``, ``w = 1`` the 3-point Laplacian."""
    size, h = 14, 0.37
    rng = np.random.default_rng(11)
    w = rng.uniform(0.5, 1.5, size=size + order)
    ones = np.ones(size + order)
    for weight in (w, ones):
        mat = np.zeros((size, size))
        for col in range(size):
            e = np.zeros((size, 1, 1))
            e[col] = 1.0
            faces = [weight[:, None, None], np.zeros((size, 1 + order, 1)), np.zeros((size, 1, 1 + order))]
            mat[:, col] = weighted_divergence(e, faces, h, order)[:, 0, 0]
        assert np.abs(mat - mat.T).max() < 1e-12 * np.abs(mat).max()
        assert np.linalg.eigvalsh(mat).min() > -1e-10 * np.abs(mat).max()
    if order == 2:
        lap = (2 * np.eye(size) - np.eye(size, k=1) - np.eye(size, k=-1)) / h**2
        assert np.abs(mat - lap).max() < 1e-12 * np.abs(lap).max()


@pytest.mark.parametrize("charge", [1.0, 2.0])
def test_checkerboard_rayleigh_quotient_pins(charge):
    grid, h = _small_operator((charge,))
    alt = ((-1.0) ** grid.box_index().sum(-1))[None, :].astype(float)
    alt = alt / h.measure.norm(alt)[0]
    q = float(h.measure.cross(alt, h.apply(alt))[0, 0])
    pin = GOLDEN["checkerboard.rayleigh.Z%g" % charge]
    assert abs(q - pin) <= 1e-9 * abs(pin)


@pytest.mark.parametrize("charges", [(1.0,), (2.0,), (1.0, 1.0)])
def test_self_adjoint_in_weighted_measure(charges):
    grid, h = _small_operator(charges)
    rng = np.random.default_rng(7)
    a, b = rng.standard_normal((2, 1, grid.n_points))
    left = float(h.measure.cross(a, h.apply(b))[0, 0])
    right = float(h.measure.cross(b, h.apply(a))[0, 0])
    assert abs(left - right) / max(abs(left), abs(right)) < 1e-13


@pytest.mark.parametrize("charge", [1.0, 2.0, 8.0])
def test_single_nucleus_potential_is_constant(charge):
    grid = Grid.from_box((6.0, 6.0, 6.0), 0.5)
    factor = CuspFactor(grid, (charge,), [[0.0, 0.0, 0.0]])
    assert np.abs(factor.transformed_potential + 0.5 * charge**2).max() < 1e-10 * charge**2


def test_identity_factor_is_the_plain_laplacian():
    """No nuclei: ``f = 1``, faces all one; the operator is ``-1/2`` times the composite second
    difference, i.e. applied to a quadratic it returns the exact ``-1/2 * 6`` inside."""
    grid = Grid.from_box((8.0, 8.0, 8.0), 0.5)
    factor = CuspFactor(grid, (), np.zeros((0, 3)))
    h = CuspHamiltonian(grid, factor)
    pts = grid.points()
    quad = (pts**2).sum(-1)[None, :]
    out = h.kinetic_apply(quad).reshape(grid.shape)
    interior = out[8:-8, 8:-8, 8:-8]
    assert np.abs(interior - (-3.0)).max() < 1e-9  # kinetic part = -1/2 lap(r^2) = -3


def test_chebyshev_recurrence_is_the_chebyshev_polynomial():
    """On a diagonal operator the filter multiplies eigenvector i by ``T_d((l_i-c)/e) / T_d((s-c)/e)``."""
    lam = np.linspace(-1.0, 5.0, 7)

    class Diag:
        def apply(self, x):
            return x * lam

    d, lower, upper, smin = 9, 0.5, 5.0, -1.0
    x = np.ones((1, lam.size))
    y = chebyshev_filter(Diag(), x, d, lower, upper, smin)
    e, c = 0.5 * (upper - lower), 0.5 * (upper + lower)
    t = np.polynomial.chebyshev.Chebyshev.basis(d)
    expect = t((lam - c) / e) / t((smin - c) / e)
    assert np.abs(y[0] - expect).max() < 1e-12 * np.abs(expect).max()
    assert len(chebyshev_coefficients(d, lower, upper, smin)) == d


def test_cusp_weight_closed_forms():
    grid = Grid.from_box((16.0, 16.0, 16.0), 0.5)
    one = CuspFactor(grid, (2.0,), [[0.0, 0.0, 0.0]])
    assert one.weight_integral_exact() == pytest.approx(math.pi / 8.0)
    two = CuspFactor(grid, (1.0, 1.0), [[0.0, 0.0, -1.0], [0.0, 0.0, 1.0]])
    # The two-centre closed form tends to the united-atom value pi/8 as R -> 0 (checked numerically).
    tiny = CuspFactor(grid, (1.0, 1.0), [[0.0, 0.0, -1e-4], [0.0, 0.0, 1e-4]])
    assert tiny.weight_integral_exact() == pytest.approx(math.pi / 8.0, rel=1e-3)
    assert 0.0 < two.weight_integral_exact() < math.pi / 8.0


def test_bare_hydrogen_golden():
    """Bare H, 33^3 at h = 1 (the derived single-centre grid), k = 5, degree 20, CheFSI from the
    seeded normal start: eigenvalue, kinetic energy and iteration count of the reference."""
    from oracle.physics import Numerics, Scenario, solve_noninteracting

    r = solve_noninteracting(Scenario((1.0,), ((0.0, 0.0, 0.0),), 1.0), Numerics())
    assert r["n_iterations"] == GOLDEN["fingerprint.h_atom.n_iterations"]
    assert abs(r["eigenvalues"][0] - GOLDEN["eigenvalue.ground.h_atom"]) < 1e-9 * 0.5
    assert abs(r["kinetic"] - GOLDEN["kinetic.h_atom"]) < 1e-9 * 0.5


@pytest.mark.slow
def test_helium_lda_ladder_rung_h05():
    """He at VWN LDA on the 16-bohr box at h = 0.5 (33^3): the reference ladder's total energy and
    SCF iteration count. This pins the quadrature, Poisson, LDA and SCF oracle together."""
    from oracle.physics import Numerics, Scenario, solve_self_consistent

    r = solve_self_consistent(Scenario((2.0,), ((0.0, 0.0, 0.0),), 2.0, "lda"),
                              Numerics(box_lengths=(16.0, 16.0, 16.0), spacing=0.5))
    pin = GOLDEN["ladder.he_lda.h0.5.total"]
    assert abs(r["energies"]["total"] - pin) < 1e-11
    assert r["scf_iterations"] == GOLDEN["ladder.he_lda.h0.5.scf_iterations"]


def test_h2plus_ladder_rung_h05():
    """Bare H2+ at R = 2 (protons at z = -1, +1) on the reference's derived grid for h = 0.5: box
    edge = nuclear span + 12 bohr vacuum = 14, lattice anchored on the first nucleus (29^3). The
    two spheres overlap, so this pins the overlapping-sphere quadrature (Becke cells, lune, 0.7
    angular node spacing) and the lumped far mass weights of ``quadrature.py:421-497`` against the
    reference ladder's eigenvalue."""
    from oracle.physics import Numerics, Scenario, solve_noninteracting

    sc = Scenario((1.0, 1.0), ((0.0, 0.0, -1.0), (0.0, 0.0, 1.0)), 1.0)
    r = solve_noninteracting(sc, Numerics(box_lengths=(14.0, 14.0, 14.0), spacing=0.5), anchor=(0.0, 0.0, -1.0))
    assert r["hamiltonian"].quadrature.far_lumped
    assert abs(r["eigenvalues"][0] - GOLDEN["ladder.h2plus_R2.h0.5.eigenvalue"]) < 1e-11


def test_pbe_limits_of_the_oracle():
    """The oracle's PBE (reference ``xc/gga.py:43-90``) against what its formulas fix independently:
    no gradient gives Slater exchange plus the modified PW92 correlation, which agrees with VWN5 (a
    separate fit of the same Ceperley-Alder data, pinned by the LDA fixtures) to 1 mHa per electron;
    the exchange enhancement tends to 1 + kappa (the Lieb-Oxford bound) at large reduced gradient and
    to 1 + mu s^2 at small; the correlation vanishes as t -> infinity and its gradient term is
    beta t^2 at small t (PBE eqs. 7, 8, 14)."""
    from oracle import physics as op

    kf = (3 * math.pi**2) ** (1 / 3)
    slater = 0.75 * (3 / math.pi) ** (1 / 3)
    n = 10 ** np.linspace(-3, 1, 40)
    e0, _, _ = op.pbe_xc(n, np.zeros_like(n))
    elda, _ = op.lda_vwn(n)
    assert np.max(np.abs(e0 - elda) / n) < 1e-3
    one = np.array([0.37])
    big, _, _ = op.pbe_xc(one, np.array([1e40]))
    assert abs(big[0] / (-slater * one[0] ** (4 / 3) * (1 + 0.804)) - 1) < 1e-10
    # small gradient: the exchange term -C n^{4/3} mu s^2 and the correlation term n beta t^2 cancel
    # (mu = beta pi^2 / 3 is chosen so, PBE eq. 12), leaving a change of second order in s^2
    s2 = 1e-7
    sigma = s2 * 4 * kf**2 * one ** (8 / 3)
    t2 = sigma * math.pi / (16 * kf * one ** (7 / 3))
    es, _, _ = op.pbe_xc(one, sigma)
    e_flat, _, _ = op.pbe_xc(one, np.array([0.0]))
    x_term = -slater * one[0] ** (4 / 3) * 0.2195149727645171 * s2
    c_term = one[0] * 0.06672455060314922 * t2[0]
    assert abs(x_term + c_term) < 1e-12 * abs(x_term)
    assert abs((es - e_flat)[0]) < 1e-4 * abs(x_term)
    # and the correlation alone: with exchange held at the uniform gas (s^2 -> 0 only in F), the
    # gradient term is n beta t^2; recovered from the total by removing the exchange enhancement
    ex_s = -slater * one[0] ** (4 / 3) * (1 + 0.804 - 0.804 / (1 + 0.2195149727645171 * s2 / 0.804))
    ex_0 = -slater * one[0] ** (4 / 3)
    assert abs(((es - e_flat)[0] - (ex_s - ex_0)) / c_term - 1) < 1e-5
    # correlation vanishes at large t: the total tends to the exchange limit alone
    huge = np.array([1e12]) * 4 * kf**2 * one ** (8 / 3)
    eh, _, _ = op.pbe_xc(one, huge)
    ex_h = -slater * one ** (4 / 3) * (1 + 0.804 - 0.804 / (1 + 0.2195149727645171 * 1e12 / 0.804))
    assert abs(eh[0] - ex_h[0]) < 1e-8 * abs(e_flat[0])


def test_pbe_derivatives_of_the_oracle_match_differences():
    """``de/dn`` and ``de/dsigma`` (by autograd, as the reference takes them) against central differences
    at 2000 seeded points, relative step 1e-5, reduced gradients s^2 in [1e-2, 10]."""
    from oracle import physics as op

    kf = (3 * math.pi**2) ** (1 / 3)
    rng = np.random.default_rng(0)
    n = 10 ** rng.uniform(-4, 2, 2000)
    sig = 10 ** rng.uniform(-2, 1, 2000) * 4 * kf**2 * n ** (8 / 3)
    e, vn, vs = op.pbe_xc(n, sig)
    h = 1e-5
    fds = (op.pbe_xc(n, sig * (1 + h))[0] - op.pbe_xc(n, sig * (1 - h))[0]) / (2 * h * sig)
    fdn = (op.pbe_xc(n * (1 + h), sig)[0] - op.pbe_xc(n * (1 - h), sig)[0]) / (2 * h * n)
    assert (np.abs(fds - vs) * sig / np.abs(e)).max() < 1e-9
    assert (np.abs(fdn - vn) * n / np.abs(e)).max() < 1e-9


def test_tpa_filter_on_a_plane_wave():
    """One Fourier mode of the box is an eigenvector of the preconditioner with eigenvalue
    ``K(x) = P/(P + 16 x^4)``, ``x = (G^2/2)/E_ref``; ``K(0) = 1`` and K falls as ``27/(16 x^4)``."""
    from oracle.eigen import tpa_filter, tpa_precondition

    grid = Grid.from_box((7.0, 9.0, 5.0), 0.5)  # 15 x 19 x 11: odd and mixed sizes
    n0, n1, n2 = grid.shape
    i, j, k = np.meshgrid(np.arange(n0), np.arange(n1), np.arange(n2), indexing="ij")
    m = (2, 3, 1)
    wave = np.cos(2 * np.pi * (m[0] * i / n0 + m[1] * j / n1 + m[2] * k / n2)).reshape(1, -1)
    g2 = sum((2 * np.pi * mm / (n * grid.spacing)) ** 2 for mm, n in zip(m, grid.shape))
    # E_ref: the mean of G^2/2 over the half spectrum, from the frequencies' closed forms
    f0 = np.fft.fftfreq(n0, grid.spacing)
    f1 = np.fft.fftfreq(n1, grid.spacing)
    f2 = np.arange(n2 // 2 + 1) / (n2 * grid.spacing)
    e_ref = 0.5 * (2 * np.pi) ** 2 * ((f0**2).mean() + (f1**2).mean() + (f2**2).mean())
    x = 0.5 * g2 / e_ref
    poly = 27 + 18 * x + 12 * x**2 + 8 * x**3
    out = tpa_precondition(wave, grid)
    assert np.abs(out - wave * poly / (poly + 16 * x**4)).max() < 1e-13
    k_table = tpa_filter(grid.shape, grid.spacing)
    assert k_table[0, 0, 0] == 1.0 and np.all(np.diff(k_table[0, 0, :]) < 0)


def test_similarity_frame_preconditioner_is_symmetric_in_the_weighted_measure():
    """The reference's check (``tests/test_cusp_factorisation.py:241-262``): ``<a, f^-1 K f b>_{f^2}``
    is symmetric, the plain frame ``<a, K b>_{f^2}`` is not."""
    from oracle.eigen import tpa_precondition

    grid = Grid.from_box((8.0, 8.0, 8.0), 0.5)
    factor = CuspFactor(grid, (1.0,), [[0.0, 0.0, 0.0]])
    f = np.maximum(np.exp(-factor.u), 1e-150)
    rng = np.random.default_rng(11)
    a, b = rng.standard_normal((2, 1, grid.n_points))
    w = factor.weight
    left = float((w * a * tpa_precondition(b, grid, f)).sum())
    right = float((w * b * tpa_precondition(a, grid, f)).sum())
    assert abs(left - right) < 1e-10 * max(1.0, abs(left))
    plain_left = float((w * a * tpa_precondition(b, grid)).sum())
    plain_right = float((w * b * tpa_precondition(a, grid)).sum())
    assert abs(plain_left - plain_right) > 1e-6 * max(1.0, abs(plain_left))


def test_lobpcg_agrees_with_chefsi_on_bare_hydrogen():
    """The two-solver check (G2.5, ``contract.py:436-449``: |Δε| < ``cross_check_tol`` = 1e-8) on the
    bare-H golden case (33^3, h = 1), LOBPCG seeded with ``seed + 1`` as the reference's SCF does
    (``scf/loop.py:485-494``); and the preconditioner frames rank as the reference's decision D-56
    records (similarity fewest iterations, then none, then transformed)."""
    from oracle import physics as ophys
    from oracle.eigen import lobpcg_solve, make_generator

    sc = ophys.Scenario((1.0,), ((0.0, 0.0, 0.0),), 1.0)
    num = ophys.Numerics(box_lengths=(32.0,) * 3, spacing=1.0)
    geometry = ophys.build_geometry(sc, num)
    grid, factor, q = geometry
    h = CuspHamiltonian(grid, factor, quadrature=q)
    iterations = {}
    for frame in ("similarity", "none", "transformed"):
        r = lobpcg_solve(h, grid, num.eigen, 1, generator=make_generator(num.seed + 1), frame=frame)
        assert r.converged
        assert abs(r.eigenvalues[0] - GOLDEN["eigenvalue.ground.h_atom"]) < 1e-8
        iterations[frame] = r.n_iterations
    assert iterations["similarity"] < iterations["none"] < iterations["transformed"]


def test_rank_deficient_block_falls_back_to_eigh():
    """``orthonormalise`` (reference ``eigen/rayleigh_ritz.py:51-72``) on a block whose rows repeat:
    the Cholesky fails, the eigh fallback keeps the independent directions (eigenvalues above 1e-14 of
    the largest) and returns fewer vectors, orthonormal in the measure and spanning the block; the
    Rayleigh-Ritz values of that span are those of the distinct rows."""
    from oracle.eigen import orthonormalise, rayleigh_ritz

    grid, h = _small_operator((1.0,))
    rng = np.random.default_rng(2)
    rows = rng.standard_normal((3, grid.n_points))
    block = np.concatenate([rows, rows[:1], 2.0 * rows[1:2]])  # rank 3 in 5 rows
    basis = orthonormalise(block, h.measure)
    assert basis.shape[0] == 3
    assert np.abs(h.measure.gram(basis) - np.eye(3)).max() < 1e-12
    # the span is the rows' span: projecting the rows on the basis loses nothing
    proj = h.measure.cross(rows, basis) @ basis
    assert np.abs(proj - rows).max() < 1e-10 * np.abs(rows).max()
    evals, _, orth = rayleigh_ritz(h, block, h.measure)
    evals_rows, _, _ = rayleigh_ritz(h, rows, h.measure)
    assert evals.shape == (3,) and orth < 1e-12
    assert np.allclose(evals, evals_rows, rtol=1e-10, atol=1e-12)
