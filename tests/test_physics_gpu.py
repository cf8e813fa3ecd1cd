"""The GPU geometry, quadrature, Poisson, LDA and the full solves against the oracle.

Tolerances (float64 throughout): geometry arrays 1e-13 relative; Poisson 1e-12 of the potential's
scale; LDA 1e-13; bare-H eigenvalue and He-LDA total energy 1e-10 Ha against the oracle and the
reference's published values (the SCF stops at |dE| < 1e-8, so paths that differ only by round-off
end within ~1e-12 of each other).
"""

from __future__ import annotations

import json

import numpy as np
import pytest
import torch

from conftest import ROOT
from oracle import core as ocore
from oracle import physics as ophys
from oracle import quadrature as oquad

pytestmark = pytest.mark.gpu
GOLDEN = json.loads((ROOT / "tests" / "golden" / "reference_values.json").read_text())


def _measure_distance(phi, phi_o, hamiltonian):
    """``||phi - phi_o||_omega`` up to sign, in the measure the eigenvectors are orthonormal in. At
    the box corners ``omega ~ f^2 h^3 ~ 1e-14`` and the components there are not determined by
    the eigenproblem, so a plain max-norm compares noise."""
    sign = np.sign(hamiltonian.measure.inner(phi, phi_o))
    return float(hamiltonian.measure.norm(sign * phi - phi_o))


def _rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="module")
def he_case(pkg, device):
    """He on the 16-bohr box at h = 0.5 (33^3): oracle and GPU geometry."""
    from cdft_amd import geometry as g

    grid_o = ocore.Grid.from_box((16.0,) * 3, 0.5)
    factor_o = ocore.CuspFactor(grid_o, (2.0,), [[0.0, 0.0, 0.0]])
    quad_o = oquad.CuspQuadrature(grid_o, factor_o)
    grid = g.UniformGrid.from_box((16.0,) * 3, 0.5, device)
    factor = g.CuspFactor(grid, (2.0,), [[0.0, 0.0, 0.0]])
    quad = g.CuspQuadrature(grid, factor)
    return grid_o, factor_o, quad_o, grid, factor, quad


@pytest.fixture(scope="module")
def h2_case(pkg, device):
    """Two centres off the lattice (H2 at R = 1.4 along x on a 0.4 grid): no symmetry to hide a
    transposed index, overlapping spheres."""
    from cdft_amd import geometry as g

    pos = [[-0.7, 0.05, 0.0], [0.7, 0.05, 0.0]]
    grid_o = ocore.Grid.from_box((14.0, 12.0, 12.8), 0.4)
    factor_o = ocore.CuspFactor(grid_o, (1.0, 1.0), pos)
    grid = g.UniformGrid.from_box((14.0, 12.0, 12.8), 0.4, device)
    factor = g.CuspFactor(grid, (1.0, 1.0), pos)
    return grid_o, factor_o, grid, factor


def test_grid_fields(he_case, h2_case):
    for case in (he_case[:2] + he_case[3:5], h2_case):
        grid_o, factor_o, grid, factor = case
        assert grid.shape == grid_o.shape and grid.origin == grid_o.origin and grid.spacing == grid_o.spacing
        assert _rel(factor.u.cpu(), factor_o.u) < 1e-14
        assert _rel(factor.weight.cpu(), factor_o.weight) < 1e-13
        assert _rel(factor.grad_u.cpu(), factor_o.grad_u) < 1e-13
        assert _rel(factor.transformed_potential.cpu(), factor_o.transformed_potential) < 1e-13
        for a, b in zip(factor.face_weights(8), factor_o.face_weights(8)):
            assert _rel(a.cpu(), b) < 1e-13


def test_quadrature_arrays(he_case):
    _, _, quad_o, grid, _, quad = he_case
    assert quad.n_nodes == quad_o.n_nodes
    perm = quad.node_order  # GPU nodes are numbered in base-cell order
    assert sorted(perm.tolist()) == list(range(quad.n_nodes))
    # the transpose's binning is built from host base cells; they are the kernel's exactly
    assert np.array_equal(quad.base.cpu().numpy(), quad._base_host)
    assert _rel(quad.points.cpu(), quad_o.points[perm]) < 1e-15
    assert _rel(quad.weights.cpu(), quad_o.weights[perm]) < 1e-13
    assert _rel(quad.f2.cpu(), quad_o.f2[perm]) < 1e-13
    assert _rel(quad.far.cpu(), quad_o.far) < 1e-13
    assert _rel(quad.mass_weights.cpu(), quad_o.mass_weights) < 1e-13
    pot, invw, q = quad.bare_operator_terms()
    pot_o, invw_o, meas_o = quad_o.bare_operator_terms()
    assert _rel(pot.cpu(), pot_o) < 1e-13
    assert _rel(invw.cpu(), invw_o) < 1e-13
    assert _rel(q.cpu(), meas_o.quadrature) < 1e-13


def test_interpolate_and_lift_are_transposes(he_case, device):
    grid_o, _, quad_o, grid, _, quad = he_case
    rng = np.random.default_rng(5)
    field = rng.standard_normal((3, grid.n_points))
    perm = quad.node_order
    got = quad.interpolate(torch.as_tensor(field).to(device)).cpu().numpy()
    assert _rel(got, quad_o.interpolate(field)[:, perm]) < 1e-13
    x = rng.standard_normal((2, quad.n_nodes))
    lifted = quad.lift(torch.as_tensor(x).to(device)).cpu().numpy()
    x_oracle = np.empty_like(x)
    x_oracle[:, perm] = x
    assert _rel(lifted, quad_o.lift(x_oracle)) < 1e-13
    # <L f, x> = <f, L^T x>
    lhs = (got[0] * x[0]).sum()
    rhs = (field[0] * lifted[0]).sum()
    assert abs(lhs - rhs) < 1e-11 * abs(lhs)


@pytest.mark.parametrize("nf", [1, 2, 3, 4])
def test_chunked_interpolation_is_the_per_node_one_bitwise(he_case, h2_case, device, nf):
    """One load of each base cell's 8^3 block per chunk of its nodes gives bitwise the per-node
    interpolant, on the single-centre rule (a dense cluster of nodes at the nucleus) and on two
    overlapping off-lattice spheres; every chunk holds nodes of one base cell."""
    from cdft_amd import geometry as g

    h2_grid, h2_factor = h2_case[2], h2_case[3]
    for quad in (he_case[5], g.CuspQuadrature(h2_grid, h2_factor)):
        bounds = quad._interp_node.cpu().numpy()
        assert bounds[0] == 0 and bounds[-1] == quad.n_nodes and np.all(np.diff(bounds) > 0)
        assert np.diff(bounds).max() <= quad.interp_chunk and quad._interp_chunks == bounds.size - 1
        base = quad._base_host
        for a, b in zip(bounds[:-1], bounds[1:]):
            assert (base[a:b] == base[a]).all()
        rng = np.random.default_rng(nf)
        field = torch.as_tensor(rng.standard_normal((nf, quad.grid.n_points))).to(device)
        chunked = quad.interpolate(field)
        assert torch.equal(chunked, quad.interpolate(field, per_node=True))


def test_poisson_matches_oracle_and_gaussian(he_case, device):
    from cdft_amd.poisson import CoulombCutoffPoisson

    grid_o, _, _, grid, _, _ = he_case
    pts = grid_o.points()
    sigma = 1.0
    rho = np.exp(-(pts**2).sum(-1) / (2 * sigma**2)) / (2 * np.pi * sigma**2) ** 1.5
    solver = CoulombCutoffPoisson(grid, 2.5)
    oracle = ophys.CoulombCutoffPoisson(grid_o, 2.5)
    assert solver.padded_shape == oracle.padded_shape
    v = solver.solve(torch.as_tensor(rho).to(device)).cpu().numpy()
    vo = oracle.solve(rho)
    assert _rel(v, vo) < 1e-12
    energy = 0.5 * grid_o.integrate(rho * v)
    assert abs(energy - 1.0 / (2 * np.sqrt(np.pi) * sigma)) < 1e-6  # Q^2/(2 sqrt(pi) sigma), grid-limited
    # further solves through the same buffers: an off-centre, asymmetric density (the pruned (y, z)
    # transforms and the untouched zero slabs), then the first density again, bitwise as before
    rng = np.random.default_rng(3)
    rho2 = np.exp(-((pts - [1.3, -0.4, 2.1]) ** 2).sum(-1) / 1.7) * (1.0 + 0.2 * rng.uniform(-1, 1, len(pts)))
    v2 = solver.solve(torch.as_tensor(rho2).to(device)).cpu().numpy()
    assert _rel(v2, oracle.solve(rho2)) < 1e-12
    v3 = solver.solve(torch.as_tensor(rho).to(device)).cpu().numpy()
    assert np.array_equal(v3, v)


@pytest.mark.parametrize("shape", [(13, 9, 29), (29, 10, 7), (10, 13, 9), (65, 12, 11), (101, 8, 9), (81, 7, 10)])
def test_poisson_column_kernel_is_the_rocfft_path(pkg, device, shape):
    """The x transforms and the multiply as one column kernel (radix 4, 2, 3, 5 Stockham stages in
    LDS) against rocFFT's strided transforms around a separate multiply, on padded lengths 36, 75, 25,
    180, 256 and 216 along x (the He 65^3, H2 101 and He 81^3 boxes' x lengths): the same potential to
    1e-14 relative, and against the oracle's numpy FFT to 1e-12; the column kernel repeats bitwise."""
    from cdft_amd import poisson
    from cdft_amd.geometry import UniformGrid

    grid = UniformGrid.from_box(tuple((n - 1) * 0.5 for n in shape), 0.5, device)
    assert grid.shape == shape
    rng = np.random.default_rng(sum(shape))
    rho = rng.uniform(0.0, 1.0, grid.n_points) * np.exp(-rng.uniform(0, 0.1, grid.n_points))
    solver = poisson.CoulombCutoffPoisson(grid, 2.5)
    x = torch.as_tensor(rho).to(device)
    try:
        poisson.set_x_path(1)
        library = solver.solve(x).cpu().numpy()
    finally:
        poisson.set_x_path(0)
    column = solver.solve(x).cpu().numpy()
    assert _rel(column, library) < 1e-14
    assert np.array_equal(solver.solve(x).cpu().numpy(), column)
    grid_o = ocore.Grid.from_box(tuple((n - 1) * 0.5 for n in shape), 0.5)
    assert _rel(column, ophys.CoulombCutoffPoisson(grid_o, 2.5).solve(rho)) < 1e-12
    print(f"[poisson] {shape} padded {solver.padded_shape}: column vs rocFFT {_rel(column, library):.2e}", flush=True)


def test_lda_matches_oracle(pkg, device):
    n = np.concatenate([np.logspace(-16, 2, 4000), [0.0, 5e-15, 1e-14, 2e-14]])
    e, v = pkg.ops.lda_vwn(torch.as_tensor(n).to(device))
    eo, vo = ophys.lda_vwn(n)
    assert np.abs(e.cpu().numpy() - eo).max() <= 1e-13 * np.abs(eo).max()
    vg = v.cpu().numpy()
    # Below n ~ 1e-10 the VWN fit cancels (para ~ 1/x against O(1) terms), so libm differences of
    # one ulp are amplified; those points carry no weight. Relative 1e-13 above, absolute below.
    dense = n > 1e-10
    assert np.all(np.abs(vg - vo)[dense] <= 1e-13 * np.abs(vo[dense]))
    assert np.abs(vg - vo).max() <= 1e-12 * np.abs(vo).max()
    assert np.all(vg[n < 1e-14] == 0.0)


def test_bare_hydrogen_solve_matches_golden(pkg, device):
    from cdft_amd import scf

    r = scf.solve_noninteracting(ophys.Scenario((1.0,), ((0.0, 0.0, 0.0),), 1.0), ophys.Numerics(), device)
    o = ophys.solve_noninteracting(ophys.Scenario((1.0,), ((0.0, 0.0, 0.0),), 1.0), ophys.Numerics())
    assert r["n_iterations"] == o["n_iterations"] == GOLDEN["fingerprint.h_atom.n_iterations"]
    assert abs(r["eigenvalues"][0] - o["eigenvalues"][0]) < 1e-12
    assert abs(r["eigenvalues"][0] - GOLDEN["eigenvalue.ground.h_atom"]) < 1e-12
    assert abs(r["kinetic"] - o["kinetic"]) < 1e-10
    # A residual-converged eigenvector is fixed only to ~ residual_tol / gap (1e-8 / 0.4), so two
    # round-off paths agree to that; the tight-tolerance test below pins eigenvectors to 1e-10.
    assert _measure_distance(r["eigenvectors"].cpu().numpy()[0], o["eigenvectors"][0], o["hamiltonian"]) < 2e-8


def test_bare_hydrogen_eigenvector_tight(pkg, device):
    """Converged to residual 1e-12 on both sides, the eigenvectors agree to 1e-10."""
    from cdft_amd import scf
    from oracle.eigen import EigenConfig

    num = ophys.Numerics(eigen=EigenConfig(residual_tol=1e-12, cross_check_tol=1e-14))
    sc = ophys.Scenario((1.0,), ((0.0, 0.0, 0.0),), 1.0)
    r = scf.solve_noninteracting(sc, num, device)
    o = ophys.solve_noninteracting(sc, num)
    assert _measure_distance(r["eigenvectors"].cpu().numpy()[0], o["eigenvectors"][0], o["hamiltonian"]) < 1e-10
    assert abs(r["eigenvalues"][0] - o["eigenvalues"][0]) < 1e-13


def test_helium_lda_scf_h05_matches_ladder(pkg, device):
    from cdft_amd import scf

    num = ophys.Numerics(box_lengths=(16.0, 16.0, 16.0), spacing=0.5)
    r = scf.solve_self_consistent(ophys.Scenario((2.0,), ((0.0, 0.0, 0.0),), 2.0, "lda"), num, device)
    assert r["scf_iterations"] == GOLDEN["ladder.he_lda.h0.5.scf_iterations"]
    assert abs(r["energies"]["total"] - GOLDEN["ladder.he_lda.h0.5.total"]) < 1e-10


def test_lumped_far_weights_match_oracle(h2_case):
    """Overlapping spheres: Becke/lune partition at the nodes, the lumped far mass weights on the
    twice-refined lattice with the edge blend, and the mass weights built from them."""
    from cdft_amd import geometry as g

    grid_o, factor_o, grid, factor = h2_case
    quad_o = oquad.CuspQuadrature(grid_o, factor_o)
    quad = g.CuspQuadrature(grid, factor)
    assert quad.far_lumped and quad_o.far_lumped
    assert quad.n_nodes == quad_o.n_nodes
    assert np.array_equal(quad.base.cpu().numpy(), quad._base_host)
    perm = quad.node_order
    assert _rel(quad.weights.cpu(), quad_o.weights[perm]) < 1e-13
    assert _rel(quad.lumped_far.cpu(), quad_o.lumped_far) < 1e-13
    assert _rel(quad.mass_weights.cpu(), quad_o.mass_weights) < 1e-13


def test_h2plus_solve_matches_ladder(pkg, device):
    """Bare H2+ at R = 2 on the reference's derived grid for h = 0.5 (box 14, anchored on the first
    nucleus, 29^3): the ladder eigenvalue of the reference."""
    from cdft_amd import scf

    sc = ophys.Scenario((1.0, 1.0), ((0.0, 0.0, -1.0), (0.0, 0.0, 1.0)), 1.0)
    num = ophys.Numerics(box_lengths=(14.0, 14.0, 14.0), spacing=0.5)
    r = scf.solve_noninteracting(sc, num, device, anchor=(0.0, 0.0, -1.0))
    assert abs(r["eigenvalues"][0] - GOLDEN["ladder.h2plus_R2.h0.5.eigenvalue"]) < 1e-11


def test_h2_lda_scf_matches_oracle(pkg, device):
    """H2 at VWN LDA, R = 1.4 along x, lattice anchored on the first proton (both protons on the
    lattice), 29 x 25 x 25 at h = 0.35: the GPU SCF against the oracle SCF."""
    from cdft_amd import scf

    sc = ophys.Scenario((1.0, 1.0), ((-0.7, 0.0, 0.0), (0.7, 0.0, 0.0)), 2.0, "lda")
    num = ophys.Numerics(box_lengths=(9.8, 8.4, 8.4), spacing=0.35)
    anchor = (-0.7, 0.0, 0.0)
    r = scf.solve_self_consistent(sc, num, device, anchor=anchor)
    o = ophys.solve_self_consistent(sc, num, anchor=anchor)
    assert r["scf_iterations"] == o["scf_iterations"]
    for term in ("total", "harris_foulkes", "kinetic", "external", "hartree", "xc"):
        print(term, r["energies"][term], float(o["energies"][term]), flush=True)
    for term in ("total", "harris_foulkes", "kinetic", "external", "hartree", "xc"):
        assert abs(r["energies"][term] - float(o["energies"][term])) < 1e-10, term


def test_captured_potentials_and_energies_match_eager(pkg, device):
    """The potentials and energy evaluations replayed from their captured graphs give bitwise the
    eager results, on every replay, with eager work between replays and fresh inputs each time."""
    from cdft_amd import scf

    sc = ophys.Scenario((1.0, 1.0), ((-0.7, 0.05, 0.0), (0.7, 0.05, 0.0)), 2.0, "lda")
    num = ophys.Numerics(box_lengths=(9.8, 8.4, 8.4), spacing=0.35)
    grid, factor, quad = scf.build_geometry(sc, num, device)
    captured = scf.KohnShamStep(sc, num, grid, factor, quad)
    eager = scf.KohnShamStep(sc, num, grid, factor, quad)
    eager.use_graphs = False
    n0 = scf.initial_density(captured)
    gen = torch.Generator(device="cpu").manual_seed(5)
    k = 3
    for call in range(5):
        bump = 1.0 + 0.05 * torch.rand(n0.shape, generator=gen, dtype=torch.float64).to(device)
        density = n0 * bump
        phi = torch.randn((k, grid.n_points), generator=gen, dtype=torch.float64).to(device)
        a = captured.potentials(density)
        b = eager.potentials(density)
        torch.cuda.synchronize()
        assert torch.equal(a["effective"], b["effective"]), call
        assert torch.equal(a["scalars"], b["scalars"]), call
        ea = captured.energy_scalars(phi, None, density[0], a["effective"])
        eb = eager.energy_scalars(phi, None, density[0], b["effective"])
        torch.cuda.synchronize()
        assert torch.equal(ea, eb), call
    assert captured.graph_record["disabled"] is None, captured.graph_record["disabled"]
    assert captured.graph_record["captures"] == 2
    assert captured.graph_record["replays"] == 2 * 4
    assert eager.graph_record["replays"] == 0


def test_speculative_filters_change_nothing(pkg, device, monkeypatch):
    """The CheFSI solve enqueues the next pass's filter before reading a pass when that pass cannot
    stop the solve; with the speculation off, an SCF gives bitwise the same energies, eigenvalues and
    iteration counts (He LDA at h = 0.5, bound reuse and float32 warm passes included)."""
    from cdft_amd import eigen, scf

    sc = ophys.Scenario((2.0,), ((0.0, 0.0, 0.0),), 2.0, "lda")
    num = ophys.Numerics(box_lengths=(16.0, 16.0, 16.0), spacing=0.5)
    runs = {}
    for label, flag, dtype in (("on", True, torch.float64), ("off", False, torch.float64),
                               ("on32", True, torch.float32), ("off32", False, torch.float32)):
        monkeypatch.setattr(eigen.ChebyshevFilteredSubspace, "speculate", flag)
        runs[label] = scf.solve_self_consistent(sc, num, device, filter_dtype=dtype)
    for a, b in (("on", "off"), ("on32", "off32")):
        assert runs[a]["energy_history"] == runs[b]["energy_history"]
        assert np.array_equal(runs[a]["eigenvalues"], runs[b]["eigenvalues"])
        assert runs[a]["applications"] == runs[b]["applications"]
    assert runs["on"]["speculative_filters"]["used"] > 0
    assert runs["off"]["speculative_filters"]["used"] == 0


def test_reused_solver_matches_a_fresh_one(pkg, device, monkeypatch):
    """An SCF solve that reuses the cached subspace solver of an earlier solve with other atoms on the
    same grid shape (its buffers and captured graphs, with the new geometry copied into its static
    arrays) gives bitwise the energies, eigenvalues and application counts of a solve with a fresh
    solver; float64 and float32 filters."""
    from cdft_amd import scf

    first = ophys.Scenario((1.0, 1.0), ((-0.7, 0.0, 0.0), (0.7, 0.0, 0.0)), 2.0, "lda")
    second = ophys.Scenario((1.0, 1.0), ((-0.75, 0.05, 0.0), (0.75, 0.05, 0.0)), 2.0, "lda")
    num = ophys.Numerics(box_lengths=(9.8, 8.4, 8.4), spacing=0.35)
    for dtype in (torch.float64, torch.float32):
        monkeypatch.setattr(scf, "_solvers", {})
        monkeypatch.setattr(scf, "SOLVER_CACHE_SIZE", 2)
        scf.solve_self_consistent(first, num, device, filter_dtype=dtype)
        reused = scf.solve_self_consistent(second, num, device, filter_dtype=dtype)
        monkeypatch.setattr(scf, "_solvers", {})
        monkeypatch.setattr(scf, "SOLVER_CACHE_SIZE", 0)
        fresh = scf.solve_self_consistent(second, num, device, filter_dtype=dtype)
        assert reused["step"].grid.shape == fresh["step"].grid.shape
        assert reused["energy_history"] == fresh["energy_history"], dtype
        assert np.array_equal(reused["eigenvalues"], fresh["eigenvalues"]), dtype
        assert reused["applications"] == fresh["applications"], dtype
        assert reused["filter_precisions"] == fresh["filter_precisions"], dtype
        assert reused["graphs"]["disabled"] is None and fresh["graphs"]["disabled"] is None
        assert reused["graphs"]["captures"] < fresh["graphs"]["captures"], (reused["graphs"], fresh["graphs"])
