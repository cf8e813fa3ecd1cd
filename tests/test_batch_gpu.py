"""The batched scan: one launch per filter step covers several grids of different sizes, and the
lockstep CheFSI reproduces each grid's own solve.

Tolerances: the batched descriptor launch runs the single-grid per-block code, so batched and
single-grid eigenvalues, iteration counts and eigenvectors agree bitwise; the test asserts 1e-14 on
eigenvalues (relative) and equality of the iteration counts.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_batched_apply_and_filter_match_single_grid(pkg, device):
    from cdft_amd.batch import BatchPlan, bond_scan_geometries
    from cdft_amd.chefsi import chebyshev_coefficients
    from cdft_amd.contract import Numerics
    from cdft_amd.hamiltonian import CuspFactoredHamiltonian
    from cdft_amd.scf import build_geometry

    geoms = []
    for r, sc, box in bond_scan_geometries(count=3, seed=4, spacing=0.5):
        grid, factor, q = build_geometry(sc, Numerics(box_lengths=box, spacing=0.5), device)
        geoms.append(CuspFactoredHamiltonian(grid, factor, q))
    assert len({h.grid.shape for h in geoms}) > 1  # uneven grids
    k, degree = 5, 20
    gen = torch.Generator().manual_seed(0)
    rows = []
    for h in geoms:
        n = h.grid.n_points
        x = (torch.rand((k, n), generator=gen, dtype=torch.float64) * 2 - 1).to(device)
        table = torch.tensor(chebyshev_coefficients(degree, -0.3, 30.0, -1.2), dtype=torch.float64,
                             device=device)
        rows.append((h.geometry, x, torch.empty_like(x), torch.empty_like(x), torch.empty_like(x), table))
    plan = BatchPlan(rows, k)
    plan.apply(rows[0][1])
    for geom, x, out, _, _, _ in rows:
        assert torch.equal(out, pkg.ops.hdiv_apply(x, geom))
    plan.filter(degree, rows[0][1])
    for geom, x, out, _, _, table in rows:
        assert torch.equal(out, pkg.ops.cheb_filter(x, geom, table, degree))


def test_batched_chefsi_matches_single_solves(pkg, device):
    from cdft_amd.batch import BatchedCheFSI, bond_scan_geometries
    from cdft_amd.contract import Numerics
    from cdft_amd.eigen import EigenConfig
    from cdft_amd.hamiltonian import CuspFactoredHamiltonian
    from cdft_amd.scf import build_geometry, solve_noninteracting

    scan = bond_scan_geometries(count=4, seed=0, spacing=0.5)
    hams, solo = [], []
    for r, sc, box in scan:
        num = Numerics(box_lengths=box, spacing=0.5)
        geometry = build_geometry(sc, num, device)
        hams.append(CuspFactoredHamiltonian(*geometry))
        solo.append(solve_noninteracting(sc, num, device, geometry=geometry))
    batched = BatchedCheFSI(hams, EigenConfig(), n_states=1).solve()
    for b, s in zip(batched, solo):
        assert b["n_iterations"] == s["n_iterations"]
        assert abs(b["eigenvalues"][0] - s["eigenvalues"][0]) <= 1e-14 * abs(s["eigenvalues"][0])
    # binding energies decrease towards the united atom: eps(R) is increasing in R here
    order = np.argsort([r for r, _, _ in scan])
    evs = np.array([b["eigenvalues"][0] for b in batched])[order]
    assert np.all(np.diff(evs) > 0)


def test_float32_batched_filter_matches_single_grid(pkg, device):
    """The float32 batch runs the float32 single-grid filter's steps, bitwise."""
    from cdft_amd.batch import BatchPlan, bond_scan_geometries
    from cdft_amd.chefsi import chebyshev_coefficients
    from cdft_amd.contract import Numerics
    from cdft_amd.hamiltonian import CuspFactoredHamiltonian
    from cdft_amd.scf import build_geometry

    rows = []
    gen = torch.Generator().manual_seed(1)
    k, degree = 5, 20
    for r, sc, box in bond_scan_geometries(count=3, seed=4, spacing=0.5):
        grid, factor, q = build_geometry(sc, Numerics(box_lengths=box, spacing=0.5), device)
        geom = CuspFactoredHamiltonian(grid, factor, q).geometry_in(torch.float32)
        x = (torch.rand((k, grid.n_points), generator=gen, dtype=torch.float32) * 2 - 1).to(device)
        table = torch.tensor(chebyshev_coefficients(degree, -0.3, 30.0, -1.2), dtype=torch.float64, device=device)
        rows.append((geom, x, torch.empty_like(x), torch.empty_like(x), torch.empty_like(x), table))
    plan = BatchPlan(rows, k)
    plan.filter(degree, rows[0][1])
    for geom, x, out, _, _, table in rows:
        assert torch.equal(out, pkg.ops.cheb_filter(x, geom, table, degree))
    plan.apply(rows[0][1])
    for geom, x, out, _, _, _ in rows:
        assert torch.equal(out, pkg.ops.hdiv_apply(x, geom))


def test_mixed_batched_chefsi_matches_float64_batch(pkg, device):
    """Mixed precision in the lockstep batch: float32 filters while every active grid's residual is
    above 1e-5, float64 after; the eigenvalues agree with the float64 batch to 1e-12 relative."""
    from cdft_amd.batch import BatchedCheFSI, bond_scan_geometries
    from cdft_amd.contract import Numerics
    from cdft_amd.eigen import EigenConfig
    from cdft_amd.hamiltonian import CuspFactoredHamiltonian
    from cdft_amd.scf import build_geometry

    hams = []
    for r, sc, box in bond_scan_geometries(count=4, seed=0, spacing=0.5):
        hams.append(CuspFactoredHamiltonian(*build_geometry(sc, Numerics(box_lengths=box, spacing=0.5), device)))
    f64 = BatchedCheFSI(hams, EigenConfig(), n_states=1).solve()
    solver = BatchedCheFSI(hams, EigenConfig(), n_states=1, filter_dtype=torch.float32)
    mixed = solver.solve()
    assert "f32" in solver.filter_precisions and solver.filter_precisions[-1] == "f64"
    for a, b in zip(mixed, f64):
        assert a["converged"]
        assert abs(a["eigenvalues"][0] - b["eigenvalues"][0]) <= 1e-12 * abs(b["eigenvalues"][0])


def test_autotuned_launch_leaves_the_eigenvalues_bitwise_unchanged(pkg, device):
    """The batch autotuner (segment length, then tile extents) changes only the launch: the tuned
    batched solve gives bitwise the eigenvalues and iteration counts of the untuned one."""
    from cdft_amd import autotune
    from cdft_amd.batch import BatchedCheFSI, bond_scan_geometries
    from cdft_amd.contract import Numerics
    from cdft_amd.eigen import EigenConfig
    from cdft_amd.hamiltonian import CuspFactoredHamiltonian
    from cdft_amd.scf import build_geometry

    scan = bond_scan_geometries(count=3, seed=5, spacing=0.5)
    hams = [CuspFactoredHamiltonian(*build_geometry(sc, Numerics(box_lengths=box, spacing=0.5), device))
            for _, sc, box in scan]
    tuned_solver = BatchedCheFSI(hams, EigenConfig(), n_states=1, autotune=True)
    tuned = tuned_solver.solve()
    plain = BatchedCheFSI(hams, EigenConfig(), n_states=1, autotune=False).solve()
    assert "tile" in tuned_solver.tuning and len(tuned_solver.tuning["tile_us"]) == len(autotune.TILES)
    for a, b in zip(tuned, plain):
        assert a["n_iterations"] == b["n_iterations"]
        assert np.array_equal(a["eigenvalues"], b["eigenvalues"])
