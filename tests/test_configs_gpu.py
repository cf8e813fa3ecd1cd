"""The benchmark configurations at their full sizes, against the reference's published energies.

- He LDA on the 16-bohr box at h = 0.25 (65^3) and h = 0.20 (81^3): the reference's convergence
  ladder (``docs/figures/convergence_ladders.json:35-52``), total energy to 1e-12 Ha and the exact
  SCF iteration counts.
- The reference's CPU fingerprint (``reports/fingerprint_cpu_reference.json``) on its derived
  grids: ``he_atom_lda`` (57^3, box 14) and ``h2_R1.4_lda`` (96^3, h = 1.4 / 6, bond along z, the
  lattice anchored on the first nucleus). Every energy term and the eigenvalue to 1e-10 Ha (the
  fingerprint's own ``--diff`` tolerance), and the iteration counts. The H2 case runs the two-centre
  Hartree split (``c = S^-1 n(R_a)``) that no single-centre case reaches.
- H2 LDA at 101 x 81 x 81 (protons at (+-0.7, 0, 0), h = 0.20, anchored on the first proton): the
  mixed solve (float32 filters, float64 Rayleigh-Ritz, Cholesky, density and energies, float64
  refinement passes) within 1e-8 Ha of the float64 solve.
- The 64-geometry H2+ scan at h = 0.25: the batched lockstep solve gives bitwise the eigenvalues and
  iteration counts of 64 single-grid solves, every grid converges, and the ground-state energy
  ``eps(R) + 1/R`` has its minimum inside the scanned range. A batch holding R = 2 on the reference's
  derived h = 0.5 grid (box 14, anchored on the first nucleus) reproduces the reference's ladder
  eigenvalue from inside the batch.
"""

from __future__ import annotations

import json

import numpy as np
import pytest
import torch

from conftest import ROOT

pytestmark = pytest.mark.gpu
GOLDEN = json.loads((ROOT / "tests" / "golden" / "reference_values.json").read_text())
TERMS = ("total", "harris_foulkes", "kinetic", "external", "hartree", "xc")


def _he_lda():
    from cdft_amd.contract import helium_lda

    return helium_lda()


@pytest.mark.parametrize("spacing, key", [(0.25, "h0.25"), (0.2, "h0.2")])
def test_helium_lda_box16_matches_reference_ladder(pkg, device, spacing, key):
    from cdft_amd import scf
    from cdft_amd.contract import Numerics

    print(f"He LDA box 16 h={spacing}", flush=True)
    r = scf.solve_self_consistent(_he_lda(), Numerics(box_lengths=(16.0,) * 3, spacing=spacing), device)
    assert r["step"].grid.shape == ((65,) * 3 if spacing == 0.25 else (81,) * 3)
    assert r["converged"]
    assert r["scf_iterations"] == GOLDEN[f"ladder.he_lda.{key}.scf_iterations"]
    assert abs(r["energies"]["total"] - GOLDEN[f"ladder.he_lda.{key}.total"]) < 1e-12


def test_helium_lda_fingerprint_terms(pkg, device):
    """``he_atom_lda`` on the reference's derived grid: h = 0.25, box 14 (57^3, origin -7)."""
    from cdft_amd import scf
    from cdft_amd.contract import Numerics

    r = scf.solve_self_consistent(_he_lda(), Numerics(box_lengths=(14.0,) * 3, spacing=0.25), device,
                                  anchor=(0.0, 0.0, 0.0))
    grid = r["step"].grid
    assert grid.shape == (57, 57, 57) and grid.origin == (-7.0, -7.0, -7.0)
    assert r["scf_iterations"] == GOLDEN["fingerprint.he_atom_lda.n_iterations"]
    for term in TERMS:
        assert abs(r["energies"][term] - GOLDEN[f"fingerprint.he_atom_lda.{term}"]) < 1e-10, term
    assert abs(r["eigenvalues"][0] - GOLDEN["fingerprint.he_atom_lda.eigenvalue"]) < 1e-10


def h2_fingerprint_case():
    """``h2_R1.4_lda``'s grid as the reference derives it (``grid.py:190-236``): h = 1.4 / round(1.4 /
    0.25) = 1.4 / 6, the edge 1.4 + 2 ln(1e9) / 2 rounded up to 95 cells, anchored on the first
    nucleus (0, 0, -0.7)."""
    from types import SimpleNamespace

    from cdft_amd.contract import Numerics

    h = 1.4 / 6
    edge = 95 * h
    sc = SimpleNamespace(charges=(1.0, 1.0), positions=((0.0, 0.0, -0.7), (0.0, 0.0, 0.7)), n_electrons=2.0,
                         xc="lda")
    return sc, Numerics(box_lengths=(edge,) * 3, spacing=h), (0.0, 0.0, -0.7)


def test_h2_lda_fingerprint_terms(pkg, device):
    """``h2_R1.4_lda`` (96^3): the two-centre Kato factor and Hartree cusp split."""
    from cdft_amd import scf

    sc, num, anchor = h2_fingerprint_case()
    r = scf.solve_self_consistent(sc, num, device, anchor=anchor)
    grid = r["step"].grid
    assert grid.shape == (96, 96, 96)
    assert np.allclose(grid.origin, (-11.2, -11.2, -10.966666666666665), atol=1e-12, rtol=0)
    assert r["scf_iterations"] == GOLDEN["fingerprint.h2_R1.4_lda.n_iterations"]
    for term in TERMS:
        print(term, r["energies"][term], GOLDEN[f"fingerprint.h2_R1.4_lda.{term}"], flush=True)
    for term in TERMS:
        assert abs(r["energies"][term] - GOLDEN[f"fingerprint.h2_R1.4_lda.{term}"]) < 1e-10, term
    assert abs(r["eigenvalues"][0] - GOLDEN["fingerprint.h2_R1.4_lda.eigenvalue"]) < 1e-10


def test_h2_lda_101x81x81_mixed_within_1e8_of_float64(pkg, device):
    from cdft_amd import scf
    from cdft_amd.contract import Numerics, hydrogen_molecule_lda

    sc = hydrogen_molecule_lda(1.4, axis=0)
    num = Numerics(box_lengths=(20.0, 16.0, 16.0), spacing=0.2)
    anchor = (-0.7, 0.0, 0.0)
    f64 = scf.solve_self_consistent(sc, num, device, anchor=anchor)
    mixed = scf.solve_self_consistent(sc, num, device, anchor=anchor, filter_dtype=torch.float32)
    assert f64["step"].grid.shape == (101, 81, 81)
    assert f64["converged"] and mixed["converged"]
    assert "f32" in mixed["filter_precisions"]
    assert mixed["filter_precisions"][-2:] == ["f64", "f64"]  # the float64 refinement passes end the solve
    assert abs(mixed["energies"]["total"] - f64["energies"]["total"]) <= 1e-8
    # The eigenvalue is first order in the density, which the SCF converges to its density tolerance
    # (1e-6 in the L1 norm) on either path: the two paths agree to a few 1e-8 Ha, not to round-off.
    assert abs(mixed["eigenvalues"][0] - f64["eigenvalues"][0]) <= 1e-7


def test_batched_scan_64_geometries_matches_sequential(pkg, device):
    from cdft_amd.batch import BatchedCheFSI, bond_scan_geometries
    from cdft_amd.contract import Numerics
    from cdft_amd.hamiltonian import CuspFactoredHamiltonian
    from cdft_amd.scf import build_geometry, solve_noninteracting

    scan = bond_scan_geometries(count=64, seed=0, spacing=0.25)
    hams, solo = [], []
    for r, sc, box in scan:
        num = Numerics(box_lengths=box, spacing=0.25)
        geometry = build_geometry(sc, num, device)
        hams.append(CuspFactoredHamiltonian(*geometry))
        solo.append(solve_noninteracting(sc, num, device, geometry=geometry))
    shapes = {h.grid.shape for h in hams}
    assert len(shapes) > 10 and all(s[1:] == (57, 57) for s in shapes)
    batched = BatchedCheFSI(hams, Numerics().eigen, n_states=1).solve()
    for b, s in zip(batched, solo):
        assert b["converged"] and s["converged"]
        assert b["n_iterations"] == s["n_iterations"]
        assert b["eigenvalues"][0] == s["eigenvalues"][0]
    bonds = np.array([r for r, _, _ in scan])
    order = np.argsort(bonds)
    evs = np.array([b["eigenvalues"][0] for b in batched])[order]
    assert np.all(np.diff(evs) > 0)  # the electronic energy rises monotonically towards separation
    total = evs + 1.0 / bonds[order]
    best = int(np.argmin(total))
    assert 0 < best < len(total) - 1 and 1.5 < bonds[order][best] < 2.6  # the H2+ minimum near R = 2


def test_batched_scan_reproduces_reference_ladder_rung(pkg, device):
    """R = 2 on the reference's derived grid for h = 0.5 (box 14, bond along z, anchored on the first
    nucleus, 29^3) solved inside a batch with two other bond lengths on their own grids."""
    from types import SimpleNamespace

    from cdft_amd.batch import BatchedCheFSI, bond_scan_geometries
    from cdft_amd.contract import Numerics
    from cdft_amd.hamiltonian import CuspFactoredHamiltonian
    from cdft_amd.scf import build_geometry

    rung = SimpleNamespace(charges=(1.0, 1.0), positions=((0.0, 0.0, -1.0), (0.0, 0.0, 1.0)), n_electrons=1.0,
                           xc="none")
    hams = [CuspFactoredHamiltonian(*build_geometry(rung, Numerics(box_lengths=(14.0,) * 3, spacing=0.5), device,
                                                    anchor=(0.0, 0.0, -1.0)))]
    for r, sc, box in bond_scan_geometries(count=2, seed=3, spacing=0.5):
        hams.append(CuspFactoredHamiltonian(*build_geometry(sc, Numerics(box_lengths=box, spacing=0.5), device)))
    assert hams[0].grid.shape == (29, 29, 29)
    out = BatchedCheFSI(hams, Numerics().eigen, n_states=1).solve()
    assert all(o["converged"] for o in out)
    assert abs(out[0]["eigenvalues"][0] - GOLDEN["ladder.h2plus_R2.h0.5.eigenvalue"]) < 1e-11
