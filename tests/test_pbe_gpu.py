"""PBE on the GPU: the analytic XC kernel against the oracle's autograd derivatives, the GGA Kohn-Sham
map (the sigma path: order-8 gradients, the flux and cross lumped weights, the transposed-gradient
divergence) against the oracle SCF, and the reference's PBE fingerprints.

Tolerances: kernel values and derivatives 1e-12 of their scale (e/n for de/dn, e/sigma for de/dsigma);
SCF energy terms 1e-10 Ha against the oracle and against the reference fingerprint
(``reports/fingerprint_cpu_reference.json``, its own ``--diff`` tolerance) with the iteration counts.
"""

from __future__ import annotations

import json
import math

import numpy as np
import pytest
import torch

from conftest import ROOT
from oracle import physics as ophys

pytestmark = pytest.mark.gpu
GOLDEN = json.loads((ROOT / "tests" / "golden" / "reference_values.json").read_text())
TERMS = ("total", "harris_foulkes", "kinetic", "external", "hartree", "xc")


def test_pbe_kernel_matches_oracle(pkg, device):
    kf = (3 * math.pi**2) ** (1 / 3)
    rng = np.random.default_rng(3)
    n = 10 ** rng.uniform(-13, 3, 4000)
    sig = 10 ** rng.uniform(-6, 2, 4000) * 4 * kf**2 * n ** (8 / 3)
    n[:10] = 10 ** rng.uniform(-18, -14.5, 10)  # below the threshold: zero energy and potentials
    sig[10:20] = 0.0                            # no gradient: the floored sigma
    e_o, vn_o, vs_o = ophys.pbe_xc(n, sig)
    nt = torch.as_tensor(n).to(device)
    e, vn, vs = pkg.ops.pbe(nt, torch.as_tensor(sig).to(device))
    e, vn, vs = e.cpu().numpy(), vn.cpu().numpy(), vs.cpu().numpy()
    assert np.all(e[:10] == 0) and np.all(vn[:10] == 0) and np.all(vs[:10] == 0)
    scale = np.abs(e_o) + 1e-300
    assert (np.abs(e - e_o) / scale).max() < 1e-12
    assert (np.abs(vn - vn_o) * n / scale).max() < 1e-12
    live = sig > 0
    assert (np.abs(vs - vs_o)[live] * sig[live] / scale[live]).max() < 1e-12
    assert np.allclose(vs[10:20], vs_o[10:20], rtol=1e-12, atol=0)


def test_pbe_scf_matches_oracle(pkg, device):
    """He PBE on the 16-bohr box at h = 0.5 (33^3) and H2 PBE off the lattice symmetry on a 29 x 25 x 25
    grid: every energy term of the GPU SCF against the oracle SCF."""
    from cdft_amd import scf

    cases = [(ophys.Scenario((2.0,), ((0.0, 0.0, 0.0),), 2.0, "pbe"), ophys.Numerics(box_lengths=(16.0,) * 3, spacing=0.5),
              None),
             (ophys.Scenario((1.0, 1.0), ((-0.7, 0.0, 0.0), (0.7, 0.0, 0.0)), 2.0, "pbe"),
              ophys.Numerics(box_lengths=(9.8, 8.4, 8.4), spacing=0.35), (-0.7, 0.0, 0.0))]
    for sc, num, anchor in cases:
        r = scf.solve_self_consistent(sc, num, device, anchor=anchor)
        o = ophys.solve_self_consistent(sc, num, anchor=anchor)
        assert r["scf_iterations"] == o["scf_iterations"]
        for term in TERMS:
            print(term, r["energies"][term], float(o["energies"][term]), flush=True)
            assert abs(r["energies"][term] - float(o["energies"][term])) < 1e-10, term


def test_helium_pbe_fingerprint(pkg, device):
    """``he_atom_pbe`` on the reference's derived grid: h = 0.25, box 14 (57^3)."""
    from cdft_amd import scf
    from cdft_amd.contract import Numerics, Scenario

    sc = Scenario("he_atom_pbe", (2.0,), ((0.0, 0.0, 0.0),), 2.0, "pbe")
    r = scf.solve_self_consistent(sc, Numerics(box_lengths=(14.0,) * 3, spacing=0.25), device, anchor=(0.0, 0.0, 0.0))
    assert r["step"].grid.shape == (57, 57, 57)
    for term in TERMS:
        print(term, r["energies"][term], GOLDEN[f"fingerprint.he_atom_pbe.{term}"], flush=True)
    assert r["scf_iterations"] == GOLDEN["fingerprint.he_atom_pbe.n_iterations"]
    for term in TERMS:
        assert abs(r["energies"][term] - GOLDEN[f"fingerprint.he_atom_pbe.{term}"]) < 1e-10, term
    assert abs(r["eigenvalues"][0] - GOLDEN["fingerprint.he_atom_pbe.eigenvalue"]) < 1e-10


def test_h2_pbe_fingerprint(pkg, device):
    """``h2_R1.4_pbe`` (96^3, h = 1.4 / 6, bond along z, anchored on the first nucleus)."""
    from cdft_amd import scf
    from test_configs_gpu import h2_fingerprint_case

    sc, num, anchor = h2_fingerprint_case()
    sc.xc = "pbe"
    r = scf.solve_self_consistent(sc, num, device, anchor=anchor)
    for term in TERMS:
        print(term, r["energies"][term], GOLDEN[f"fingerprint.h2_R1.4_pbe.{term}"], flush=True)
    assert r["scf_iterations"] == GOLDEN["fingerprint.h2_R1.4_pbe.n_iterations"]
    for term in TERMS:
        assert abs(r["energies"][term] - GOLDEN[f"fingerprint.h2_R1.4_pbe.{term}"]) < 1e-10, term
    assert abs(r["eigenvalues"][0] - GOLDEN["fingerprint.h2_R1.4_pbe.eigenvalue"]) < 1e-10
