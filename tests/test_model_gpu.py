"""The model systems' plain operator on the device against the oracle, and the two model systems of
the reference's registry against the reference's CPU fingerprint.

Tolerances: apply, filter and residual norms against the oracle to 1e-12 relative (float64, sums in
another order); the solves against ``reports/fingerprint_cpu_reference.json`` (the reference's values,
as data): eigenvalue, kinetic and total energy to 1e-10 Ha (the fingerprint's own tolerance), iteration
counts equal (15 for the harmonic well, 36 for the infinite well).
"""

from __future__ import annotations

import json
import pathlib

import numpy as np
import pytest
import torch

from oracle.core import Grid
from oracle.model import PlainHamiltonian, harmonic_potential

pytestmark = pytest.mark.gpu

FINGERPRINT = json.loads((pathlib.Path(__file__).parent / "golden" / "fingerprint_cpu_reference.json").read_text())


def _geometry(pkg, shape, h, order, boundary, pot):
    from cdft_amd import ops

    return ops.LocalGeometry(shape, order, boundary, ops.local_coefficients(order, h), pot)


@pytest.mark.parametrize("boundary", [0, 1])
@pytest.mark.parametrize("order", [2, 8, 12])
def test_apply_filter_and_residuals_match_the_oracle(pkg, device, boundary, order):
    from cdft_amd import ops

    shape, h = (23, 19, 26), 0.3
    grid = Grid(shape, h, (-3.0, -2.5, -3.5))
    pot = harmonic_potential(grid, 1.3)
    oh = PlainHamiltonian(grid, pot, "odd" if boundary else "zero", order)
    geom = _geometry(pkg, shape, h, order, boundary, torch.tensor(pot, device=device))
    gen = np.random.default_rng(5)
    x = gen.standard_normal((3, grid.n_points))
    print(f"[model] order {order} boundary {boundary} shape {shape}", flush=True)
    xd = torch.tensor(x, device=device)
    got = ops.hdiv_apply(xd, geom).cpu().numpy()
    want = oh.apply(x)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * np.abs(want).max())
    # a degree-4 filter with a fixed table: (c, a, b) rows as the solver builds them
    table = np.array([[0.7, 0.05, 0.0], [0.7, 0.1, 0.3], [0.7, 0.1, 0.4], [0.7, 0.1, 0.45]])
    y = ops.cheb_filter(xd, geom, torch.tensor(table, device=device), 4).cpu().numpy()
    prev, cur = x, (oh.apply(x) - table[0, 0] * x) * table[0, 1]
    for c, a, b in table[1:]:
        prev, cur = cur, (oh.apply(cur) - c * cur) * a - b * prev
    np.testing.assert_allclose(y, cur, rtol=0, atol=1e-12 * np.abs(cur).max())
    evals = np.array([0.3, 1.1, -0.2])
    q = np.full(grid.n_points, h**3)
    norms = torch.empty(3, dtype=torch.float64, device=device)
    ops.hdiv_resnorm(xd, geom, torch.tensor(evals, device=device), torch.tensor(q, device=device), norms)
    r = want - evals[:, None] * x
    np.testing.assert_allclose(norms.cpu().numpy(), np.sqrt((r * r * q).sum(1)), rtol=1e-12)


@pytest.mark.parametrize("sid", ["harmonic_w1", "box_L10"])
def test_model_systems_match_the_reference_fingerprint(pkg, device, sid):
    from cdft_amd import fingerprint as fp

    print(f"[model] {sid}", flush=True)
    got = fp.model_record(sid, device)
    ref = FINGERPRINT[sid]
    print(f"[model] {sid}: {got['eigenvalues']} ({got['n_iterations']}) reference {ref['eigenvalues']} "
          f"({ref['n_iterations']}) kinetic {got['energies']['kinetic']} / {ref['energies']['kinetic']}", flush=True)
    assert got["status"] == "valid" and got["n_iterations"] == ref["n_iterations"]
    assert got["grid"] == ref["grid"]
    for key in ("kinetic", "external", "total", "harris_foulkes"):
        assert abs(float(got["energies"][key]) - float(ref["energies"][key])) < 1e-10, key
    assert abs(float(got["eigenvalues"][0]) - float(ref["eigenvalues"][0])) < 1e-10
    assert got["occupations"] == ref["occupations"]
    failing = fp.diff({sid: ref}, {sid: got}, out=lambda *_: None)
    assert failing == 0


def test_harmonic_spectrum_against_the_analytic_levels(pkg, device):
    """The reference's G1.2 check (``gates/tier1.py:46-90``): the lowest harmonic-well eigenvalues against
    ``(n + 3/2) omega`` with their degeneracies (1, then the three-fold 5/2 shell) to 1e-6 Ha."""
    from cdft_amd import model

    r = model.solve_model("harmonic_w1", device, n_states=4)
    print(f"[model] harmonic spectrum {r['eigenvalues'].tolist()} ({r['n_iterations']})", flush=True)
    assert r["converged"]
    np.testing.assert_allclose(r["eigenvalues"], [1.5, 2.5, 2.5, 2.5], rtol=0, atol=1e-6)
