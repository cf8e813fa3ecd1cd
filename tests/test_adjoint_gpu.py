"""The backward of the hot path: the transpose H-apply, HApply's autograd, and the Sternheimer
gradient of a density loss against torch autograd through a dense symmetric eigendecomposition of
the oracle operator, and against central finite differences of the GPU forward map.

Tolerances (float64): transpose identity 1e-13 relative; HApply gradients 1e-12; the Sternheimer
gradient against the dense autograd gradient 1e-7 of its max-norm (both eigensolves converge to
residual ~1e-11, the CG to 1e-11); finite differences 1e-6 relative (step 1e-4 along smooth
directions, central differences, error O(step^2) plus eigensolve noise / step).
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle.core import CuspFactor, CuspHamiltonian, Grid

pytestmark = pytest.mark.gpu


def _small_case(pkg, device, n=13, spacing=0.5, charge=2.0, position=(0.11, -0.07, 0.05)):
    """Plain-measure cusp operator (no quadrature) of one nucleus on an n^3 grid: oracle and GPU. The
    nucleus sits off the lattice so no two eigenvalues are degenerate (torch's eigh backward divides
    by every eigenvalue difference)."""
    from cdft_amd.geometry import UniformGrid

    length = (n - 1) * spacing
    grid_o = Grid.from_box((length,) * 3, spacing)
    factor = CuspFactor(grid_o, (charge,), [list(position)])
    ham_o = CuspHamiltonian(grid_o, factor)
    grid = UniformGrid.from_box((length,) * 3, spacing, device)
    geom = pkg.geometry_from_arrays(grid_o.shape, 8, spacing, ham_o.face, ham_o.inverse_weight, ham_o.potential, device)
    q = torch.as_tensor(ham_o.measure.quadrature).to(device)
    return grid_o, ham_o, grid, geom, q


def _bumps(points, seed, count=8, amplitude=0.05, width=1.0, radius=2.0):
    rng = np.random.default_rng(seed)
    a = rng.uniform(-amplitude, amplitude, count)
    d = rng.normal(size=(count, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    c = d * (radius * rng.uniform(0, 1, count) ** (1.0 / 3.0))[:, None]
    out = np.zeros(points.shape[0])
    for j in range(count):
        out += a[j] * np.exp(-((points - c[j]) ** 2).sum(-1) / (2 * width**2))
    return out


def test_adjoint_apply_is_the_transpose(pkg, device):
    grid_o, ham_o, grid, geom, q = _small_case(pkg, device, n=17)
    gen = torch.Generator().manual_seed(3)
    x = (torch.rand((2, grid.n_points), generator=gen, dtype=torch.float64) * 2 - 1).to(device)
    y = (torch.rand((2, grid.n_points), generator=gen, dtype=torch.float64) * 2 - 1).to(device)
    ax = pkg.ops.hdiv_apply(x, geom)
    aty = pkg.ops.hdiv_apply_adjoint(y, geom)
    lhs = (y * ax).sum(dim=1)
    rhs = (aty * x).sum(dim=1)
    assert torch.allclose(lhs, rhs, rtol=1e-13, atol=0.0)
    # against the oracle: A^T y = Q A Q^-1 y (A is self-adjoint in the measure q)
    expect = ham_o.apply(y.cpu().numpy() / ham_o.measure.quadrature) * ham_o.measure.quadrature
    got = aty.cpu().numpy()
    assert np.abs(got - expect).max() <= 1e-12 * np.abs(expect).max()


def test_happly_autograd(pkg, device):
    from cdft_amd.adjoint import HApply

    _, _, grid, geom, _ = _small_case(pkg, device, n=13)
    gen = torch.Generator().manual_seed(4)
    x = (torch.rand((3, grid.n_points), generator=gen, dtype=torch.float64) * 2 - 1).to(device).requires_grad_()
    v = (0.1 * torch.rand(grid.n_points, generator=gen, dtype=torch.float64)).to(device).requires_grad_()
    g = (torch.rand((3, grid.n_points), generator=gen, dtype=torch.float64) * 2 - 1).to(device)
    y = HApply.apply(x, v, geom)
    (y * g).sum().backward()
    geom_v = geom.with_potential(v.detach())
    assert torch.allclose(x.grad, pkg.ops.hdiv_apply_adjoint(g, geom_v), rtol=1e-12, atol=1e-12)
    assert torch.allclose(v.grad, (g * x.detach()).sum(dim=0), rtol=1e-12, atol=1e-14)


def test_sternheimer_gradient_matches_dense_autograd(pkg, device):
    from cdft_amd.adjoint import KohnShamDensity, KohnShamProblem, density_loss

    grid_o, ham_o, grid, geom, q = _small_case(pkg, device, n=13)
    n = grid_o.n_points
    qn = ham_o.measure.quadrature
    weights = qn * ham_o.factor.weight  # omega f^2 of the plain measure
    v_t = _bumps(grid_o.points(), seed=0)
    # dense reference: B = Q^1/2 A Q^-1/2 (symmetric), autograd through eigh
    eye = np.eye(n)
    a_dense = ham_o.apply(eye).T  # column j = A e_j
    sq = np.sqrt(qn)
    b0 = (sq[:, None] * a_dense) / sq[None, :]
    b0 = 0.5 * (b0 + b0.T)
    b0_t = torch.as_tensor(b0)

    def dense_rho(vv):
        _, vecs = torch.linalg.eigh(b0_t + torch.diag(vv))
        phi = vecs[:, 0] / torch.as_tensor(sq)
        return 2.0 * phi**2

    rho_t_dense = dense_rho(torch.as_tensor(v_t)).detach()
    v0 = torch.zeros(n, dtype=torch.float64, requires_grad=True)
    loss_dense = density_loss(dense_rho(v0), rho_t_dense, torch.as_tensor(weights))
    loss_dense.backward()
    grad_dense = v0.grad.numpy()

    problem = KohnShamProblem(grid, geom, q, [2.0])
    rho_t, _ = problem.forward(torch.as_tensor(v_t).to(device))
    problem._warm = None
    v = torch.zeros(n, dtype=torch.float64, device=device, requires_grad=True)
    rho = KohnShamDensity.apply(v, problem)
    assert torch.allclose(rho.detach().cpu(), dense_rho(torch.zeros(n, dtype=torch.float64)).detach(), rtol=0,
                          atol=1e-9 * float(rho.detach().abs().max()))
    loss = density_loss(rho, rho_t.detach(), torch.as_tensor(weights).to(device))
    loss.backward()
    assert abs(float(loss) - float(loss_dense)) <= 1e-8 * abs(float(loss_dense))
    grad = v.grad.cpu().numpy()
    assert problem.history[-1]["sternheimer"][0]["converged"]
    assert np.abs(grad - grad_dense).max() <= 1e-7 * np.abs(grad_dense).max()


def test_sternheimer_gradient_matches_finite_differences(pkg, device):
    from cdft_amd.adjoint import KohnShamDensity, KohnShamProblem, density_loss

    grid_o, ham_o, grid, geom, q = _small_case(pkg, device, n=17)
    pts = grid_o.points()
    weights = torch.as_tensor(ham_o.measure.quadrature * ham_o.factor.weight).to(device)
    problem = KohnShamProblem(grid, geom, q, [2.0])
    rho_t, _ = problem.forward(torch.as_tensor(_bumps(pts, seed=1)).to(device))
    v = torch.zeros(grid.n_points, dtype=torch.float64, device=device, requires_grad=True)
    loss = density_loss(KohnShamDensity.apply(v, problem), rho_t, weights)
    loss.backward()
    grad = v.grad.detach()
    step = 1e-4
    for seed in (11, 12, 13):
        d = torch.as_tensor(_bumps(pts, seed=seed, amplitude=1.0)).to(device)
        lp = density_loss(problem.forward(step * d)[0], rho_t, weights)
        lm = density_loss(problem.forward(-step * d)[0], rho_t, weights)
        fd = float(lp - lm) / (2 * step)
        an = float((grad * d).sum())
        assert abs(fd - an) <= 1e-6 * max(abs(an), 1e-30), (seed, fd, an)


def test_sternheimer_gradient_pointwise_finite_differences(pkg, device):
    """dL/dV at 64 seeded grid points (inside the region where the density lives) against central
    differences of the GPU forward map with V perturbed at that one point (step 1e-3)."""
    from cdft_amd.adjoint import KohnShamDensity, KohnShamProblem, density_loss

    grid_o, ham_o, grid, geom, q = _small_case(pkg, device, n=17)
    pts = grid_o.points()
    weights = torch.as_tensor(ham_o.measure.quadrature * ham_o.factor.weight).to(device)
    problem = KohnShamProblem(grid, geom, q, [2.0])
    rho_t, _ = problem.forward(torch.as_tensor(_bumps(pts, seed=1)).to(device))
    v = torch.zeros(grid.n_points, dtype=torch.float64, device=device, requires_grad=True)
    loss = density_loss(KohnShamDensity.apply(v, problem), rho_t, weights)
    loss.backward()
    grad = v.grad.detach().cpu().numpy()
    scale = np.abs(grad).max()
    inner = np.nonzero(np.linalg.norm(pts, axis=1) < 2.5)[0]
    picks = np.random.default_rng(7).choice(inner, size=64, replace=False)
    step = 1e-3
    worst = 0.0
    for p in picks:
        d = torch.zeros(grid.n_points, dtype=torch.float64, device=device)
        d[int(p)] = 1.0
        lp = density_loss(problem.forward(step * d)[0], rho_t, weights)
        lm = density_loss(problem.forward(-step * d)[0], rho_t, weights)
        fd = float(lp - lm) / (2 * step)
        worst = max(worst, abs(fd - grad[p]))
    print(f"pointwise FD: worst |fd - grad| = {worst:.3e}, max |grad| = {scale:.3e}", flush=True)
    assert worst <= 1e-6 * scale


def test_sternheimer_gradient_on_converged_lda_potential(pkg, device):
    """The inversion step's own setting at reduced size: He LDA on the 16-bohr box at h = 0.5 (33^3),
    the quadrature measure (omega / h^3) and the converged LDA effective potential as V_0, the loss
    in the cusp-aware measure (omega f^2). Central differences along three smooth seeded directions
    agree with the Sternheimer gradient, relative to ``sum |dL/dV_p d_p|`` (the scale of the directional
    derivative before its cancellations): at the configuration's step 1e-6 to 1e-5 (the forward solves
    converge to residual 1e-11, so the loss carries about 1e-13 of noise, which the difference quotient
    divides by 2e-6) and at step 1e-4 to 1e-6. The conjugate-gradient solves converge (below their
    1e-11 tolerance, well before the cap)."""
    from cdft_amd import scf
    from cdft_amd.adjoint import KohnShamDensity, KohnShamProblem, density_loss
    from cdft_amd.contract import Numerics, helium_lda

    r = scf.solve_self_consistent(helium_lda(), Numerics(box_lengths=(16.0,) * 3, spacing=0.5), device)
    step_ = r["step"]
    ham = step_.hamiltonian({"effective": r["effective_potential"]})
    grid = step_.grid
    n0, n1, n2 = grid.shape
    axes = [np.arange(n) * grid.spacing + o for n, o in zip(grid.shape, grid.origin)]
    pts = np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1).reshape(-1, 3)
    weights = step_.q.mass_weights * step_.f2
    problem = KohnShamProblem(grid, ham.geometry, ham.measure.q, [2.0])
    rho_t, _ = problem.forward(torch.as_tensor(_bumps(pts, seed=1)).to(device))
    v = torch.zeros(grid.n_points, dtype=torch.float64, device=device, requires_grad=True)
    loss = density_loss(KohnShamDensity.apply(v, problem), rho_t, weights)
    loss.backward()
    grad = v.grad.detach()
    cg = problem.history[-1]["sternheimer"][0]
    assert cg["converged"] and cg["iterations"] < problem.cg.max_iterations, cg
    for step, tol in ((1e-6, 1e-5), (1e-4, 1e-6)):
        for seed in (11, 12, 13):
            d = torch.as_tensor(_bumps(pts, seed=seed, amplitude=1.0)).to(device)
            lp = float(density_loss(problem.forward(step * d)[0], rho_t, weights))
            lm = float(density_loss(problem.forward(-step * d)[0], rho_t, weights))
            fd = (lp - lm) / (2 * step)
            an = float((grad * d).sum())
            scale = float((grad * d).abs().sum())
            print(f"step {step:g} seed {seed}: fd {fd:.12e} analytic {an:.12e} error/scale {abs(fd - an) / scale:.2e}",
                  flush=True)
            assert abs(fd - an) <= tol * scale, (step, seed, fd, an, scale)


def test_checkpointed_filter_gradient(pkg, device):
    """The differentiable Chebyshev filter (degree 20) with checkpointing of its iterates: the forward is
    bitwise the fused filter; the gradients for the block and the potential with a checkpoint every 4
    steps are bitwise those with every iterate stored (the recomputed iterates are the forward's), agree
    with torch autograd through a chain of HApply calls to 1e-10 of their max-norm, and with central
    differences (step 1e-6) along a seeded potential direction to 1e-7 relative."""
    from cdft_amd.adjoint import ChebyshevFilterFunction, HApply
    from cdft_amd.chefsi import scaled_chebyshev_table

    grid_o, ham_o, grid, geom, q = _small_case(pkg, device, n=13)
    n = grid.n_points
    rows = scaled_chebyshev_table(20, 0.5, 40.0, -2.5)
    gen = torch.Generator().manual_seed(21)
    x0 = (torch.rand((3, n), generator=gen, dtype=torch.float64) * 2 - 1).to(device)
    v0 = (0.1 * torch.rand(n, generator=gen, dtype=torch.float64)).to(device) + geom.potential
    weight = (torch.rand((3, n), generator=gen, dtype=torch.float64) * 2 - 1).to(device)

    def run(every):
        x = x0.clone().requires_grad_(True)
        v = v0.clone().requires_grad_(True)
        y = ChebyshevFilterFunction.apply(x, v, geom, rows, every)
        (weight * y).sum().backward()
        return y.detach(), x.grad, v.grad

    y4, gx4, gv4 = run(4)
    y1, gx1, gv1 = run(1)
    table = torch.tensor(rows, dtype=torch.float64, device=device)
    assert torch.equal(y4, pkg.ops.cheb_filter(x0, geom.with_potential(v0), table, 20))
    assert torch.equal(gx4, gx1) and torch.equal(gv4, gv1)

    # autograd through the recurrence written with HApply (every iterate kept by autograd)
    x = x0.clone().requires_grad_(True)
    v = v0.clone().requires_grad_(True)
    prev, y = None, x
    for j, (c, a, b) in enumerate(rows):
        nxt = (HApply.apply(y, v, geom) - c * y) * a
        if j > 0:
            nxt = nxt - b * prev
        prev, y = y, nxt
    (weight * y).sum().backward()
    assert (gx4 - x.grad).abs().max() <= 1e-10 * x.grad.abs().max()
    assert (gv4 - v.grad).abs().max() <= 1e-10 * v.grad.abs().max()

    d = torch.as_tensor(_bumps(grid_o.points(), seed=5, amplitude=1.0)).to(device)
    step = 1e-6

    def loss(vv):
        return float((weight * pkg.ops.cheb_filter(x0, geom.with_potential(vv.contiguous()), table, 20)).sum())

    fd = (loss(v0 + step * d) - loss(v0 - step * d)) / (2 * step)
    an = float((gv4 * d).sum())
    assert abs(fd - an) <= 1e-7 * abs(an), (fd, an)
