"""K3/K4 and the small dense kernels on the GPU against numpy float64.

Tolerances: reductions and rotations 1e-13 relative (summation order differs); Jacobi eigenvalues
1e-13 relative to the spectral radius, eigenvectors up to sign.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ka,kb,n", [(1, 1, 1), (5, 5, 274625), (3, 7, 1000), (8, 8, 4097), (11, 9, 5000)])
def test_wgram_matches_numpy(pkg, device, ka, kb, n):
    rng = np.random.default_rng(ka * 100 + kb)
    a = rng.standard_normal((ka, n))
    b = rng.standard_normal((kb, n))
    q = rng.uniform(0.1, 2.0, n)
    got = pkg.ops.wgram(torch.as_tensor(a).to(device), torch.as_tensor(b).to(device), torch.as_tensor(q).to(device),
                        scale=0.5).cpu().numpy()
    expect = 0.5 * (a * q) @ b.T
    assert np.abs(got - expect).max() <= 1e-13 * np.abs(expect).max() * max(1.0, np.sqrt(n) / 100)
    again = pkg.ops.wgram(torch.as_tensor(a).to(device), torch.as_tensor(b).to(device), torch.as_tensor(q).to(device),
                          scale=0.5).cpu().numpy()
    assert np.array_equal(got, again)


def test_rotate_and_lincomb(pkg, device):
    rng = np.random.default_rng(1)
    x = rng.standard_normal((5, 3001))
    r = rng.standard_normal((4, 5))
    xd = torch.as_tensor(x).to(device)
    got = pkg.ops.rotate(torch.as_tensor(r).to(device), xd).cpu().numpy()
    assert np.abs(got - r @ x).max() < 1e-13 * np.abs(r @ x).max()
    rt = rng.standard_normal((5, 3))
    got = pkg.ops.rotate(torch.as_tensor(rt).to(device), xd, transpose=True).cpu().numpy()
    assert np.abs(got - rt.T @ x).max() < 1e-13 * np.abs(rt.T @ x).max()
    dev = torch.tensor([2.0, 4.0], dtype=torch.float64, device=device)
    out = torch.empty(3001, dtype=torch.float64, device=device)
    pkg.ops.lincomb(out, [(xd[0], 1.0), (xd[1], -0.5, 1, 0), (xd[2], 3.0, 2, 1)], dev)
    expect = x[0] - 0.5 * 2.0 * x[1] + 3.0 / 4.0 * x[2]
    assert np.abs(out.cpu().numpy() - expect).max() < 1e-14 * np.abs(expect).max()


def test_resnorm(pkg, device):
    rng = np.random.default_rng(2)
    v = rng.standard_normal((5, 7777))
    av = rng.standard_normal((5, 7777))
    e = rng.standard_normal(5)
    q = rng.uniform(0.5, 1.5, 7777)
    got = pkg.ops.resnorm(torch.as_tensor(av).to(device), torch.as_tensor(v).to(device),
                          torch.as_tensor(e).to(device), torch.as_tensor(q).to(device)).cpu().numpy()
    expect = np.sqrt(((av - e[:, None] * v) ** 2 * q).sum(axis=1))
    assert np.abs(got - expect).max() < 1e-13 * expect.max()


@pytest.mark.parametrize("k", [1, 2, 5, 8, 16])
def test_cholesky_inverse_and_jacobi(pkg, device, k):
    rng = np.random.default_rng(k)
    a = rng.standard_normal((k, k + 3))
    s = a @ a.T
    sd = torch.as_tensor(s).to(device)
    t = torch.empty_like(sd)
    flag = torch.zeros(1, dtype=torch.int32, device=device)
    pkg.ops.chol_inv(sd, t, flag)
    tn = t.cpu().numpy()
    assert int(flag.item()) == 0
    assert np.abs(tn @ s @ tn.T - np.eye(k)).max() < 1e-12
    assert np.abs(np.triu(tn, 1)).max() == 0.0
    m = rng.standard_normal((k, k))
    m = m + m.T
    evals = torch.empty(k, dtype=torch.float64, device=device)
    vecs = torch.empty((k, k), dtype=torch.float64, device=device)
    pkg.ops.eigh(torch.as_tensor(m).to(device), evals, vecs)
    w, u = np.linalg.eigh(m)
    assert np.abs(evals.cpu().numpy() - w).max() < 1e-13 * np.abs(w).max()
    got = vecs.cpu().numpy()
    for j in range(k):
        assert min(np.abs(got[:, j] - u[:, j]).max(), np.abs(got[:, j] + u[:, j]).max()) < 1e-10
    assert np.abs(got.T @ got - np.eye(k)).max() < 1e-13


def test_cholesky_flags_indefinite(pkg, device):
    s = torch.tensor([[1.0, 2.0], [2.0, 1.0]], dtype=torch.float64, device=device)
    t = torch.empty_like(s)
    flag = torch.zeros(1, dtype=torch.int32, device=device)
    pkg.ops.chol_inv(s, t, flag)
    assert int(flag.item()) == 1


def test_device_chebyshev_table_is_the_host_table(pkg, device):
    """The filter's table built on the device equals the host rows bitwise, including the clamp of
    the lower edge into the bounds."""
    from cdft_amd.chefsi import chebyshev_coefficients

    rng = np.random.default_rng(5)
    for trial in range(20):
        smin = float(rng.uniform(-3.0, -0.5))
        smax = float(rng.uniform(20.0, 400.0))
        evals = np.sort(rng.uniform(smin - 1.0, 5.0, 5))
        if trial == 0:
            evals[-1] = smax + 3.0  # clamped from above
        if trial == 1:
            evals[-1] = smin - 2.0  # clamped from below
        degree = int(rng.integers(1, 25))
        table = torch.empty((degree, 3), dtype=torch.float64, device=device)
        pkg.ops.cheb_table(torch.as_tensor(evals).to(device), 4,
                           torch.tensor([smin, smax], dtype=torch.float64, device=device), degree, table)
        lower = min(max(float(evals[4]), smin + 1e-8), smax - 1e-8)
        expect = np.asarray(chebyshev_coefficients(degree, lower, smax, smin))
        assert np.array_equal(table.cpu().numpy(), expect), trial


@pytest.mark.parametrize("k,n", [(5, 274625), (3, 1001), (8, 4097), (1, 77)])
def test_rotate_gram_is_rotate_then_gram(pkg, device, k, n):
    """The fused rotation + weighted Gram gives bitwise the rotation and the Gram matrix of the two
    separate kernels."""
    rng = np.random.default_rng(k + n)
    x = torch.as_tensor(rng.standard_normal((k, n))).to(device)
    r = torch.as_tensor(rng.standard_normal((k, k))).to(device)
    q = torch.as_tensor(rng.uniform(0.1, 2.0, n)).to(device)
    y = torch.empty_like(x)
    gram = torch.empty((k, k), dtype=torch.float64, device=device)
    pkg.ops.rotate_gram(r, x, y, q, gram)
    y2 = pkg.ops.rotate(r, x, transpose=True)
    assert torch.equal(y, y2)
    assert torch.equal(gram, pkg.ops.wgram(y2, None, q))


@pytest.mark.parametrize("n", [1, 1000, 42244, 274625])
def test_reduce_dots_are_the_separate_reductions(pkg, device, n):
    """Several dot products in one pass give bitwise the separate reductions (scales, accumulation onto
    an earlier value and a third weight factor included), and match numpy."""
    ops = pkg.ops
    rng = np.random.default_rng(n)
    t = lambda: torch.as_tensor(rng.standard_normal(n)).to(device)  # noqa: E731
    a, b, c, d, e, w = t(), t(), t(), t(), t(), t()
    ref = torch.as_tensor(rng.standard_normal(6)).to(device)
    got = ref.clone()
    ops.reduce(ops.RED_DOT, a, ref, 0, b=b, scale=0.5)
    ops.reduce(ops.RED_DOT, c, ref, 2, b=d, accumulate=True)
    ops.reduce(ops.RED_DOT, e, ref, 3, b=a, w=w)
    ops.reduce(ops.RED_DOT, d, ref, 5)
    ops.reduce_dots([(a, b, None, 0, 0.5, False), (c, d, None, 2, 1.0, True), (e, a, w, 3, 1.0, False),
                     (d, None, None, 5, 1.0, False)], got)
    assert torch.equal(got, ref)
    host = lambda x: x.cpu().numpy()  # noqa: E731
    assert abs(float(got[3]) - float((host(e) * host(a) * host(w)).sum())) <= 1e-12 * max(1.0, np.abs(host(e) * host(a) * host(w)).sum())
    for nd in (1, 2):
        got2 = ref.clone()
        ops.reduce_dots([(a, b, None, 0, 0.5, False), (c, d, None, 2, 1.0, True)][:nd], got2)
        assert float(got2[0]) == float(ref[0])


@pytest.mark.parametrize("n,weighted", [(1, True), (3001, True), (274625, True), (4097, False)])
def test_lincomb_dots_is_lincomb_then_gram(pkg, device, n, weighted):
    """A linear combination fused with the weighted products of its result gives bitwise the separate
    combination and Gram calls (two products, and the norm, in place as the Lanczos step uses them)."""
    ops = pkg.ops
    rng = np.random.default_rng(n + weighted)
    t = lambda: torch.as_tensor(rng.standard_normal(n)).to(device)  # noqa: E731
    w, v, vp = t(), t(), t()
    q = torch.as_tensor(rng.uniform(0.1, 2.0, n)).to(device) if weighted else None
    dev = torch.as_tensor(rng.uniform(0.5, 2.0, 4)).to(device)
    terms_of = lambda out: [(out, 1.0), (v, -1.0, 1, 0), (vp, -0.5, 2, 1)]  # noqa: E731
    ref_w = w.clone()
    ops.lincomb(ref_w, terms_of(ref_w), dev)
    ref_dots = torch.cat([ops.wgram(ref_w.view(1, -1), v.view(1, -1), q).view(1),
                          ops.wgram(ref_w.view(1, -1), vp.view(1, -1), q).view(1)])
    got_w = w.clone()
    dots = torch.empty(2, dtype=torch.float64, device=device)
    ops.lincomb_dots(got_w, terms_of(got_w), dev, q, [v, vp], dots)
    assert torch.equal(got_w, ref_w) and torch.equal(dots, ref_dots)
    # the norm of the result, in place, with the root taken
    ref2 = ref_w.clone()
    ops.lincomb(ref2, [(ref2, 1.0), (v, -1.0, 1, 2)], dev)
    ref_norm = ops.wgram(ref2.view(1, -1), None, q, sqrt=True).view(1)
    got2 = ref_w.clone()
    nrm = torch.empty(1, dtype=torch.float64, device=device)
    ops.lincomb_dots(got2, [(got2, 1.0), (v, -1.0, 1, 2)], dev, q, [got2], nrm, sqrt=True)
    assert torch.equal(got2, ref2) and torch.equal(nrm, ref_norm)


def test_fused_lanczos_is_the_unfused_recurrence(pkg, device):
    """The Lanczos recurrence with its updates fused into the following products gives bitwise the
    scalars of the recurrence written with separate combination and product calls."""
    from cdft_amd import hamiltonian as hm

    ops = pkg.ops
    n, steps = 20000, 9
    rng = np.random.default_rng(21)
    diag = torch.as_tensor(rng.uniform(-2.0, 5.0, n)).to(device)
    q = torch.as_tensor(rng.uniform(0.2, 1.5, n)).to(device)
    start = torch.as_tensor(rng.standard_normal(n)).to(device)
    m = hm.DeviceMeasure(None, q)

    def apply(x, y):
        ops.ewise(ops.EW_MUL, y.view(-1), x.view(-1), diag)

    got = hm.lanczos_enqueue(apply, m, start, steps).clone()
    scal = torch.zeros(2 * steps + 3, dtype=torch.float64, device=device)
    a_slot, b_slot = (lambda i: i), (lambda i: steps + i)
    s1, s2, s0 = 2 * steps, 2 * steps + 1, 2 * steps + 2
    m.norm_into(start, scal[s0:s0 + 1])
    v = ops.lincomb(torch.empty(n, dtype=torch.float64, device=device), [(start, 1.0, 2, s0)], scal)
    v_prev = torch.zeros_like(v)
    w = torch.empty_like(v)
    for i in range(steps):
        apply(v.view(1, -1), w.view(1, -1))
        m.inner_into(v, w, scal[a_slot(i):a_slot(i) + 1])
        terms = [(w, 1.0), (v, -1.0, 1, a_slot(i))]
        if i > 0:
            terms.append((v_prev, -1.0, 1, b_slot(i - 1)))
        ops.lincomb(w, terms, scal)
        m.inner_into(w, v, scal[s1:s1 + 1])
        m.inner_into(w, v_prev, scal[s2:s2 + 1])
        ops.lincomb(w, [(w, 1.0), (v, -1.0, 1, s1), (v_prev, -1.0, 1, s2)], scal)
        m.norm_into(w, scal[b_slot(i):b_slot(i) + 1])
        v_new = ops.lincomb(torch.empty_like(v), [(w, 1.0, 2, b_slot(i))], scal)
        v_prev, v = v, v_new
    assert torch.equal(got, scal)
