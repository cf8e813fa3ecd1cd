"""K1/K2 parity on the GPU: the divergence-form apply and the fused Chebyshev filter against the oracle.

Tolerances: float64 kernels agree with the float64 oracle to 1e-13 of the largest output entry
(summation order and FMA contraction differ; the arithmetic is the same); float32 kernels to 2e-5.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle.core import CuspFactor, CuspHamiltonian, Grid, weighted_divergence, staggered_coefficients
from oracle.eigen import chebyshev_filter

pytestmark = pytest.mark.gpu


def _case(shape, spacing, charges, positions, seed=0):
    lengths = tuple((n - 1) * spacing for n in shape)
    grid = Grid.from_box(lengths, spacing)
    factor = CuspFactor(grid, charges, positions)
    return grid, CuspFactoredRef(grid, factor)


class CuspFactoredRef(CuspHamiltonian):
    pass


@pytest.mark.parametrize(
    "shape,spacing,charges,positions",
    [
        ((13, 13, 13), 0.5, (1.0,), [[0.0, 0.0, 0.0]]),
        ((17, 21, 19), 0.4, (2.0,), [[0.1, -0.2, 0.05]]),
        ((33, 29, 37), 0.5, (1.0, 1.0), [[0.0, 0.0, -0.7], [0.0, 0.0, 0.7]]),
        ((5, 3, 70), 0.3, (1.0,), [[0.0, 0.0, 0.0]]),
    ],
)
@pytest.mark.parametrize("k", [1, 3])
def test_apply_matches_oracle_f64(pkg, device, shape, spacing, charges, positions, k):
    grid, ref = _case(shape, spacing, charges, positions)
    geom = pkg.geometry_from_arrays(grid.shape, 8, grid.spacing, ref.face, ref.inverse_weight, ref.potential, device)
    gen = torch.Generator().manual_seed(k)
    phi = torch.rand((k, grid.n_points), generator=gen, dtype=torch.float64) * 2 - 1
    out = pkg.ops.hdiv_apply(phi.to(device), geom).cpu().numpy()
    expect = ref.apply(phi.numpy())
    scale = np.abs(expect).max(axis=1, keepdims=True)
    assert (np.abs(out - expect) / scale).max() < 1e-13


@pytest.mark.parametrize("order,shape", [(2, (11, 18, 23)), (4, (11, 18, 23)), (6, (11, 18, 23)), (8, (11, 18, 23)),
                                         (12, (11, 18, 23)), (12, (3, 5, 30)), (12, (20, 4, 9)), (12, (14, 17, 2)),
                                         (12, (29, 33, 40))])
def test_orders_match_oracle_with_random_weights(pkg, device, order, shape):
    """Random positive face weights, ``invw`` and ``V``: every point is well conditioned, so the
    comparison is per point. Order 12 (reach 11, three halo slots a thread, the 12-plane loop) also
    on grids thinner than its reach along each axis and on a grid of several tiles and segments."""
    rng = np.random.default_rng(order + sum(shape))
    n = int(np.prod(shape))
    faces = [rng.uniform(0.5, 1.5, size=tuple(s + (order if d == a else 0) for d, s in enumerate(shape))) for a in range(3)]
    invw = rng.uniform(0.5, 2.0, size=n)
    pot = rng.uniform(-1.0, 1.0, size=n)
    h = 0.37
    geom = pkg.geometry_from_arrays(shape, order, h, faces, invw, pot, device)
    phi = rng.uniform(-1, 1, size=(2, n))
    out = pkg.ops.hdiv_apply(torch.as_tensor(phi).to(device), geom).cpu().numpy()
    box = phi.reshape(2, *shape)
    expect = (weighted_divergence(box, faces, h, order).reshape(2, -1) * 0.5) * invw + phi * pot
    assert np.abs(out - expect).max() / np.abs(expect).max() < 1e-13


def test_apply_f32(pkg, device):
    grid, ref = _case((21, 21, 21), 0.5, (1.0,), [[0.0, 0.0, 0.0]])
    geom = pkg.geometry_from_arrays(grid.shape, 8, grid.spacing, ref.face, ref.inverse_weight, ref.potential, device,
                                    dtype=torch.float32)
    gen = torch.Generator().manual_seed(3)
    phi = torch.rand((4, grid.n_points), generator=gen, dtype=torch.float64) * 2 - 1
    out = pkg.ops.hdiv_apply(phi.to(device=device, dtype=torch.float32), geom).double().cpu().numpy()
    expect = ref.apply(phi.numpy())
    # float32 carries ~7 digits; the kinetic sum cancels about one of them at the cusp.
    assert np.abs(out - expect).max() / np.abs(expect).max() < 2e-5


def test_apply_is_bitwise_repeatable(pkg, device):
    grid, ref = _case((19, 19, 19), 0.5, (2.0,), [[0.0, 0.0, 0.0]])
    geom = pkg.geometry_from_arrays(grid.shape, 8, grid.spacing, ref.face, ref.inverse_weight, ref.potential, device)
    phi = torch.randn((5, grid.n_points), dtype=torch.float64, generator=torch.Generator().manual_seed(1)).to(device)
    a = pkg.ops.hdiv_apply(phi, geom)
    b = pkg.ops.hdiv_apply(phi, geom)
    assert torch.equal(a, b)


@pytest.mark.parametrize("degree", [1, 2, 3, 20])
def test_filter_matches_oracle(pkg, device, degree):
    grid, ref = _case((17, 17, 17), 0.5, (1.0,), [[0.0, 0.0, 0.0]])
    geom = pkg.geometry_from_arrays(grid.shape, 8, grid.spacing, ref.face, ref.inverse_weight, ref.potential, device)
    gen = torch.Generator().manual_seed(degree)
    block = torch.rand((5, grid.n_points), generator=gen, dtype=torch.float64) * 2 - 1
    lower, upper, smin = -0.3, 60.0, -0.55
    out = pkg.chebyshev_filter_device(block.to(device), geom, degree, lower, upper, smin).cpu().numpy()
    expect = chebyshev_filter(ref, block.numpy(), degree, lower, upper, smin)
    assert np.abs(out - expect).max() / np.abs(expect).max() < 1e-12


def test_order12_filter_and_gradient_match_oracle(pkg, device):
    """The order-12 cusp operator (the comparison order of the stencil-order gate) on an off-lattice
    two-centre geometry: a degree-20 filter against the oracle's, and the order-12 central gradient
    against the oracle's; order 12 is refused in float32."""
    from oracle.core import CuspFactor as OracleFactor, CuspHamiltonian as OracleHamiltonian, Grid as OracleGrid
    from cdft_amd.geometry import UniformGrid

    grid = OracleGrid.from_box((8.0, 7.0, 9.0), 0.5, anchor=(0.1, -0.2, 0.3), fd_order=12)
    ref = OracleHamiltonian(grid, OracleFactor(grid, (1.0, 2.0), [[0.1, -0.2, 0.3], [0.1, 0.4, -1.2]]))
    geom = pkg.geometry_from_arrays(grid.shape, 12, grid.spacing, ref.face, ref.inverse_weight, ref.potential, device)
    gen = torch.Generator().manual_seed(12)
    block = torch.rand((3, grid.n_points), generator=gen, dtype=torch.float64) * 2 - 1
    out = pkg.chebyshev_filter_device(block.to(device), geom, 20, -0.3, 90.0, -2.1).cpu().numpy()
    expect = chebyshev_filter(ref, block.numpy(), 20, -0.3, 90.0, -2.1)
    assert np.abs(out - expect).max() / np.abs(expect).max() < 1e-12
    mine = UniformGrid(grid.shape, grid.spacing, grid.origin, device, fd_order=12, gradient_order=12)
    got = mine.gradient(block.to(device)).cpu().numpy()
    want = grid.gradient(block.numpy()).reshape(3, 3, -1)
    assert np.abs(got - want).max() / np.abs(want).max() < 1e-13
    geom32 = pkg.geometry_from_arrays(grid.shape, 12, grid.spacing, ref.face, ref.inverse_weight, ref.potential,
                                      device, dtype=torch.float32)
    with pytest.raises(pkg._lib.HipLibraryError, match="float64 only"):
        pkg.ops.hdiv_apply(block.to(device=device, dtype=torch.float32), geom32)


def test_invalid_arguments_raise(pkg, device):
    grid, ref = _case((9, 9, 9), 0.5, (1.0,), [[0.0, 0.0, 0.0]])
    geom = pkg.geometry_from_arrays(grid.shape, 8, grid.spacing, ref.face, ref.inverse_weight, ref.potential, device)
    phi = torch.zeros((2, grid.n_points), dtype=torch.float64, device=device)
    with pytest.raises(ValueError):
        pkg.ops.hdiv_apply(phi[:, :-1], geom)
    with pytest.raises(pkg._lib.HipLibraryError):
        pkg._lib.call("cdft_hdiv_apply_f64", phi.data_ptr(), 0, 0, 0, 0, 0, 0, phi.data_ptr(), 2, 9, 9, 9,
                      0, 8, 0)


@pytest.mark.parametrize("order,shape,k,segment", [(8, (65, 65, 65), 5, 0), (8, (29, 25, 25), 3, 0),
                                                   (6, (17, 9, 40), 2, 4), (8, (6, 5, 7), 5, 0),
                                                   (4, (30, 31, 32), 4, 11), (12, (23, 19, 27), 3, 0),
                                                   (12, (9, 11, 13), 2, 5)])
def test_fused_residual_norms_match_apply_then_norm(pkg, device, order, shape, k, segment):
    """K1 with the residual-norm epilogue (nothing stored, block partials combined in fixed order)
    against the apply followed by the weighted norm of ``A x - e x``, with the oracle's numpy norm
    beside them; every segment split gives the same norms to 1e-13."""
    rng = np.random.default_rng(order + k)
    n = int(np.prod(shape))
    faces = [rng.uniform(0.5, 1.5, size=tuple(s + (order if d == a else 0) for d, s in enumerate(shape))) for a in range(3)]
    geom = pkg.geometry_from_arrays(shape, order, 0.31, faces, rng.uniform(0.5, 2.0, n), rng.uniform(-1, 1, n), device)
    x = torch.as_tensor(rng.uniform(-1, 1, size=(k, n))).to(device)
    q = torch.as_tensor(rng.uniform(0.1, 1.5, size=n)).to(device)
    evals = torch.as_tensor(rng.uniform(-2, 2, size=k)).to(device)
    pkg.ops.hdiv_set_segment(segment)
    try:
        fused = pkg.ops.hdiv_resnorm(x, geom, evals, q, torch.empty(k, dtype=torch.float64, device=device))
    finally:
        pkg.ops.hdiv_set_segment(0)
    ax = pkg.ops.hdiv_apply(x, geom)
    two_pass = pkg.ops.resnorm(ax, x, evals, q, out=torch.empty(k, dtype=torch.float64, device=device))
    r = ax.cpu().numpy() - evals.cpu().numpy()[:, None] * x.cpu().numpy()
    numpy_norm = np.sqrt((r * r * q.cpu().numpy()).sum(axis=1))
    assert np.abs(fused.cpu().numpy() - two_pass.cpu().numpy()).max() <= 1e-13 * numpy_norm.max()
    assert np.abs(fused.cpu().numpy() - numpy_norm).max() <= 1e-13 * numpy_norm.max()
    again = pkg.ops.hdiv_resnorm(x, geom, evals, q, torch.empty(k, dtype=torch.float64, device=device))
    assert torch.equal(again, pkg.ops.hdiv_resnorm(x, geom, evals, q, torch.empty(k, dtype=torch.float64, device=device)))


@pytest.mark.parametrize("k,degree", [(5, 20), (1, 3), (8, 12)])
def test_residual_norm_tail_is_the_separate_calls(pkg, device, k, degree):
    """The residual-norm launch with the pass's tail (the next filter's Chebyshev table and the
    orthonormality error in the combine launch) gives bitwise the residual norms, ``cheb_table`` and
    ``orth_error`` of the separate calls; the table also equals the oracle's host rows."""
    from oracle.eigen import chebyshev_coefficients

    rng = np.random.default_rng(40 + k)
    shape = (29, 25, 27)
    n = int(np.prod(shape))
    faces = [rng.uniform(0.5, 1.5, size=tuple(s + (8 if d == a else 0) for d, s in enumerate(shape))) for a in range(3)]
    geom = pkg.geometry_from_arrays(shape, 8, 0.35, faces, rng.uniform(0.5, 2.0, n), rng.uniform(-1, 1, n), device)
    f = dict(dtype=torch.float64, device=device)
    x = torch.as_tensor(rng.uniform(-1, 1, size=(k, n))).to(device)
    q = torch.as_tensor(rng.uniform(0.1, 1.5, size=n)).to(device)
    evals = torch.as_tensor(np.sort(rng.uniform(-2, 2, size=k))).to(device)
    bounds = torch.tensor([-2.5, 40.0], **f)
    gram = torch.eye(k, **f) + torch.as_tensor(rng.uniform(-1e-9, 1e-9, size=(k, k))).to(device)
    norms_a, norms_b = torch.empty(k, **f), torch.empty(k, **f)
    table_a, table_b = torch.full((degree, 3), 7.0, **f), torch.full((degree, 3), -7.0, **f)
    orth_a, orth_b = torch.full((1,), 7.0, **f), torch.full((1,), -7.0, **f)
    pkg.ops.hdiv_resnorm(x, geom, evals, q, norms_a)
    pkg.ops.cheb_table(evals, k - 1, bounds, degree, table_a)
    pkg.ops.orth_error(gram, orth_a)
    pkg.ops.hdiv_resnorm(x, geom, evals, q, norms_b, tail=(k - 1, bounds, degree, table_b, gram, orth_b))
    torch.cuda.synchronize()
    assert torch.equal(norms_a, norms_b)
    assert torch.equal(table_a, table_b)
    assert torch.equal(orth_a, orth_b)
    assert orth_b.item() == (gram - torch.eye(k, **f)).abs().max().item()
    ev = evals.cpu().numpy()
    host = chebyshev_coefficients(degree, float(ev[k - 1]), 40.0, -2.5)  # lower inside the clamp
    assert np.array_equal(table_b.cpu().numpy(), np.asarray(host, dtype=np.float64).reshape(degree, 3))
    with pytest.raises(Exception):
        pkg.ops.hdiv_resnorm(x, geom, evals, q, norms_b, tail=(k, bounds, degree, table_b, gram, orth_b))


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_geometry_records_hold_the_separate_arrays(pkg, device, dtype):
    """The per-plane geometry records are the face weights and invw at each column of the face-extended
    rows (the newest x-face, the y- and z-faces, invw), zero where an array has no such entry."""
    shape, order = (7, 5, 9), 8
    rng = np.random.default_rng(4)
    n = int(np.prod(shape))
    faces = [rng.uniform(0.5, 1.5, size=tuple(s + (order if d == a else 0) for d, s in enumerate(shape))) for a in range(3)]
    invw = rng.uniform(0.5, 2.0, n)
    geom = pkg.geometry_from_arrays(shape, order, 0.25, faces, invw, rng.uniform(-1, 1, n), device, dtype=dtype)
    n0, n1, n2 = shape
    rec = geom.packed.view(n0, n1 + order, n2 + order, 4).double().cpu().numpy()
    cast = lambda a: np.asarray(torch.as_tensor(a).to(dtype).double())  # noqa: E731
    w0, w1, w2, iw = cast(faces[0]), cast(faces[1]), cast(faces[2]), cast(invw).reshape(shape)
    expect = np.zeros_like(rec)
    expect[:, :n1, :n2, 0] = w0[order - 1: order - 1 + n0]
    expect[:, :, :n2, 1] = w1
    expect[:, :n1, :, 2] = w2
    expect[:, :n1, :n2, 3] = iw
    assert np.array_equal(rec, expect)


def test_outputs_do_not_depend_on_the_tiling(pkg, device):
    """Every tiling (largest tile rows along y 1..16, width 13 or 16 along z) and segment length gives
    bitwise the same apply, transpose and Chebyshev steps (each point's sums take the same operands
    in the same order); the residual norms, summed over per-block partials, agree to round-off."""
    from cdft_amd.chefsi import scaled_chebyshev_table

    ops = pkg.ops
    for shape, k in (((61, 57, 57), 2), ((20, 23, 31), 3), ((33, 33, 33), 5)):
        grid, ref = _case(shape, 0.4, (1.0, 1.0), [[-0.7, 0.05, 0.0], [0.7, 0.05, 0.0]])
        geom = pkg.geometry_from_arrays(grid.shape, 8, grid.spacing, ref.face, ref.inverse_weight, ref.potential,
                                        device)
        gen = torch.Generator().manual_seed(shape[0])
        x = (torch.rand((k, grid.n_points), generator=gen, dtype=torch.float64) * 2 - 1).to(device)
        p = (torch.rand((k, grid.n_points), generator=gen, dtype=torch.float64) * 2 - 1).to(device)
        q = torch.rand(grid.n_points, generator=gen, dtype=torch.float64).to(device)
        evals = torch.linspace(-1.0, 1.0, k, dtype=torch.float64, device=device)
        table = torch.tensor(scaled_chebyshev_table(3, -0.3, 30.0, -1.2), dtype=torch.float64, device=device)

        def run():
            norms = torch.empty(k, dtype=torch.float64, device=device)
            ops.hdiv_resnorm(x, geom, evals, q, norms)
            return (ops.hdiv_apply(x, geom), ops.hdiv_apply_adjoint(x, geom),
                    ops.cheb_step(x, None, geom, table, 0, torch.empty_like(x)),
                    ops.cheb_step(x, p, geom, table, 1, torch.empty_like(x)), norms)

        base = run()
        try:
            ops.hdiv_set_kernel(1)  # the column kernel for every launch (its tiling is what varies)
            for tile, segment in (((16, 13), 0), ((13, 13), 5), ((8, 16), 0), ((3, 13), 4), ((1, 16), 11)):
                ops.hdiv_set_tile(*tile)
                ops.hdiv_set_segment(segment)
                shape_used = ops.hdiv_launch_shape(grid.shape, k)
                assert shape_used["tile_z"] == tile[1] and shape_used["tile_y"] <= tile[0]
                got = run()
                for a, b in zip(got[:4], base[:4]):
                    assert torch.equal(a, b), (shape, tile, segment)
                assert torch.allclose(got[4], base[4], rtol=1e-13, atol=0), (shape, tile, segment)
        finally:
            ops.hdiv_set_kernel(0)
            ops.hdiv_set_tile(0, 0)
            ops.hdiv_set_segment(0)
    with pytest.raises(Exception):
        ops.hdiv_set_tile(17, 0)
    with pytest.raises(Exception):
        ops.hdiv_set_tile(0, 14)


def _random_geometry(pkg, device, shape, seed, dtype):
    rng = np.random.default_rng(seed)
    n = int(np.prod(shape))
    faces = [rng.uniform(0.5, 1.5, size=tuple(s + (8 if d == a else 0) for d, s in enumerate(shape))) for a in range(3)]
    invw = rng.uniform(0.5, 2.0, size=n)
    pot = rng.uniform(-1.0, 1.0, size=n)
    geom = pkg.geometry_from_arrays(shape, 8, 0.3, faces, invw, pot, device)
    if dtype == torch.float32:
        geom = geom.to(torch.float32)
    return geom


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_vector_kernel_is_the_column_kernel_bitwise(pkg, device, dtype):
    """The vector kernel (each thread owns a 16-byte vector of adjacent z-columns) gives bitwise the
    column kernel's apply and Chebyshev steps (modes 0, 1, 2), for grids whose rows end inside a
    vector (n2 odd or not a multiple of four), grids thinner than the stencil's reach along y, and
    forced launch settings of every tile width, several tile heights and segment lengths."""
    from cdft_amd.autotune import VEC_WIDTHS
    from cdft_amd.chefsi import scaled_chebyshev_table

    ops = pkg.ops
    cases = [((65, 65, 65), 5), ((5, 3, 70), 3), ((11, 18, 23), 1), ((20, 7, 33), 2), ((9, 41, 57), 8)]
    for shape, k in cases:
        geom = _random_geometry(pkg, device, shape, shape[2], dtype)
        gen = torch.Generator().manual_seed(shape[1])
        x = (torch.rand((k, geom.n_points), generator=gen, dtype=torch.float64) * 2 - 1).to(device, dtype)
        p = (torch.rand((k, geom.n_points), generator=gen, dtype=torch.float64) * 2 - 1).to(device, dtype)
        table = torch.tensor(scaled_chebyshev_table(3, -0.3, 30.0, -1.2), dtype=torch.float64, device=device)

        def run():
            return (ops.hdiv_apply(x, geom), ops.cheb_step(x, None, geom, table, 0, torch.empty_like(x)),
                    ops.cheb_step(x, p, geom, table, 1, torch.empty_like(x)))

        try:
            ops.hdiv_set_kernel(1)
            column = run()
            ops.hdiv_set_kernel(0)
            settings = [None] + [(ty, tz, seg) for tz in VEC_WIDTHS[dtype] for ty, seg in ((17, 0), (5, 3), (11, 7))]
            for setting in settings:
                if setting is None:
                    ops.hdiv_vec_set(shape, k, dtype, 0, 0, 0)
                else:
                    ops.hdiv_vec_set(shape, k, dtype, *setting)
                    used = ops.hdiv_vec_shape(shape, k, dtype)
                    assert used["tile_z"] == setting[1] and used["tile_y"] <= setting[0]
                for a, b in zip(run(), column):
                    assert torch.equal(a, b), (shape, k, dtype, setting)
        finally:
            ops.hdiv_set_kernel(0)
            ops.hdiv_vec_set(shape, k, dtype, 0, 0, 0)


def test_tuned_launch_is_bitwise_the_default(pkg, device):
    """The single-grid tuner times candidate vector-kernel settings on the device and puts the fastest
    in force for the shape; the apply and the steps are bitwise those of the built-in setting."""
    from cdft_amd import autotune
    from cdft_amd.chefsi import scaled_chebyshev_table

    ops = pkg.ops
    shape, k = (33, 29, 37), 5
    geom = _random_geometry(pkg, device, shape, 7, torch.float64)
    gen = torch.Generator().manual_seed(3)
    x = (torch.rand((k, geom.n_points), generator=gen, dtype=torch.float64) * 2 - 1).to(device)
    p = (torch.rand((k, geom.n_points), generator=gen, dtype=torch.float64) * 2 - 1).to(device)
    table = torch.tensor(scaled_chebyshev_table(3, -0.3, 30.0, -1.2), dtype=torch.float64, device=device)
    ops.hdiv_vec_set(shape, k, torch.float64, 0, 0, 0)
    base = (ops.hdiv_apply(x, geom), ops.cheb_step(x, p, geom, table, 1, torch.empty_like(x)))
    try:
        tuning = autotune.tune_stencil(geom, k)
        assert tuning is not None and len(tuning["us"]) >= 2
        assert all(t > 0 for t in tuning["us"].values())
        got = (ops.hdiv_apply(x, geom), ops.cheb_step(x, p, geom, table, 1, torch.empty_like(x)))
        for a, b in zip(got, base):
            assert torch.equal(a, b)
    finally:
        autotune._stencil_cache.pop((shape, k, torch.float64, geom.device), None)
        ops.hdiv_vec_set(shape, k, torch.float64, 0, 0, 0)
