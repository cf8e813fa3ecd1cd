"""Host-side logic that needs no GPU: the memoised seeded draws (the reference's start vectors),
the stencil weights, the Chebyshev table, the batched-scan geometry rule and the lumped-weight
axis tables."""

from __future__ import annotations

import numpy as np
import torch


def _torch_sequence(seed, sizes):
    gen = torch.Generator().manual_seed(seed)
    return [torch.randn(s, generator=gen, dtype=torch.float64) for s in sizes]


def test_seeded_draws_reproduce_torch_with_memo_and_prefetch(pkg):
    from cdft_amd.hamiltonian import SeededDraws, prefetch_draws, seeded_randn

    store: dict = {}
    prefetch_draws(store, 0, [((5, 300), (300,)), ((300,),)])
    for sizes in ([(5, 300), (300,)], [(300,)], [(5, 300), (300,), (2, 300)], [(300,), (300,)]):
        expect = _torch_sequence(0, sizes)
        for _ in range(2):  # second pass is served from the memo
            draws = SeededDraws(0, store)
            got = [seeded_randn(s, draws, "cpu") for s in sizes]
            assert all(torch.equal(a, b) for a, b in zip(got, expect))


def test_chunked_draws_are_the_serial_draws(pkg, monkeypatch):
    """Draws split into chunks of whole 16-value groups, each drawn on a host thread from the
    Mersenne-Twister state at its first uniform, equal the serial torch draw bit for bit: sizes
    with and without a partial last group, chunks that straddle twists, seeds above 2^32."""
    from cdft_amd import hamiltonian as hm

    assert hm._layout_verified()  # otherwise parallel_randn would fall back to the serial draw
    monkeypatch.setattr(hm, "_CHUNK_VALUES", 48)
    for seed in (0, 5, 2**32 + 17):
        sizes = [(3, 1001), (624,), (16,), (17,), (5, 320), (311,)]
        got = hm.parallel_randn(seed, sizes)
        expect = _torch_sequence(seed, sizes)
        assert all(torch.equal(a, b) for a, b in zip(got, expect))
    with_small = hm.parallel_randn(3, [(40,), (7,)])  # a size below 16 takes torch's scalar path
    assert all(torch.equal(a, b) for a, b in zip(with_small, _torch_sequence(3, [(40,), (7,)])))


def test_batch_starts_share_the_block_prefix_bitwise(pkg, monkeypatch):
    """The batched solver's start vectors: for every grid size the (k, n) block and the (n,) Lanczos
    start equal a fresh generator's serial draws, although the blocks' common prefix is drawn once
    (sizes with and without a partial last group, k = 1 and 5, chunks that straddle twists)."""
    from cdft_amd import hamiltonian as hm

    monkeypatch.setattr(hm, "_CHUNK_VALUES", 64)
    for seed, k, sizes in ((0, 5, [17, 35, 100, 1000, 3111]), (2**32 + 3, 1, [16, 17, 31, 33, 1025]),
                           (9, 3, [250, 343])):
        got = hm.seeded_starts(seed, k, sizes, torch.device("cpu"))
        for n, (block, start) in zip(sizes, got):
            expect_block, expect_start = _torch_sequence(seed, [(k, n), (n,)])
            assert torch.equal(block, expect_block) and torch.equal(start, expect_start), (seed, k, n)
    small = hm.seeded_starts(1, 1, [7, 40], torch.device("cpu"))  # below 16: the serial draw
    for n, (block, start) in zip([7, 40], small):
        expect_block, expect_start = _torch_sequence(1, [(1, n), (n,)])
        assert torch.equal(block, expect_block) and torch.equal(start, expect_start)


def test_mersenne_states_match_the_stream(pkg):
    """The states the library reports continue the stream exactly: numbers drawn from a generator
    placed at position p equal the numbers after p outputs of the seeded generator."""
    from cdft_amd import hamiltonian as hm

    for seed, skip in ((1, 0), (1, 312), (1, 313), (99, 1000), (2**33 + 1, 4)):
        ref = torch.Generator().manual_seed(seed)
        torch.rand(skip, generator=ref, dtype=torch.float64)
        key, index = hm._states(seed, [2 * skip])
        placed = hm._generator_at(seed, key[0], int(index[0]))
        assert torch.equal(torch.rand(700, generator=placed, dtype=torch.float64),
                           torch.rand(700, generator=ref, dtype=torch.float64))


def test_staggered_weights_closed_forms(pkg):
    from cdft_amd.stencil import staggered_coefficients

    assert list(staggered_coefficients(2)) == [-1.0, 1.0]
    assert np.allclose(staggered_coefficients(4), [1 / 24, -9 / 8, 9 / 8, -1 / 24], rtol=0, atol=1e-16)


def test_chebyshev_rows_match_the_oracle(pkg):
    from cdft_amd.chefsi import chebyshev_coefficients
    from oracle.eigen import chebyshev_coefficients as oracle_rows

    assert np.array_equal(np.asarray(chebyshev_coefficients(20, -0.3, 40.0, -2.0)),
                          np.asarray(oracle_rows(20, -0.3, 40.0, -2.0)))


def test_bond_scan_rule(pkg):
    from cdft_amd.batch import bond_scan_geometries

    scan = bond_scan_geometries(count=64, seed=0)
    bonds = [r for r, _, _ in scan]
    assert len(bonds) == 64 and 0.8 <= min(bonds) and max(bonds) <= 6.0
    for r, sc, box in scan:
        cells = box[0] / 0.25
        assert abs(cells - round(cells)) < 1e-9 and box[0] >= r + 14.0 and box[1:] == (14.0, 14.0)
        assert sc.positions[0][0] == -0.5 * r and sc.positions[1][0] == 0.5 * r


def test_lumped_axis_tables_match_the_oracle_matrices(pkg):
    """The device's per-axis (first, weights, ranges) tables rebuild the oracle's dense transposed
    interpolation matrix of ``quadrature.py:447-466``."""
    from cdft_amd.lumped import axis_table

    n, refine, h, origin = 21, 2, 0.4, -4.0
    fine, first, w, ranges = axis_table(origin, h, n, refine, 7)
    mat = np.zeros((n, n * refine))
    for f in range(n * refine):
        mat[first[f]:first[f] + 8, f] = w[f]
    # every coarse row's nonzeros are exactly its fine range
    for c in range(n):
        nz = np.nonzero(mat[c])[0]
        assert nz.min() >= ranges[c, 0] and nz.max() < ranges[c, 1]
    # degree-7 reproduction: sum_c M[c, f] p(x_c) = p(x_f) for cubic p
    xc = origin + h * np.arange(n)
    p = lambda x: 1.0 + 0.3 * x - 0.05 * x**2 + 0.01 * x**3  # noqa: E731
    assert np.allclose(mat.T @ p(xc), p(fine), rtol=0, atol=1e-12)


def _loop_node_table(gm, centres, radius, h, node_spacing, origin):
    """The node table built shell by shell with (N, 3) arrays, as a plain loop."""
    points, weights, owners = [], [], []
    for atom, centre in enumerate(centres):
        n_panels = max(2, int(np.ceil(radius / h)))
        edges = np.linspace(0.0, radius, n_panels + 1)
        gl_nodes, gl_weights = gm.gauss_legendre(6)
        for lo, hi in zip(edges[:-1], edges[1:]):
            mid, half = 0.5 * (hi + lo), 0.5 * (hi - lo)
            for node, weight in zip(gl_nodes, gl_weights):
                r = mid + half * node
                wr = half * weight * r * r
                n_theta = max(4, int(np.ceil(np.sqrt(2.0 * np.pi) * r / (node_spacing * h))))
                directions, angular = gm.sphere_directions(n_theta)
                points.append(r * directions + np.asarray(centre))
                weights.append(wr * angular)
                owners.append(np.full(directions.shape[0], atom, dtype=np.int32))
    points, weights, owners = np.concatenate(points), np.concatenate(weights), np.concatenate(owners)
    base = np.floor((points - origin) / h).astype(np.int32)
    lo = base.min(axis=0)
    ext = base.max(axis=0) - lo + 1
    rel = base - lo
    key = (rel[:, 0].astype(np.int64) * int(ext[1]) + rel[:, 1]) * int(ext[2]) + rel[:, 2]
    perm = np.argsort(key, kind="stable")
    return points[perm], owners[perm], weights[perm], base[perm], perm, lo, ext, np.bincount(key, minlength=int(ext.prod()))


def test_node_table_is_the_shell_loop(pkg):
    """The vectorised node table (coordinate rows, one radix sort) equals, bitwise, the table built
    shell by shell and sorted by a 64-bit stable sort, for one and two nuclei."""
    from cdft_amd import geometry as gm
    origin = np.array([-8.0, -8.0, -8.0])
    for centres, radius, h, ns in (([(0.0, 0.0, 0.0)], 2.875, 0.25, 1.0),
                                   ([(0.0, 0.0, -0.7), (0.0, 0.0, 0.7)], 2.0, 0.2, 0.7),
                                   ([(0.3, -0.2, 0.1)], 1.1, 0.3, 1.0)):
        ref = _loop_node_table(gm, centres, radius, h, ns, origin)
        t = gm.node_table(centres, [radius] * len(centres), h, 6, [ns] * len(centres), origin)
        got = (t["point_rows"].T, t["owners"], t["bare_weights"], t["base_rows"].T, t["perm"], t["lo"], t["ext"],
               t["counts"])
        for a, b in zip(ref, got):
            assert np.array_equal(a, b)
        points, weights = gm.sphere_nodes(centres[0], radius, h, 6, ns)
        assert points.shape == (weights.shape[0], 3)


def test_dropped_graphs_wait_for_a_capture_free_moment(pkg):
    """A dropped captured graph is not destroyed where the last reference goes (which may be while
    another host thread records a capture) but queued, and destroyed under the capture lock."""
    import threading

    from cdft_amd import ops

    destroyed = []

    class FakeGraph:
        def __init__(self, name):
            self.name = name

        def replay(self):
            pass

        def __del__(self):
            destroyed.append(self.name)

    ops.release_graphs()
    g = ops.CapturedGraph(FakeGraph("a"))
    g.replay()
    del g
    assert destroyed == [] and len(ops._released) == 1
    # while another thread holds the capture lock, release_graphs waits
    held, done = threading.Event(), threading.Event()

    def hold():
        with ops._graph_lock:
            held.set()
            done.wait(5)

    t = threading.Thread(target=hold)
    t.start()
    held.wait(5)
    waiter = threading.Thread(target=ops.release_graphs)
    waiter.start()
    waiter.join(0.2)
    assert waiter.is_alive() and destroyed == []
    done.set()
    t.join(5)
    waiter.join(5)
    assert destroyed == ["a"] and not ops._released
    # deferred device-side releases run at the same moment, in the lock
    calls = []
    ops.release_later(lambda: calls.append("plan"))
    assert calls == []
    ops.release_graphs()
    assert calls == ["plan"] and not ops._released
    # finalisers run during interpreter shutdown, when module globals are already None
    queue = ops._released
    try:
        ops._released = None
        ops.release_later(lambda: calls.append("late"))
    finally:
        ops._released = queue
    ops.release_graphs()
    assert calls == ["plan", "late"] and not ops._released
