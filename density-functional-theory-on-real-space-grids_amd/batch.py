"""Many geometries in one launch: batched stencil plans and a lockstep CheFSI for independent
non-interacting solves (the bond-length scan of ``scripts/scan_throughput.py:212-313``, which the
reference runs as separate solves in worker processes or streams).

Every grid of a batch gets a descriptor (its buffers, geometry, Chebyshev table, tile plan and
first block). One launch of the batched stencil kernel covers all grids: a block finds its grid by
a binary search over the first-block offsets, then runs exactly the per-block code of the
single-grid kernel. Each batched step is therefore bitwise the single-grid step. A CheFSI
iteration of the whole batch is one HIP graph:
- one batched degree-d filter;
- each grid's Gram, Cholesky and rotation;
- one batched apply;
- each grid's projection, Jacobi, rotation and orthonormality check;
- one batched apply and each grid's residual norms.
The host then reads all grids' eigenvalues and residuals in one transfer. It runs each grid's stop
rule (the reference's, ``eigen/chefsi.py:442-585``) and freezes the results of grids that have
stopped, while the batch keeps iterating until every grid has.
"""

from __future__ import annotations

import ctypes
import time

import numpy as np
import torch

from . import _lib, ops
from .eigen import BOUND_REUSE_RITZ_FRACTION, EigenConfig, RitzWork, _fraction  # noqa: F401
from .chefsi import scaled_chebyshev_table
from .hamiltonian import lanczos_finish, lanczos_launch, seeded_starts
from .ops import stream_of


class BatchPlan:
    """A descriptor table on the device for ``grids``: per grid ``(geometry, x, out, work1, work2,
    table)``. The plan keeps every tensor it points at alive."""

    def __init__(self, grids: list, k: int):
        n = len(grids)
        if n < 1:
            raise ValueError("a batch needs at least one grid")
        order = grids[0][0].order
        dtype = grids[0][0].dtype
        if dtype not in (torch.float64, torch.float32):
            raise ValueError("a batch runs in float64 or float32")
        device = grids[0][1].device
        shapes = (ctypes.c_int * (3 * n))()
        coefs = (ctypes.c_double * (order * n))()
        ptrs = (ctypes.c_void_p * (11 * n))()
        for i, (geom, x, out, w1, w2, table) in enumerate(grids):
            if geom.order != order or geom.dtype != dtype:
                raise ValueError("a batch needs one stencil order and one geometry dtype")
            for t, name in ((x, "x"), (out, "out"), (w1, "work1"), (w2, "work2")):
                ops.require(t, (k, geom.n_points), dtype, device, name)
            ops.require(table, (table.shape[0], 3), torch.float64, device, "table")
            shapes[3 * i: 3 * i + 3] = list(geom.shape)
            for j in range(order):
                coefs[order * i + j] = geom.taps[j]
            vals = [x, out, w1, w2, geom.w0, geom.w1, geom.w2, geom.invw, geom.potential, geom.packed, table]
            for j, t in enumerate(vals):
                ptrs[11 * i + j] = t.data_ptr()
        nbytes = int(_lib.load().cdft_batch_desc_bytes(n))
        host = (ctypes.c_uint8 * nbytes)()
        launch = (ctypes.c_int * 3)()
        self.suffix = "f64" if dtype == torch.float64 else "f32"
        _lib.call(f"cdft_batch_plan_{self.suffix}", n, ctypes.addressof(shapes), k, order, ctypes.addressof(coefs),
                  ctypes.addressof(ptrs), ctypes.addressof(host), nbytes, ctypes.addressof(launch))
        self.desc = torch.frombuffer(bytearray(host), dtype=torch.uint8).to(device)
        self.launch = launch
        self.n = n
        self.order = order
        self.k = k
        self.blocks = int(launch[0])
        self._keep = grids

    def filter(self, degree: int, stream_tensor: torch.Tensor) -> None:
        _lib.call(f"cdft_cheb_filter_batched_{self.suffix}", self.desc.data_ptr(), self.n,
                  ctypes.addressof(self.launch), self.order, int(degree), stream_of(stream_tensor))

    def apply(self, stream_tensor: torch.Tensor) -> None:
        _lib.call(f"cdft_hdiv_apply_batched_{self.suffix}", self.desc.data_ptr(), self.n,
                  ctypes.addressof(self.launch), self.order, stream_of(stream_tensor))


class _Instance:
    """Host state of one grid's CheFSI (mirrors the locals of ``ChebyshevFilteredSubspace.solve``)."""

    def __init__(self, hamiltonian, work: RitzWork, n_states: int):
        self.h = hamiltonian
        self.work = work
        self.n_states = n_states
        self.iteration = 0
        self.done = False
        self.stop_reason = "iteration limit"
        self.stable = 0
        self.result = None


class BatchedCheFSI:
    """Lockstep CheFSI for the lowest states of many independent operators (one per grid)."""

    #: Mixed precision (as the single-grid solver): float32 filters while every active grid's worst
    #: residual is above this and no float32 pass has stalled, float64 after that.
    switch_residual: float = 1.0e-5
    stall_ratio: float = 0.7
    fp64_tail: int = 1

    def __init__(self, hamiltonians: list, config: EigenConfig, n_states: int = 1, ritz_stop: bool = True,
                 seed: int = 0, use_graphs: bool = True, autotune: bool = True, filter_dtype=torch.float64):
        self.h = hamiltonians
        self.filter_dtype = filter_dtype
        self.mixed = filter_dtype != torch.float64
        self.filter_precisions: list[str] = []
        self.config = config
        self.n_states = n_states
        self.ritz_stop = ritz_stop
        self.seed = seed
        self.use_graphs = use_graphs
        self.graph_record = {"captures": 0, "replays": 0}
        self.timings: dict[str, float] = {}
        self.autotune = autotune
        self.tuning: dict | None = None

    def _iteration_body(self, insts, plan_filter, plan_rr, plan_res, degree, use32=False):
        cfg = self.config
        if use32:  # float32 filter of the float64 Ritz vectors, back to float64 for Rayleigh-Ritz
            for it in insts:
                it.work.x32.copy_(it.work.vectors)
            plan_filter.filter(degree, insts[0].work.vectors)
            for it in insts:
                it.work.filtered.copy_(it.work.y32)
        else:
            plan_filter.filter(degree, insts[0].work.vectors)
        for it in insts:  # orthonormalise and rotate into the basis
            w, m = it.work, it.h.measure
            m.gram(w.filtered, out=w.S)
            ops.chol_inv(w.S, w.T, w.flag)
            ops.rotate(w.T, w.filtered, out=w.basis)
        plan_rr.apply(insts[0].work.basis)  # applied = A basis, all grids
        for it in insts:
            w, m = it.work, it.h.measure
            m.cross(w.basis, w.applied, out=w.M)
            ops.eigh(w.M, w.evals, w.C)
            ops.rotate(w.C, w.basis, transpose=True, out=w.vectors)
            m.gram(w.vectors, out=w.G)
            ops.orth_error(w.G, w.scalars[2 * w.k: 2 * w.k + 1])
        plan_res.apply(insts[0].work.vectors)  # applied = A vectors, all grids
        for it in insts:
            w = it.work
            ops.resnorm(w.applied, w.vectors, w.evals, it.h.measure.q, out=w.norms)
        del cfg

    def solve(self) -> list[dict]:
        cfg = self.config
        t0 = time.perf_counter()
        extra = max(cfg.n_extra_states, int(cfg.extra_states_fraction * self.n_states))
        k = self.n_states + extra
        degree = cfg.chebyshev_degree
        n = len(self.h)
        dev = self.h[0].grid.device
        scal = torch.zeros((n, 2 * k + 1), dtype=torch.float64, device=dev)
        flags = torch.zeros(n, dtype=torch.int32, device=dev)
        tables = torch.zeros((n, degree, 3), dtype=torch.float64, device=dev)
        tables_host = torch.zeros((n, degree, 3), dtype=torch.float64).pin_memory()
        insts = []
        for i, h in enumerate(self.h):
            work = RitzWork(k, h.grid.n_points, degree, dev, mixed=self.mixed)
            work.scalars = scal[i]
            work.flag = flags[i: i + 1]
            work.table = tables[i]
            insts.append(_Instance(h, work, self.n_states))
        # Start: the reference's seeded draws per grid (the (k, n) block, then the Lanczos start, from
        # a fresh CPU generator each, precision.py:246-263; the blocks' common prefix drawn once,
        # hamiltonian.seeded_starts); every grid's Lanczos run enqueued without a read, then one
        # transfer for all of them; the first Rayleigh-Ritz per grid.
        ta = time.perf_counter()
        sizes = [it.h.grid.n_points for it in insts]
        draws = seeded_starts(self.seed, k, sizes, dev)
        t_draw = time.perf_counter() - ta
        tb = time.perf_counter()
        lanczos = []
        for it, (block, start) in zip(insts, draws):
            it.block = block
            lanczos.append(lanczos_launch(it.h, start, cfg.lanczos_steps))
        values = torch.stack(lanczos).cpu().numpy()
        for it, vals in zip(insts, values):
            it.smin, it.smax = lanczos_finish(vals, cfg.lanczos_steps, cfg.bound_safety)
        t_lanczos = time.perf_counter() - tb
        for it in insts:
            block = it.block
            w = it.work
            it.h.measure.gram(block, out=w.S)
            ops.chol_inv(w.S, w.T, w.flag)
            ops.rotate(w.T, block, out=w.basis)
            ops.hdiv_apply(w.basis, it.h.geometry, w.applied)
            it.h.measure.cross(w.basis, w.applied, out=w.M)
            ops.eigh(w.M, w.evals, w.C)
            ops.rotate(w.C, w.basis, transpose=True, out=w.vectors)
            w.G.zero_()
            ops.hdiv_apply(w.vectors, it.h.geometry, w.applied)
            ops.resnorm(w.applied, w.vectors, w.evals, it.h.measure.q, out=w.norms)
        host = scal.cpu().numpy()
        for i, it in enumerate(insts):
            it.evals = host[i, :k].copy()
            it.worst = float(host[i, k:k + self.n_states].max())
            it.previous = it.evals.copy()
            it.next_lower = float(it.evals[k - 1])
        self.timings["start_s"] = time.perf_counter() - t0
        self.timings["start_draws_s"] = t_draw
        self.timings["start_lanczos_s"] = t_lanczos
        geoms = [(it.h.geometry, it.work) for it in insts]
        filter_rows = [(g, w.vectors, w.filtered, w.work1, w.work2, w.table) for g, w in geoms]
        from . import autotune

        tuning = None
        if self.autotune:
            # The tables are still zero here, so the timed filters only overwrite scratch blocks.
            signature = ("batched-filter", tuple(it.h.grid.shape for it in insts), k, degree)
            self.tuning = tuning = autotune.tune(signature,
                                                 lambda: BatchPlan(filter_rows, k).filter(degree, insts[0].work.vectors))
        autotune.apply(tuning)
        try:
            plan_filter = BatchPlan(filter_rows, k)
        finally:
            autotune.apply(None)
        plan_rr = BatchPlan([(g, w.basis, w.applied, w.work1, w.work2, w.table) for g, w in geoms], k)
        plan_res = BatchPlan([(g, w.vectors, w.applied, w.work1, w.work2, w.table) for g, w in geoms], k)
        plan32 = None
        if self.mixed:
            autotune.apply(tuning)
            try:
                plan32 = BatchPlan([(it.h.geometry_in(self.filter_dtype), it.work.x32, it.work.y32, it.work.a32,
                                     it.work.b32, it.work.table) for it in insts], k)
            finally:
                autotune.apply(None)
        self.blocks_per_launch = plan_filter.blocks
        graphs: dict = {}
        stalled32 = False
        worst_before = max(it.worst for it in insts)
        self.timings["plans_s"] = time.perf_counter() - t0 - self.timings["start_s"]
        weak_form = self.ritz_stop
        eigenvalue_tol = 0.1 * cfg.cross_check_tol
        t1 = time.perf_counter()
        for iteration in range(1, cfg.max_iterations + 1):
            for i, it in enumerate(insts):
                if it.done:
                    continue
                lower = min(max(it.next_lower, it.smin + 1e-8), it.smax - 1e-8)
                tables_host[i] = torch.as_tensor(scaled_chebyshev_table(degree, lower, it.smax, it.smin))
            tables.copy_(tables_host, non_blocking=True)
            active_worst = max(it.worst for it in insts if not it.done)
            use32 = (self.mixed and not stalled32 and active_worst > self.switch_residual
                     and iteration <= cfg.max_iterations - self.fp64_tail)
            self.filter_precisions.append("f32" if use32 else "f64")
            body_filter = plan32 if use32 else plan_filter
            if self.use_graphs:
                graph = graphs.get(use32)
                if graph is None:
                    graph = ops.capture_graph(
                        lambda: self._iteration_body(insts, body_filter, plan_rr, plan_res, degree, use32))
                    graphs[use32] = graph
                    self.graph_record["captures"] += 1
                graph.replay()
                self.graph_record["replays"] += 1
            else:
                self._iteration_body(insts, body_filter, plan_rr, plan_res, degree, use32)
            fl = flags.to("cpu", non_blocking=True)
            host = scal.cpu().numpy()
            if int(fl.max()) != 0:
                raise RuntimeError("Rayleigh-Ritz: a filtered block's Gram matrix is not positive definite")
            if use32:
                # a float32 pass that no longer shrinks the worst active residual has reached the
                # float32 floor: the rest of the solve filters in float64
                now = max(float(host[i, k:k + self.n_states].max()) for i, it in enumerate(insts) if not it.done)
                if now > self.stall_ratio * worst_before:
                    stalled32 = True
                worst_before = now
            for i, it in enumerate(insts):
                if it.done:
                    continue
                it.iteration = iteration
                evals = host[i, :k].copy()
                norms = host[i, k:2 * k]
                it.worst = float(norms[: self.n_states].max())
                drift = float(np.abs(evals[: self.n_states] - it.previous[: self.n_states]).max())
                it.next_lower = float(evals[k - 1])
                it.previous = evals
                stop = None
                if it.worst < cfg.residual_tol and not use32:
                    stop = "residual below residual_tol"
                else:
                    # the weak-form stop counts float64 passes only (float32 Ritz values jitter)
                    it.stable = it.stable + 1 if (weak_form and drift < eigenvalue_tol and not use32) else 0
                    if it.stable >= 2:
                        stop = "Ritz values stable on the weak-form path"
                if stop is not None:
                    it.done = True
                    it.stop_reason = stop
                    it.result = {"eigenvalues": evals[: self.n_states], "residuals": norms[: self.n_states].copy(),
                                 "eigenvectors": it.work.vectors[: self.n_states].clone(), "n_iterations": iteration,
                                 "converged": True, "stop_reason": stop}
            if all(it.done for it in insts):
                break
        torch.cuda.synchronize()
        self.timings["iterations_s"] = time.perf_counter() - t1
        self.iterations = iteration
        out = []
        for it in insts:
            if it.result is None:
                it.result = {"eigenvalues": it.previous[: self.n_states], "residuals": None,
                             "eigenvectors": it.work.vectors[: self.n_states].clone(), "n_iterations": it.iteration,
                             "converged": False, "stop_reason": it.stop_reason}
            out.append(it.result)
        return out


def bond_scan_geometries(count: int = 64, seed: int = 0, spacing: float = 0.25, low: float = 0.8, high: float = 6.0,
                         across: float = 14.0, vacuum: float = 14.0):
    """The seeded H2+ bond-length scan: ``R = low + (high - low) U(0, 1)`` (numpy generator ``seed``),
    protons at ``(-R/2, 0, 0)`` and ``(R/2, 0, 0)``, box ``(R + vacuum, across, across)`` with the
    bond edge rounded up to whole cells so every grid keeps the spacing exactly."""
    from types import SimpleNamespace

    rng = np.random.default_rng(seed)
    bonds = low + (high - low) * rng.uniform(0.0, 1.0, count)
    out = []
    for r in bonds:
        edge = float(np.ceil((r + vacuum) / spacing - 1e-9)) * spacing
        sc = SimpleNamespace(charges=(1.0, 1.0), positions=((-0.5 * r, 0.0, 0.0), (0.5 * r, 0.0, 0.0)),
                             n_electrons=1.0, xc="none")
        out.append((float(r), sc, (edge, across, across)))
    return out
