"""Rayleigh-Ritz and Chebyshev-filtered subspace iteration on the GPU.

Mirrors ``src/cdft/eigen/rayleigh_ritz.py:51-115`` and ``src/cdft/eigen/chefsi.py:442-585``:
Cholesky orthonormalisation in the operator's measure, the projected k x k problem, rotation,
residual norms, and the outer loop with one host read per iteration. Every block-sized operation is
a HIP kernel (K1-K4); the k x k algebra runs in one-thread kernels on the device.

One CheFSI iteration -- the degree-d filter, Rayleigh-Ritz and the residual norms, about 35 kernel
launches -- is captured once into a HIP graph over static buffers (the block, the recurrence table,
the diagonal potential) and replayed every iteration, across SCF iterations too; the host only
uploads the next recurrence table and reads the iteration's scalars. This plays the role of the
reference's ``GraphedChebyshevFilter`` (``chefsi.py:185-392``), extended to the whole iteration.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch

from . import autotune, ops
from .hamiltonian import DeviceMeasure, lanczos_enqueue, lanczos_finish, seeded_randn

BOUND_REUSE_RITZ_FRACTION = 0.9


@dataclass
class EigenConfig:
    """Eigensolver settings (reference ``contract.py:430-462``; ``ALL_ELECTRON_NUMERICS`` values)."""

    n_extra_states: int = 4
    extra_states_fraction: float = 0.20
    chebyshev_degree: int = 20
    lanczos_steps: int = 12
    bound_safety: float = 1.02
    max_iterations: int = 300
    residual_tol: float = 1.0e-8
    cross_check: bool = False
    cross_check_tol: float = 1.0e-8


@dataclass
class EigenResult:
    eigenvalues: np.ndarray  # host copy, ascending
    eigenvectors: torch.Tensor  # device (n_states, N)
    residuals: np.ndarray
    n_iterations: int
    converged: bool
    info: dict = field(default_factory=dict)
    eigenvalues_device: torch.Tensor | None = None


@dataclass(frozen=True)
class SpectralBoundsHint:
    """Carried bounds and a Weyl shift (reference ``eigen/chefsi.py:169-182``)."""

    lower: float
    upper: float
    shift: float = 0.0  # or a DeviceScalar, read when the solve first needs it


class DeviceArray:
    """A device array copied to pinned host memory behind an event as soon as it is made;
    :meth:`numpy` waits only for that copy, not for work enqueued after it."""

    def __init__(self, value: torch.Tensor):
        self._host = torch.empty(value.numel(), dtype=value.dtype).pin_memory()
        self._host.copy_(value.reshape(-1), non_blocking=True)
        self._ready = torch.cuda.Event()
        self._ready.record()

    def numpy(self) -> np.ndarray:
        self._ready.synchronize()
        return self._host.numpy()


class DeviceScalar:
    """A device scalar copied to pinned host memory behind an event as soon as it is made; ``float()``
    waits only for that copy. Lets a caller hand the solver a value still being computed without a
    synchronising read of its own (the solver reads it after its first Rayleigh-Ritz, by which time
    the copy has long finished)."""

    def __init__(self, value: torch.Tensor):
        self._host = torch.empty(1, dtype=torch.float64).pin_memory()
        self._host.copy_(value.reshape(1), non_blocking=True)
        self._ready = torch.cuda.Event()
        self._ready.record()

    def __float__(self) -> float:
        self._ready.synchronize()
        return float(self._host[0])


class RitzWork:
    """Device buffers of one CheFSI iteration for a block of k vectors (allocated once)."""

    def __init__(self, k: int, n: int, degree: int, device, mixed: bool):
        f = dict(dtype=torch.float64, device=device)
        self.k = k
        self.S = torch.empty((k, k), **f)
        self.T = torch.empty((k, k), **f)
        self.M = torch.empty((k, k), **f)
        self.C = torch.empty((k, k), **f)
        self.G = torch.empty((k, k), **f)
        self.flag = torch.zeros(1, dtype=torch.int32, device=device)
        # scalars read once per iteration: evals (k) | residual norms (k) | orth error (1)
        self.scalars = torch.zeros(2 * k + 1, **f)
        self.basis = torch.empty((k, n), **f)
        self.applied = torch.empty((k, n), **f)
        self.vectors = torch.empty((k, n), **f)
        self.filtered = torch.empty((k, n), **f)
        self.work1 = torch.empty((k, n), **f)
        self.work2 = torch.empty((k, n), **f)
        self.table = torch.zeros((degree, 3), **f)
        # [spectrum_min, upper] of the filter interval: the device builds the table from these and the
        # last Ritz value, so only a change of bounds crosses the bus
        self.bounds = torch.zeros(2, **f)
        self.bounds_host = torch.zeros(2, dtype=torch.float64).pin_memory()
        self.bounds_set: tuple | None = None
        # the per-pass scalars land in pinned host memory behind an event, so the host can wait for
        # them while a speculative filter runs
        self.scalars_host = torch.zeros(2 * k + 1, dtype=torch.float64).pin_memory()
        self.scalars_ready = torch.cuda.Event() if torch.cuda.is_available() else None
        if mixed:
            g = dict(dtype=torch.float32, device=device)
            self.x32 = torch.empty((k, n), **g)
            self.y32 = torch.empty((k, n), **g)
            self.a32 = torch.empty((k, n), **g)
            self.b32 = torch.empty((k, n), **g)

    @property
    def evals(self) -> torch.Tensor:
        return self.scalars[: self.k]

    @property
    def norms(self) -> torch.Tensor:
        return self.scalars[self.k: 2 * self.k]


def rayleigh_ritz(geometry, measure, block: torch.Tensor, work: RitzWork, orth: bool = True) -> None:
    """Orthonormalise ``block`` (Cholesky, ``T = L^-1``), project, diagonalise, rotate: eigenvalues
    in ``work.evals``, Ritz vectors in ``work.vectors``, ``max|G - I|`` in ``work.scalars[2k]``; a
    non-positive-definite Gram matrix sets ``work.flag``. ``orth=False`` leaves ``max|G - I|`` to the
    caller (the residual-norm launch can compute it from ``work.G``)."""
    k = work.k
    measure.gram(block, out=work.S)
    ops.chol_inv(work.S, work.T, work.flag)
    ops.rotate(work.T, block, out=work.basis)
    ops.hdiv_apply(work.basis, geometry, work.applied)
    measure.cross(work.basis, work.applied, out=work.M)
    ops.eigh(work.M, work.evals, work.C)
    if k <= 8:
        ops.rotate_gram(work.C, work.basis, work.vectors, measure.q, work.G)
    else:
        ops.rotate(work.C, work.basis, transpose=True, out=work.vectors)
        measure.gram(work.vectors, out=work.G)
    if orth:
        ops.orth_error(work.G, work.scalars[2 * k: 2 * k + 1])


def residual_norms(geometry, measure, work: RitzWork) -> None:
    """``||A phi_i - eps_i phi_i||`` in the measure into ``work.norms`` (a fresh apply, as the
    reference's ``residual_norms``; the apply keeps only the norm partials, nothing is stored)."""
    ops.hdiv_resnorm(work.vectors, geometry, work.evals, measure.q, work.norms)


def _read(work: RitzWork) -> tuple[np.ndarray, np.ndarray, float]:
    """The pass's scalars (eigenvalues, residual norms, orthonormality error), read now."""
    _read_async(work)
    return _collect(work)


def _read_async(work: RitzWork) -> None:
    """Start the pass's scalar read: a copy into pinned host memory and an event behind it."""
    work.scalars_host.copy_(work.scalars, non_blocking=True)
    work.scalars_ready.record()


def _collect(work: RitzWork) -> tuple[np.ndarray, np.ndarray, float]:
    """Wait for the read started by :func:`_read_async` (not for later work on the stream)."""
    work.scalars_ready.synchronize()
    host = work.scalars_host.numpy()
    k = work.k
    orth = float(host[2 * k])
    if not orth < 0.5 and int(work.flag.cpu()[0]) != 0:
        raise RuntimeError(
            "Rayleigh-Ritz: the filtered block's Gram matrix is not positive definite (the block has "
            "collapsed in rank); the eigh fallback of the reference that drops vectors is not implemented")
    return host[:k].copy(), host[k:2 * k].copy(), orth


def _fraction(value, low, high):
    width = high - low
    return (value - low) / width if width > 0.0 else float("inf")


#: Speculate the next filter only while the last residual is above this many times the tolerance (a
#: pass cuts the residual by far less than this at the filter degrees used).
SPECULATE_RESIDUAL = 1.0e3


class ChebyshevFilteredSubspace:
    """CheFSI (reference ``eigen/chefsi.py:408-585``) on the device, one HIP graph per iteration."""

    #: Mixed precision: filter in float32 while the worst residual is above this...
    switch_residual: float = 1.0e-5
    #: ...and always finish an iteration-limited solve with this many float64 filter passes.
    fp64_tail: int = 1
    #: A float32 pass whose residual is above this fraction of the previous one has stalled.
    stall_ratio: float = 0.7
    #: Enqueue the next pass's filter before reading a pass's scalars when that pass cannot stop
    #: the solve (see ``SPECULATE_RESIDUAL``); results are the same either way.
    speculate: bool = True

    def __init__(self, grid, config: EigenConfig, ritz_stop: bool = True, filter_dtype=torch.float64,
                 filter_hook=None, use_graphs: bool = True):
        self.grid = grid
        self.config = config
        self.ritz_stop = ritz_stop
        self.filter_dtype = filter_dtype
        self.filter_hook = filter_hook
        self.use_graphs = use_graphs
        self.graph_record: dict = {"captures": 0, "replays": 0, "disabled": None}
        self.history: list[dict] = []
        self.filter_precisions: list[str] = []
        self.speculation = {"used": 0, "redone": 0, "wasted": 0}
        self._work: dict[int, RitzWork] = {}
        self._graphs: dict[tuple, ops.CapturedGraph] = {}
        n = grid.n_points
        self._pot = torch.empty(n, dtype=torch.float64, device=grid.device)
        self._pot32 = torch.empty(n, dtype=torch.float32, device=grid.device) if filter_dtype != torch.float64 else None
        self._sources = None
        self._layout = None
        self._geom64 = self._geom32 = None
        self._measure = None

    @property
    def mixed(self) -> bool:
        return self.filter_dtype != torch.float64

    def work_for(self, k: int) -> RitzWork:
        if k not in self._work:
            self._work[k] = RitzWork(k, self.grid.n_points, self.config.chebyshev_degree, self.grid.device, self.mixed)
        return self._work[k]

    def _bind(self, hamiltonian) -> None:
        """Copy this operator into the static geometry the captured graphs read: its diagonal
        potential on every call; its face weights, records, ``invw`` and ``q`` when they are other
        arrays than the last bound ones (a new solve). The graphs are recaptured only when the
        shape, order, taps or device change: the taps are kernel arguments, captured by value.
        The last bound arrays stay referenced, so their addresses cannot be handed to new arrays."""
        geom = hamiltonian.geometry
        q = hamiltonian.measure.q
        if isinstance(geom, ops.LocalGeometry):
            # the plain operator of the model systems: its weights are kernel arguments (captured by
            # value), its potential and measure device arrays; float64 filters only
            if self.mixed:
                raise ValueError("the plain model-system operator filters in float64 only")
            layout = ("local", geom.shape, geom.order, geom.boundary, tuple(geom.coef), geom.device)
            if layout != self._layout:
                self._layout = layout
                self._graphs.clear()
                self._geom64 = geom.with_potential(self._pot)
                self._measure = DeviceMeasure(self.grid, q.clone())
            elif self._sources is None or self._sources[0] is not q:
                self._measure.q.copy_(q)
            self._sources = (q,)
            self._pot.copy_(hamiltonian.potential)
            return
        sources = (geom.w0, geom.w1, geom.w2, geom.invw, geom.packed, q)
        if self._sources is None or any(a is not b for a, b in zip(sources, self._sources)):
            layout = (geom.shape, geom.order, tuple(geom.taps), geom.device, geom.dtype)
            if layout != self._layout:
                self._layout = layout
                self._graphs.clear()
                clone = lambda t: t.clone()  # noqa: E731
                self._geom64 = ops.StencilGeometry(geom.shape, geom.order, clone(geom.w0), clone(geom.w1),
                                                   clone(geom.w2), clone(geom.invw), self._pot, geom.taps,
                                                   clone(geom.packed))
                self._measure = DeviceMeasure(self.grid, q.clone())
                if self.mixed:
                    g, cast = self._geom64, lambda t: t.to(self.filter_dtype)  # noqa: E731
                    # the records hold the arrays' values and zeros only, so their cast is the records
                    # of the cast arrays
                    self._geom32 = ops.StencilGeometry(g.shape, g.order, cast(g.w0), cast(g.w1), cast(g.w2),
                                                       cast(g.invw), self._pot32, g.taps, cast(g.packed))
            else:
                targets = (self._geom64.w0, self._geom64.w1, self._geom64.w2, self._geom64.invw,
                           self._geom64.packed, self._measure.q)
                for dst, src in zip(targets, sources):
                    dst.copy_(src)
                if self.mixed:
                    g, g32 = self._geom64, self._geom32
                    for dst, src in ((g32.w0, g.w0), (g32.w1, g.w1), (g32.w2, g.w2), (g32.invw, g.invw),
                                     (g32.packed, g.packed)):
                        dst.copy_(src)
            self._sources = sources
        self._pot.copy_(hamiltonian.potential)
        if self.mixed:
            self._pot32.copy_(self._pot)

    def begin_solve(self, grid, config: EigenConfig, filter_hook=None) -> None:
        """Start a new solve with this solver (its buffers, static geometry and captured graphs are
        kept): clear the per-solve records and take the solve's grid, configuration and hook."""
        self.grid = grid
        self.config = config
        self.filter_hook = filter_hook
        self.graph_record = {"captures": 0, "replays": 0, "disabled": self.graph_record["disabled"]}
        self.history = []
        self.filter_precisions = []
        self.speculation = {"used": 0, "redone": 0, "wasted": 0}

    # ---- one iteration ----

    def _filter_into(self, work: RitzWork, use32: bool) -> None:
        degree = self.config.chebyshev_degree
        if not use32:
            ops.cheb_filter(work.vectors, self._geom64, work.table, degree, out=work.filtered,
                            work=(work.work1, work.work2))
        else:
            work.x32.copy_(work.vectors)
            ops.cheb_filter(work.x32, self._geom32, work.table, degree, out=work.y32, work=(work.a32, work.b32))
            work.filtered.copy_(work.y32)

    def _ritz(self, work: RitzWork) -> None:
        rayleigh_ritz(self._geom64, self._measure, work.filtered, work, orth=False)
        # the residual norms, and in the same combine launch the next filter's table (from this
        # pass's last Ritz value and the current bounds) and the orthonormality error
        k = work.k
        ops.hdiv_resnorm(work.vectors, self._geom64, work.evals, self._measure.q, work.norms,
                         tail=(k - 1, work.bounds, self.config.chebyshev_degree, work.table, work.G,
                               work.scalars[2 * k: 2 * k + 1]))

    def _replay(self, key: tuple, body) -> None:
        """Run ``body`` through its captured graph (captured on first use), or eagerly when capture
        is off or failed (the reason is recorded)."""
        if self.use_graphs and self.graph_record["disabled"] is None:
            graph = self._graphs.get(key)
            if graph is None:
                try:
                    graph = ops.capture_graph(body)
                    self._graphs[key] = graph
                    self.graph_record["captures"] += 1
                except Exception as exc:  # noqa: BLE001 - recorded; the iteration then runs eagerly
                    self.graph_record["disabled"] = f"capture failed: {type(exc).__name__}: {exc}"
                    graph = None
            if graph is not None:
                graph.replay()
                self.graph_record["replays"] += 1
                return
        body()

    def _filter(self, work: RitzWork, use32: bool, degree: int) -> None:
        """The filter graph, with the timing hook's events (if any) around it on the same stream."""
        if self.filter_hook is not None:
            self.filter_hook("start", self, work.x32 if use32 else work.vectors, degree)
        self._replay(("filter", work.k, use32), lambda: self._filter_into(work, use32))
        if self.filter_hook is not None:
            self.filter_hook("end", self, work.x32 if use32 else work.vectors, degree)

    def _run_iteration(self, work: RitzWork, use32: bool, degree: int, filtered: bool = False) -> None:
        """Filter (unless a speculative filter already did, with these settings), then Rayleigh-Ritz
        and residuals: two captured graphs."""
        if not filtered:
            self._filter(work, use32, degree)
        self._replay(("ritz", work.k), lambda: self._ritz(work))

    def _lanczos(self, hamiltonian, generator) -> tuple[float, float]:
        """The operator's Lanczos enclosure (``hamiltonian.spectral_bounds``: the same seeded start,
        recurrence and widening) run through a captured graph on the solver's static geometry,
        which holds this operator's potential; one read of the recurrence scalars."""
        return self._lanczos_end(self._lanczos_begin(hamiltonian, generator))

    def _lanczos_begin(self, hamiltonian, generator):
        """Enqueue the Lanczos run of :meth:`_lanczos` and the copy of its scalars to pinned host
        memory; :meth:`_lanczos_end` waits for that copy and finishes on the host."""
        cfg = self.config
        steps = cfg.lanczos_steps
        start = seeded_randn(self.grid.n_points, generator, self.grid.device)
        buffers = self.__dict__.setdefault("_lanczos_buffers", {})
        if steps not in buffers:
            f = dict(dtype=torch.float64, device=self.grid.device)
            buffers[steps] = (torch.empty(self.grid.n_points, **f), torch.zeros(2 * steps + 3, **f),
                              torch.zeros(2 * steps + 3, dtype=torch.float64).pin_memory(), torch.cuda.Event())
        start_s, scal, host, ready = buffers[steps]
        start_s.copy_(start)
        geom, measure = self._geom64, self._measure
        self._replay(("lanczos", steps),
                     lambda: lanczos_enqueue(lambda x, y: ops.hdiv_apply(x, geom, y), measure, start_s, steps, scal))
        hamiltonian.note_applications(steps)
        host.copy_(scal, non_blocking=True)
        ready.record()
        return steps, host, ready

    def _lanczos_end(self, pending) -> tuple[float, float]:
        steps, host, ready = pending
        ready.synchronize()
        return lanczos_finish(host.numpy().copy(), steps, self.config.bound_safety)

    def _set_bounds(self, work: RitzWork, smin: float, smax: float) -> bool:
        """Point the filter at ``[smin, smax]``. The table itself is built on the device from the
        last Ritz value (at the end of every Rayleigh-Ritz graph); a change of bounds uploads them
        and rebuilds the table at once. Returns whether the bounds changed."""
        if work.bounds_set != (smin, smax):
            work.bounds_set = (smin, smax)
            work.bounds_host[0], work.bounds_host[1] = smin, smax
            work.bounds.copy_(work.bounds_host, non_blocking=True)
            ops.cheb_table(work.evals, work.k - 1, work.bounds, self.config.chebyshev_degree, work.table)
            return True
        return False

    # ---- the solver ----

    def solve(self, hamiltonian, n_states: int, initial: torch.Tensor | None = None, generator=None,
              bounds: SpectralBoundsHint | None = None, precision: str = "auto",
              min_fp64_passes: int = 0, max_iterations: int | None = None) -> EigenResult:
        """``precision`` (mixed solvers only): ``"auto"`` filters in float32 while the worst residual
        is above ``switch_residual``; ``"f32"`` filters every pass in float32 (the warm passes of an
        SCF iteration, whose vectors the outer loop refines); ``"f64"`` filters in float64 and
        runs at least ``min_fp64_passes`` passes (the refinement of a mixed solve).
        ``max_iterations`` caps this solve's passes in place of ``config.max_iterations`` (the SCF's
        warm iterations); the configuration, which concurrent solves may share, is only read."""
        cfg = self.config
        limit = cfg.max_iterations if max_iterations is None else int(max_iterations)
        g = self.grid
        extra = max(cfg.n_extra_states, int(cfg.extra_states_fraction * n_states))
        size = min(n_states + extra, g.n_points)
        if initial is None:
            block = seeded_randn((size, g.n_points), generator, g.device)
        else:
            block = initial.reshape(-1, g.n_points).to(torch.float64)
            if block.shape[0] < size:
                pad = seeded_randn((size - block.shape[0], g.n_points), generator, g.device)
                block = torch.cat([block, pad], dim=0)
        block = block.contiguous()
        work = self.work_for(size)
        self._bind(hamiltonian)
        # launch settings of the stencil kernels for this shape: timed once per process, on one of
        # its first solves (autotune.TUNE_ON_SOLVE), before that solve captures any graph; no result
        # depends on them
        retuned = autotune.tune_for_solve(self._geom64, size)
        if self.mixed:
            retuned |= autotune.tune_for_solve(self._geom32, size)
        if retuned:
            self._graphs.clear()  # captured with the earlier launches

        def lanczos():
            return self._lanczos(hamiltonian, generator)

        source, shift_used = "unset", 0.0
        pending = self._lanczos_begin(hamiltonian, generator) if bounds is None else None
        # the first Rayleigh-Ritz is the iteration's: the start block goes where a filter leaves its
        # output, and the captured Rayleigh-Ritz graph runs on it (behind the Lanczos run, whose
        # scalars are read with its own)
        work.filtered.copy_(block)
        self._replay(("ritz", work.k), lambda: self._ritz(work))
        hamiltonian.note_applications(2)
        work.bounds_set = None  # new Ritz values: the first pass rebuilds its table
        evals, norms, _ = _read(work)
        if pending is not None:
            spectrum_min, spectrum_max = self._lanczos_end(pending)
            source = "lanczos"
        worst = float(norms[:n_states].max())
        stalled32 = False
        previous = evals.copy()
        stable = 0
        stop_reason = "iteration limit"
        weak_form = self.ritz_stop and getattr(hamiltonian, "uses_divergence_form", False)
        eigenvalue_tol = 0.1 * cfg.cross_check_tol
        next_lower = float(evals[min(size, evals.shape[0]) - 1])
        if bounds is not None:
            shift_used = float(bounds.shift)
            spectrum_min, spectrum_max = bounds.lower - shift_used, bounds.upper + shift_used
            source = "reused"
        top_fraction = float("-inf")
        converged = False
        drift = float("inf")
        iteration = 0
        speculative = None  # precision of a speculative filter of the coming pass, if one is enqueued
        for iteration in range(1, limit + 1):
            lower = next_lower
            fraction = _fraction(lower, spectrum_min, spectrum_max)
            if source == "reused" and not fraction < BOUND_REUSE_RITZ_FRACTION:
                spectrum_min, spectrum_max = lanczos()
                source = f"lanczos (reuse rejected at iteration {iteration})"
                fraction = _fraction(lower, spectrum_min, spectrum_max)
            top_fraction = max(top_fraction, fraction)
            if precision == "f32":
                use32 = self.mixed
            elif precision == "f64":
                use32 = False
            else:
                use32 = (self.mixed and not stalled32 and worst > self.switch_residual
                         and iteration <= limit - self.fp64_tail)
            # the device-built table takes the last Ritz value (clamped into the bounds) as its lower
            # edge; the host copy of it above decides only the bound reuse
            degree = cfg.chebyshev_degree
            changed = self._set_bounds(work, spectrum_min, spectrum_max)
            self.filter_precisions.append("f32" if use32 else "f64")
            # A speculative filter of this pass (enqueued before the previous pass's read) stands
            # unless the bounds or the precision have changed since; it then runs again.
            ready = speculative == use32 and not changed
            if speculative is not None:
                self.speculation["used" if ready else "redone"] += 1
            self._run_iteration(work, use32, degree, filtered=ready)
            _read_async(work)
            # Speculate the next pass's filter while the host waits for this pass's scalars, when
            # this pass cannot be the last: not at the iteration limit, the residual stop far away
            # (the last residual above SPECULATE_RESIDUAL x tol) and the weak-form stop impossible
            # (it needs two stable passes in a row and the last pass was not stable).
            speculative = None
            if (self.speculate and iteration < limit and worst > SPECULATE_RESIDUAL * cfg.residual_tol
                    and stable == 0):
                self._filter(work, use32, degree)
                speculative = use32
            evals, norms, orth = _collect(work)
            hamiltonian.note_applications(degree + 2)
            before = worst
            worst = float(norms[:n_states].max())
            # A float32 pass that no longer shrinks the residual has reached float32's floor
            # (about eps32 times the spectral width); the rest of an "auto" solve runs in float64.
            if use32 and worst > self.stall_ratio * before:
                stalled32 = True
            drift = float(np.abs(evals[:n_states] - previous[:n_states]).max())
            next_lower = float(evals[min(size, evals.shape[0]) - 1])
            previous = evals.copy()
            self.history.append({"iteration": iteration, "residual": worst, "orthonormality": orth,
                                 "eigenvalue_drift": drift})
            if worst < cfg.residual_tol and iteration >= min_fp64_passes:
                converged = True
                stop_reason = "residual below residual_tol"
                break
            stable = stable + 1 if (weak_form and drift < eigenvalue_tol) else 0
            if stable >= 2:
                converged = True
                stop_reason = "Ritz values stable on the weak-form path"
                break
        if speculative is not None and (converged or iteration == limit):
            self.speculation["wasted"] += 1
        info = {"spectral_bounds": (spectrum_min, spectrum_max), "bounds_source": source, "bounds_shift": shift_used,
                "top_ritz_fraction": top_fraction, "stop_reason": stop_reason, "eigenvalue_drift": drift,
                "graphs": dict(self.graph_record)}
        vectors = work.vectors[:n_states].clone()
        return EigenResult(evals[:n_states], vectors, norms[:n_states], iteration, converged, info,
                           eigenvalues_device=work.evals[:n_states].clone())
