"""Rayleigh-Ritz and Chebyshev-filtered subspace iteration on the GPU, float64 throughout.

Mirrors ``src/cdft/eigen/rayleigh_ritz.py:51-115`` and ``src/cdft/eigen/chefsi.py:442-585``:
Cholesky orthonormalisation in the operator's measure, the projected k x k problem, rotation,
residual norms, and the outer loop with one host read per iteration. Every block-sized operation
is a HIP kernel (K1-K4); the k x k algebra runs in one-thread kernels on the device.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch

from . import ops
from .chefsi import chebyshev_coefficients
from .hamiltonian import seeded_randn

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
    shift: float = 0.0


class RitzWork:
    """Device buffers of one Rayleigh-Ritz step for a block of k vectors (allocated once)."""

    def __init__(self, k: int, n: int, device):
        f = dict(dtype=torch.float64, device=device)
        self.k = k
        self.S = torch.empty((k, k), **f)
        self.T = torch.empty((k, k), **f)
        self.M = torch.empty((k, k), **f)
        self.C = torch.empty((k, k), **f)
        self.G = torch.empty((k, k), **f)
        self.flag = torch.zeros(1, dtype=torch.int32, device=device)
        # scalars read once per iteration: evals (k) | residual norms (k) | orth error (1) | flag (1)
        self.scalars = torch.zeros(2 * k + 2, **f)
        self.basis = torch.empty((k, n), **f)
        self.applied = torch.empty((k, n), **f)
        self.vectors = torch.empty((k, n), **f)
        self.filtered = torch.empty((k, n), **f)
        self.work1 = torch.empty((k, n), **f)
        self.work2 = torch.empty((k, n), **f)

    @property
    def evals(self) -> torch.Tensor:
        return self.scalars[: self.k]

    @property
    def norms(self) -> torch.Tensor:
        return self.scalars[self.k: 2 * self.k]


def rayleigh_ritz(hamiltonian, block: torch.Tensor, work: RitzWork) -> None:
    """Orthonormalise ``block`` (Cholesky, ``T = L^-1``), project, diagonalise, rotate; leaves the
    eigenvalues in ``work.evals``, the Ritz vectors in ``work.vectors`` and ``max|G - I|`` in
    ``work.scalars[2k]``. A non-positive-definite Gram matrix sets ``work.flag``."""
    m = hamiltonian.measure
    k = work.k
    m.gram(block, out=work.S)
    ops.chol_inv(work.S, work.T, work.flag)
    ops.rotate(work.T, block, out=work.basis)
    hamiltonian.apply(work.basis, work.applied)
    m.cross(work.basis, work.applied, out=work.M)
    ops.eigh(work.M, work.evals, work.C)
    ops.rotate(work.C, work.basis, transpose=True, out=work.vectors)
    m.gram(work.vectors, out=work.G)
    ops.orth_error(work.G, work.scalars[2 * k: 2 * k + 1])


def residual_norms(hamiltonian, work: RitzWork) -> None:
    """``||A phi_i - eps_i phi_i||`` in the measure, into ``work.norms`` (fresh apply, as the
    reference's ``residual_norms``)."""
    hamiltonian.apply(work.vectors, work.applied)
    ops.resnorm(work.applied, work.vectors, work.evals, hamiltonian.measure.q, out=work.norms)


def _read(work: RitzWork) -> tuple[np.ndarray, np.ndarray, float]:
    flag = work.flag.to("cpu", non_blocking=True)
    host = work.scalars.cpu().numpy()  # the iteration's one synchronising read (the flag rides along)
    if int(flag[0]) != 0:
        raise RuntimeError(
            "Rayleigh-Ritz: the filtered block's Gram matrix is not positive definite (the block has "
            "collapsed in rank); the eigh fallback of the reference that drops vectors is not implemented")
    k = work.k
    return host[:k].copy(), host[k:2 * k].copy(), float(host[2 * k])


def _fraction(value, low, high):
    width = high - low
    return (value - low) / width if width > 0.0 else float("inf")


class ChebyshevFilteredSubspace:
    """CheFSI (reference ``eigen/chefsi.py:408-585``) on the device."""

    #: Mixed precision: filter in float32 while the worst residual is above this...
    switch_residual: float = 1.0e-5
    #: ...and always finish an iteration-limited solve with this many float64 filter passes.
    fp64_tail: int = 1

    def __init__(self, grid, config: EigenConfig, ritz_stop: bool = True, filter_dtype=torch.float64,
                 filter_hook=None):
        self.grid = grid
        self.config = config
        self.ritz_stop = ritz_stop
        self.filter_dtype = filter_dtype
        self.filter_hook = filter_hook
        self.history: list[dict] = []
        self.filter_precisions: list[str] = []
        self._work: dict[int, RitzWork] = {}

    def work_for(self, k: int) -> RitzWork:
        if k not in self._work:
            self._work[k] = RitzWork(k, self.grid.n_points, self.grid.device)
        return self._work[k]

    def filter(self, hamiltonian, block: torch.Tensor, out: torch.Tensor, work: RitzWork, lower, upper, smin,
               dtype=torch.float64):
        rows = chebyshev_coefficients(self.config.chebyshev_degree, lower, upper, smin)
        table = torch.tensor(rows, dtype=torch.float64).pin_memory().to(block.device, non_blocking=True)
        if self.filter_hook is not None:
            self.filter_hook("start", self, block if dtype == torch.float64 else block.to(dtype), len(rows))
        if dtype == torch.float64:
            ops.cheb_filter(block, hamiltonian.geometry, table, len(rows), out=out, work=(work.work1, work.work2))
            self.filter_precisions.append("f64")
        else:
            g32 = hamiltonian.geometry_in(dtype)
            x32 = block.to(dtype)
            y32 = ops.cheb_filter(x32, g32, table, len(rows))
            out.copy_(y32)
            self.filter_precisions.append("f32")
        if self.filter_hook is not None:
            self.filter_hook("end", self, block if dtype == torch.float64 else x32, len(rows))
        hamiltonian.note_applications(len(rows))
        return out

    def solve(self, hamiltonian, n_states: int, initial: torch.Tensor | None = None, generator=None,
              bounds: SpectralBoundsHint | None = None) -> EigenResult:
        cfg = self.config
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

        def lanczos():
            return hamiltonian.spectral_bounds(steps=cfg.lanczos_steps, safety=cfg.bound_safety, generator=generator)

        source, shift_used = "unset", 0.0
        if bounds is None:
            spectrum_min, spectrum_max = lanczos()
            source = "lanczos"
        rayleigh_ritz(hamiltonian, block, work)
        residual_norms(hamiltonian, work)
        evals, norms, _ = _read(work)
        worst = float(norms[:n_states].max())
        previous = evals.copy()
        stable = 0
        stop_reason = "iteration limit"
        weak_form = self.ritz_stop and getattr(hamiltonian, "uses_divergence_form", False)
        eigenvalue_tol = 0.1 * cfg.cross_check_tol
        next_lower = float(evals[min(size, evals.shape[0]) - 1])
        if bounds is not None:
            spectrum_min, spectrum_max = bounds.lower - bounds.shift, bounds.upper + bounds.shift
            source = "reused"
            shift_used = float(bounds.shift)
        top_fraction = float("-inf")
        converged = False
        drift = float("inf")
        iteration = 0
        for iteration in range(1, cfg.max_iterations + 1):
            lower = next_lower
            fraction = _fraction(lower, spectrum_min, spectrum_max)
            if source == "reused" and not fraction < BOUND_REUSE_RITZ_FRACTION:
                spectrum_min, spectrum_max = lanczos()
                source = f"lanczos (reuse rejected at iteration {iteration})"
                fraction = _fraction(lower, spectrum_min, spectrum_max)
            top_fraction = max(top_fraction, fraction)
            lower = min(max(lower, spectrum_min + 1e-8), spectrum_max - 1e-8)
            # The Ritz vectors of the previous step are the filter input.
            src = work.vectors
            dtype = torch.float64
            if self.filter_dtype != torch.float64 and worst > self.switch_residual and \
                    iteration <= cfg.max_iterations - self.fp64_tail:
                dtype = self.filter_dtype
            self.filter(hamiltonian, src, work.filtered, work, lower, spectrum_max, spectrum_min, dtype)
            rayleigh_ritz(hamiltonian, work.filtered, work)
            residual_norms(hamiltonian, work)
            evals, norms, orth = _read(work)
            worst = float(norms[:n_states].max())
            drift = float(np.abs(evals[:n_states] - previous[:n_states]).max())
            next_lower = float(evals[min(size, evals.shape[0]) - 1])
            previous = evals.copy()
            self.history.append({"iteration": iteration, "residual": worst, "orthonormality": orth, "eigenvalue_drift": drift})
            if worst < cfg.residual_tol:
                converged = True
                stop_reason = "residual below residual_tol"
                break
            stable = stable + 1 if (weak_form and drift < eigenvalue_tol) else 0
            if stable >= 2:
                converged = True
                stop_reason = "Ritz values stable on the weak-form path"
                break
        info = {"spectral_bounds": (spectrum_min, spectrum_max), "bounds_source": source, "bounds_shift": shift_used,
                "top_ritz_fraction": top_fraction, "stop_reason": stop_reason, "eigenvalue_drift": drift}
        vectors = work.vectors[:n_states].clone()
        return EigenResult(evals[:n_states], vectors, norms[:n_states], iteration, converged, info,
                           eigenvalues_device=work.evals[:n_states].clone())
