"""LOBPCG, the second eigensolver of the two-solver cross-check (reference ``src/cdft/eigen/lobpcg.py``).

Matrix-free block LOBPCG in the operator's measure, on the device: the operator is the stencil
kernel, inner products and Gram matrices are the weighted-Gram kernel, rotations the rotation kernel,
the small symmetric eigenproblems (at most 3 n_states <= 16) the register Jacobi kernel, and the
Teter-Payne-Allan kinetic preconditioner the FFT of ``csrc/precond.hip``. As in the reference, the
host reads a few scalars per iteration (the residual norms, the Ritz values and the Gram spectra of
the stable-basis steps) and decides convergence, deflation and the directions dropped.

Gate G2.5 (``docs/03_METHOD.md:433`` of the reference) compares its eigenvalues with CheFSI's to
``cross_check_tol``; ``scf.solve_self_consistent`` and ``scf.solve_noninteracting`` run it with
``cross_check=True``.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib, ops
from .eigen import EigenConfig, EigenResult
from .hamiltonian import seeded_randn

#: Where the kinetic preconditioner acts (reference ``lobpcg.py:28-41``).
PRECONDITIONER_FRAMES = ("similarity", "transformed", "none")

#: Gram eigenvalues at most this fraction of the largest are dropped from a basis (``lobpcg.py:254``).
DEPENDENCE_THRESHOLD = 1e-12

#: The largest block the small dense kernels take.
MAX_BASIS = 16


class KineticPreconditioner:
    """``K`` on the box by FFT, optionally in the similarity frame ``f^-1 K f`` (``lobpcg.py:207-239``).
    Plans, the filter and the work buffers belong to one grid and device; build them through
    :func:`preconditioner_for`, which keeps one per grid. The finaliser does not call the device
    (it can run at any allocation, including inside a graph capture, where a free invalidates the
    capture): it hands the plans' destruction to :func:`ops.release_later`."""

    def __init__(self, grid):
        self.grid = grid
        n0, n1, n2 = grid.shape
        handle = ctypes.c_void_p()
        _lib.call("cdft_tpa_create", n0, n1, n2, ctypes.byref(handle))
        self._handle = handle
        half = n0 * n1 * (n2 // 2 + 1)
        dev = grid.device
        self.filter = torch.empty(half, dtype=torch.float64, device=dev)
        self._real = torch.empty(grid.n_points, dtype=torch.float64, device=dev)
        self._spectrum = torch.empty(2 * half, dtype=torch.float64, device=dev)
        e_ref = ctypes.c_double()
        _lib.call("cdft_tpa_filter_f64", handle, float(grid.spacing), self.filter.data_ptr(), ctypes.byref(e_ref),
                  ops.stream_of(self.filter))
        #: the mean kinetic scale of the grid (``lobpcg.py:225-227``)
        self.kinetic_reference = e_ref.value

    def apply(self, block: torch.Tensor, u: torch.Tensor | None = None, out: torch.Tensor | None = None) -> torch.Tensor:
        """``f^-1 K (f r)`` for every row ``r`` of ``block`` (``u`` the cusp exponent, ``f = exp(-u)``
        clamped at 1e-150), or ``K r`` when ``u`` is None."""
        n = self.grid.n_points
        rows = block.reshape(-1, n)
        ops.require(rows, (rows.shape[0], n), torch.float64, self.grid.device, "block")
        if u is not None:
            ops.require(u, (n,), torch.float64, self.grid.device, "u")
        out = torch.empty_like(rows) if out is None else out
        _lib.call("cdft_tpa_apply_f64", self._handle, rows.data_ptr(), rows.shape[0], ops._ptr(u), self.filter.data_ptr(),
                  self._real.data_ptr(), self._spectrum.data_ptr(), out.data_ptr(), ops.stream_of(rows))
        return out

    def __del__(self, _release_later=ops.release_later, _lib=_lib):
        handle = getattr(self, "_handle", None)
        if handle is not None and handle.value:
            _release_later(lambda h=handle: _lib.call("cdft_tpa_destroy", h))
            self._handle = None


_preconditioners: dict = {}


def preconditioner_for(grid) -> KineticPreconditioner:
    """The preconditioner of a grid shape, spacing and device, built once."""
    key = (tuple(grid.shape), float(grid.spacing), ops.workspace_key(grid.device))
    held = _preconditioners.get(key)
    if held is None:
        held = _preconditioners[key] = KineticPreconditioner(grid)
    return held


class LOBPCG:
    """Locally optimal block preconditioned conjugate gradient (reference ``lobpcg.py:44-196``)."""

    def __init__(self, grid, config: EigenConfig, precondition: bool | str = True) -> None:
        self.grid = grid
        self.config = config
        if precondition is True:
            precondition = "similarity"
        elif precondition is False:
            precondition = "none"
        if precondition not in PRECONDITIONER_FRAMES:
            raise ValueError(f"unknown preconditioner frame {precondition!r}; one of {PRECONDITIONER_FRAMES}")
        self.frame: str = precondition
        self.history: list[dict[str, float]] = []
        #: Ritz-value stability that also ends the solve: an order tighter than ``cross_check_tol``.
        self.eigenvalue_tol = 0.1 * config.cross_check_tol
        self.eigenvalue_drift = float("inf")

    # -- small dense steps ------------------------------------------------------------------------

    def _stable_basis(self, block: torch.Tensor, measure) -> tuple[torch.Tensor, int]:
        """An orthonormal basis of the rows' span, near-dependent directions dropped (``lobpcg.py:254-271``)."""
        k = block.shape[0]
        if k == 0:
            return block, 0
        if k > MAX_BASIS:
            raise ValueError(f"LOBPCG basis of {k} vectors exceeds the small-matrix kernels' {MAX_BASIS}")
        gram = measure.gram(block)
        evals = torch.empty(k, dtype=torch.float64, device=block.device)
        vecs = torch.empty((k, k), dtype=torch.float64, device=block.device)
        ops.eigh(gram, evals, vecs)
        host = torch.cat([evals, vecs.reshape(-1)]).cpu().numpy()  # one read: spectrum and vectors
        lam, c = host[:k], host[k:].reshape(k, k)
        largest = lam[-1]
        if not largest > 0.0:
            return block[:0], k
        keep = lam > largest * DEPENDENCE_THRESHOLD
        scaled = c[:, keep] / np.sqrt(lam[keep])  # (k, kept): basis = scaled^T block
        coeff = torch.as_tensor(np.ascontiguousarray(scaled), device=block.device)
        return ops.rotate(coeff, block, transpose=True), int((~keep).sum())

    @staticmethod
    def _project_out(block: torch.Tensor, basis: torch.Tensor, measure) -> torch.Tensor:
        """``block - <block, basis> basis`` (two classical Gram-Schmidt passes, ``lobpcg.py:242-251``)."""
        for _ in range(2):
            c = measure.cross(block, basis)
            along = ops.rotate(c, basis)
            ops.lincomb(block, [(block, 1.0), (along, -1.0)])
        return block

    def _orthonormalise(self, block: torch.Tensor, measure) -> torch.Tensor:
        """Cholesky in the measure, the stable basis when the Gram matrix is not positive definite
        (``lobpcg.py:242-251`` via ``_orthonormalise``)."""
        k = block.shape[0]
        gram = measure.gram(block)
        t = torch.empty((k, k), dtype=torch.float64, device=block.device)
        flag = torch.zeros(1, dtype=torch.int32, device=block.device)
        ops.chol_inv(gram, t, flag)
        if int(flag.item()) != 0:
            return self._stable_basis(block, measure)[0]
        return ops.rotate(t, block)

    def _ritz(self, basis: torch.Tensor, h_basis: torch.Tensor, measure, n_states: int):
        """The lowest ``n_states`` Ritz pairs of the basis: values (device) and ``(X, HX)``."""
        kb = basis.shape[0]
        sub = measure.cross(basis, h_basis)
        evals = torch.empty(kb, dtype=torch.float64, device=basis.device)
        vecs = torch.empty((kb, kb), dtype=torch.float64, device=basis.device)
        ops.eigh(sub, evals, vecs)  # the symmetric part, ascending, vectors in the columns
        coeff = vecs[:, :n_states].contiguous()
        return evals[:n_states].clone(), ops.rotate(coeff, basis, transpose=True), ops.rotate(coeff, h_basis, transpose=True)

    # -- the solve --------------------------------------------------------------------------------

    def solve(self, hamiltonian, n_states: int, initial: torch.Tensor | None = None,
              generator: torch.Generator | None = None) -> EigenResult:
        """The lowest ``n_states`` eigenpairs (reference ``lobpcg.py:71-165``)."""
        grid = self.grid
        cfg = self.config
        measure = hamiltonian.measure
        if 3 * n_states > MAX_BASIS:
            raise ValueError(f"LOBPCG on the device takes at most {MAX_BASIS // 3} states, not {n_states}")
        factor = getattr(hamiltonian, "factor", None)
        u = factor.u if (self.frame == "similarity" and factor is not None and not factor.is_identity) else None
        pre = preconditioner_for(grid) if self.frame != "none" else None
        n = grid.n_points
        if initial is None:
            x = seeded_randn((n_states, n), generator, grid.device)
        else:
            x = initial.to(torch.float64).reshape(-1, n)[:n_states].contiguous().clone()
        x = self._orthonormalise(x, measure)
        hx = hamiltonian.apply(x)
        evals, x, hx = self._ritz(x, hx, measure, n_states)

        p = None
        converged = False
        iteration = 0
        stable = 0
        previous = None
        residual_host = np.full(n_states, np.inf)
        r = torch.empty_like(x)
        norms = torch.empty(n_states, dtype=torch.float64, device=x.device)
        self.eigenvalue_drift = float("inf")
        self.history = []
        for iteration in range(1, cfg.max_iterations + 1):
            for i in range(n_states):
                ops.lincomb(r[i], [(hx[i], 1.0), (x[i], -1.0, 1, i)], evals)
            ops.resnorm(hx, x, evals, measure.q, norms)
            host = torch.cat([norms, evals]).cpu().numpy()  # the iteration's one read of scalars
            residual_host, ev_host = host[:n_states], host[n_states:]
            keep = residual_host > 0.1 * cfg.residual_tol
            worst = float(residual_host.max())
            drift = float(np.abs(ev_host - previous).max()) if previous is not None else float("inf")
            previous = ev_host
            self.eigenvalue_drift = drift
            self.history.append({"iteration": float(iteration), "residual": worst, "eigenvalue_drift": drift})
            if worst < cfg.residual_tol:
                converged = True
                break
            stable = stable + 1 if drift < self.eigenvalue_tol else 0
            if stable >= 2:
                converged = True
                break
            # deflate converged directions (a search along one at the noise floor injects noise)
            w = r[torch.as_tensor(np.flatnonzero(keep), device=r.device)] if keep.any() else r.clone()
            w = w.contiguous()
            if pre is not None:
                pre.apply(w, u, out=w)
            w = self._project_out(w, x, measure)
            w, _ = self._stable_basis(w, measure)
            if w.shape[0] == 0:
                converged = True
                break
            blocks = [x, w] if p is None else [x, w, p]
            basis, dropped = self._stable_basis(torch.cat(blocks, dim=0), measure)
            h_basis = hamiltonian.apply(basis)
            x_new_evals, x_new, hx_new = self._ritz(basis, h_basis, measure, n_states)
            # the conjugate direction: the part of the update not already in X
            p = x_new.clone()
            self._project_once(p, x, measure)
            x, hx, evals = x_new, hx_new, x_new_evals
            if dropped:
                self.history[-1]["dropped_directions"] = float(dropped)
        return EigenResult(eigenvalues=evals.cpu().numpy(), eigenvectors=x, residuals=residual_host,
                           n_iterations=iteration, converged=converged,
                           info={"history": self.history, "eigenvalue_drift": self.eigenvalue_drift,
                                 "frame": self.frame},
                           eigenvalues_device=evals)

    @staticmethod
    def _project_once(block: torch.Tensor, basis: torch.Tensor, measure) -> None:
        c = measure.cross(block, basis)
        ops.lincomb(block, [(block, 1.0), (ops.rotate(c, basis), -1.0)])
