"""Differentiation of the Kohn-Sham hot path: the adjoint H-apply and the Sternheimer backward.

The reference defines the Kohn-Sham inversion seam (``InversionProtocol``,
``src/cdft/contract.py:1263-1283``) and leaves it a stub (``src/cdft/inversion/__init__.py:1``).
This module differentiates the discrete Kohn-Sham map
``V -> rho = sum_i f_i phi_i^2``, where ``A(V) phi_i = eps_i phi_i`` with ``A(V) = A_0 + diag(V)``.
Here ``A_0`` is the cusp-factorised operator and ``V`` the grid potential on its diagonal.

* :class:`HApply` wraps K1 as a ``torch.autograd.Function``. The backward applies the hand-written
  transpose ``A^T g = 0.5 sum_d D_d^T[w_d D_d (invw g)] + V g`` (``cdft_hdiv_apply_adjoint_f64``),
  and the potential's gradient is ``sum_k g_k x_k``.
* :class:`KohnShamDensity` wraps the eigensolve. Its backward is the Sternheimer equation.
  ``A`` is self-adjoint in the measure ``<a, b>_q = sum q a b``. For a loss ``L(rho)`` with
  ``G_i = dL/dphi_i = 2 f_i (dL/drho) phi_i``, each occupied state solves
  ``(A - eps_i) zeta_i = -P (G_i / q)`` on the unoccupied complement, where
  ``P = 1 - sum_occ phi_j <phi_j, .>_q``. The solver is conjugate gradients in the q inner product.
  ``A - eps_i`` is positive definite there when the occupied states are the lowest. The gradient
  is ``dL/dV = q * sum_i zeta_i phi_i``.
  With equal occupations the density is invariant under rotations of the occupied states, so the
  occupied-occupied couplings cancel. The occupied projector is therefore exact.

The gradient is checked against central finite differences of the GPU forward map, and against
torch autograd through a dense symmetric eigendecomposition of the oracle operator on a small
grid.
"""

from __future__ import annotations

from dataclasses import dataclass, replace

import numpy as np
import torch

from . import ops
from .eigen import ChebyshevFilteredSubspace, EigenConfig
from .hamiltonian import DeviceMeasure, lanczos_bounds, seeded_randn
from .ops import StencilGeometry


class StencilHamiltonian:
    """``A = 0.5 invw sum_d D_d^T[w_d D_d .] + diag(V)`` with its measure ``q``: the operator seam
    the eigensolver reads (reference ``operators/hamiltonian.py:178-396``), built from a stencil
    geometry rather than from a cusp factor."""

    uses_divergence_form = True

    def __init__(self, grid, geometry: StencilGeometry, q: torch.Tensor):
        self.grid = grid
        self.geometry = geometry
        self.potential = geometry.potential
        self.measure = DeviceMeasure(grid, q)
        self.apply_count = 0

    def apply(self, phi: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        self.apply_count += 1
        return ops.hdiv_apply(phi, self.geometry, out)

    def geometry_in(self, dtype: torch.dtype) -> StencilGeometry:
        cache = self.__dict__.setdefault("_cast", {})
        if dtype not in cache:
            cache[dtype] = self.geometry.to(dtype)
        return cache[dtype]

    def note_applications(self, count: int) -> None:
        self.apply_count += int(count)

    def spectral_bounds(self, steps: int = 12, safety: float = 1.02, generator=None) -> tuple[float, float]:
        start = seeded_randn(self.grid.n_points, generator, self.grid.device)
        return lanczos_bounds(self, start, steps, safety)


class HApply(torch.autograd.Function):
    """``y = A(V) x`` for a block ``(k, N)``; gradients for ``x`` (by the transpose apply) and for the
    diagonal potential ``V`` (``sum_k g_k x_k``)."""

    @staticmethod
    def forward(ctx, x: torch.Tensor, potential: torch.Tensor, geometry: StencilGeometry):
        geom = geometry.with_potential(potential.detach().contiguous())
        ctx.geometry = geom
        ctx.save_for_backward(x)
        return ops.hdiv_apply(x.detach().contiguous(), geom)

    @staticmethod
    def backward(ctx, grad):
        (x,) = ctx.saved_tensors
        g = grad.contiguous()
        gx = ops.hdiv_apply_adjoint(g, ctx.geometry) if ctx.needs_input_grad[0] else None
        gv = None
        if ctx.needs_input_grad[1]:
            gv = (g.reshape(-1, x.shape[-1]) * x.reshape(-1, x.shape[-1])).sum(dim=0)
        return gx, gv, None


@dataclass
class SternheimerConfig:
    """Conjugate-gradient settings of the backward: relative residual (in the q norm), iteration cap,
    and how often the residual is read back to the host."""

    tol: float = 1.0e-11
    max_iterations: int = 2000
    check_every: int = 10


@dataclass
class ForwardState:
    """What the backward needs: the occupied eigenpairs and the operator they belong to."""

    hamiltonian: StencilHamiltonian
    vectors: torch.Tensor  # (n_occ, N), q-orthonormal
    eigenvalues: np.ndarray
    occupations: np.ndarray
    result: object


class KohnShamProblem:
    """The map ``V -> rho`` on one geometry: base operator ``A_0`` (stencil geometry whose diagonal
    holds the fixed part of the potential), measure ``q`` and occupations."""

    def __init__(self, grid, base_geometry: StencilGeometry, q: torch.Tensor, occupations,
                 eigen: EigenConfig | None = None, sternheimer: SternheimerConfig | None = None, seed: int = 0):
        self.grid = grid
        self.base = base_geometry
        self.q = q
        self.occupations = np.asarray(occupations, dtype=np.float64)
        self.n_occ = int(np.count_nonzero(self.occupations))
        self.eigen = eigen if eigen is not None else EigenConfig(residual_tol=1.0e-11)
        self.cg = sternheimer if sternheimer is not None else SternheimerConfig()
        self.seed = seed
        self._warm: torch.Tensor | None = None
        self.history: list[dict] = []

    def hamiltonian(self, v: torch.Tensor) -> StencilHamiltonian:
        pot = ops.ewise(ops.EW_ADD, torch.empty_like(self.base.potential), self.base.potential, v.detach().contiguous())
        return StencilHamiltonian(self.grid, self.base.with_potential(pot), self.q)

    def forward(self, v: torch.Tensor) -> tuple[torch.Tensor, ForwardState]:
        """Solve ``A(V)``'s lowest states and return ``rho = sum_i f_i phi_i^2``."""
        ham = self.hamiltonian(v)
        solver = ChebyshevFilteredSubspace(self.grid, self.eigen, ritz_stop=False)
        gen = torch.Generator(device="cpu").manual_seed(self.seed)
        result = solver.solve(ham, self.n_occ, initial=self._warm, generator=gen)
        self._warm = result.eigenvectors.detach().clone()
        vectors = result.eigenvectors[: self.n_occ].contiguous()
        rho = torch.empty(self.grid.n_points, dtype=torch.float64, device=self.grid.device)
        ops.density(vectors, self.occupations[: self.n_occ], None, rho, None)
        return rho, ForwardState(ham, vectors, np.asarray(result.eigenvalues[: self.n_occ]),
                                 self.occupations[: self.n_occ], result)

    # ---- backward ----

    def _project(self, state: ForwardState, r: torch.Tensor, coef: torch.Tensor, dev: torch.Tensor) -> None:
        """``r -= sum_j phi_j <phi_j, r>_q`` in place (coefficients in ``coef``, copied to ``dev``)."""
        ops.wgram(state.vectors, r.reshape(1, -1), self.q, out=coef.view(-1, 1))
        dev[8:8 + self.n_occ].copy_(coef.view(-1))
        terms = [(r, 1.0)] + [(state.vectors[j], -1.0, 1, 8 + j) for j in range(self.n_occ)]
        ops.lincomb(r, terms, dev)

    def sternheimer(self, state: ForwardState, i: int, rhs: torch.Tensor) -> tuple[torch.Tensor, dict]:
        """Solve ``(A - eps_i) zeta = rhs`` for ``zeta`` in the q-orthogonal complement of the occupied
        states by conjugate gradients; ``rhs`` is projected first. Scalars stay on the device; the
        residual is read every ``check_every`` iterations."""
        n = self.grid.n_points
        dev = torch.zeros(8 + self.n_occ, dtype=torch.float64, device=self.grid.device)
        coef = torch.empty(self.n_occ, dtype=torch.float64, device=self.grid.device)
        eps = float(state.eigenvalues[i])
        r = rhs.clone()
        self._project(state, r, coef, dev)
        zeta = torch.zeros(n, dtype=torch.float64, device=self.grid.device)
        p = r.clone()
        ap = torch.empty_like(r)
        ops.reduce(ops.RED_DOT, r, dev, 0, b=r, w=self.q)  # rr
        rr0 = float(dev[0].item())
        info = {"iterations": 0, "relative_residual": 0.0 if rr0 == 0.0 else 1.0, "converged": rr0 == 0.0}
        if rr0 == 0.0:
            return zeta, info
        target = (self.cg.tol ** 2) * rr0
        for it in range(1, self.cg.max_iterations + 1):
            state.hamiltonian.apply(p.view(1, -1), ap.view(1, -1))
            ops.lincomb(ap, [(ap, 1.0), (p, -eps)])
            ops.reduce(ops.RED_DOT, p, dev, 1, b=ap, w=self.q)  # pAp
            ops.ewise(ops.EW_DIV, dev[2:3], dev[0:1], dev[1:2])  # alpha = rr / pAp
            ops.lincomb(zeta, [(zeta, 1.0), (p, 1.0, 1, 2)], dev)
            ops.lincomb(r, [(r, 1.0), (ap, -1.0, 1, 2)], dev)
            self._project(state, r, coef, dev)
            ops.reduce(ops.RED_DOT, r, dev, 3, b=r, w=self.q)  # rr_new
            ops.ewise(ops.EW_DIV, dev[4:5], dev[3:4], dev[0:1])  # beta
            ops.lincomb(p, [(r, 1.0), (p, 1.0, 1, 4)], dev)
            ops.ewise(ops.EW_SCALE, dev[0:1], dev[3:4], s=1.0)
            if it % self.cg.check_every == 0 or it == self.cg.max_iterations:
                rr = float(dev[0].item())
                info = {"iterations": it, "relative_residual": float(np.sqrt(rr / rr0)), "converged": rr <= target}
                if rr <= target:
                    break
        self._project(state, zeta, coef, dev)
        return zeta, info

    def backward(self, state: ForwardState, grad_rho: torch.Tensor) -> torch.Tensor:
        """``dL/dV`` from ``dL/drho``."""
        grad_v = torch.zeros(self.grid.n_points, dtype=torch.float64, device=self.grid.device)
        g = grad_rho.detach().contiguous().reshape(-1)
        solves = []
        for i in range(self.n_occ):
            phi = state.vectors[i]
            # rhs = -(G_i / q) with G_i = 2 f_i g phi_i (the projector is applied inside the solve)
            rhs = ops.ewise(ops.EW_MUL, torch.empty_like(g), g, phi)
            ops.ewise(ops.EW_DIV, rhs, rhs, self.q)
            ops.ewise(ops.EW_SCALE, rhs, rhs, s=-2.0 * float(state.occupations[i]))
            zeta, info = self.sternheimer(state, i, rhs)
            solves.append(info)
            # grad_v += q zeta phi
            ops.ewise(ops.EW_MUL3, rhs, zeta, phi, self.q)
            ops.ewise(ops.EW_ADD, grad_v, grad_v, rhs)
        self.history.append({"sternheimer": solves})
        return grad_v


class ChebyshevFilterFunction(torch.autograd.Function):
    """``Y = p_d(A(V)) X``: the scaled Chebyshev filter of degree ``d`` (the fused K2 steps, bitwise
    :func:`ops.cheb_filter`), differentiable in the block ``X`` and the diagonal potential ``V`` by the
    adjoint recurrence, with checkpointing of the filter iterates.

    With rows ``(c_j, a_j, b_j)`` the forward runs ``y_1 = a_0 (A - c_0) y_0`` and
    ``y_{j+1} = a_j (A - c_j) y_j - b_j y_{j-1}`` from ``y_0 = X``. For ``t = a_j ybar_{j+1}`` the
    backward adds ``(A^T - c_j) t`` to ``ybar_j`` (the hand-written transpose apply),
    ``-b_j ybar_{j+1}`` to ``ybar_{j-1}`` and ``sum_k t_k y_{j,k}`` to ``Vbar``, for ``j = d-1 .. 0``; it
    needs the iterates in reverse order. The forward keeps only every ``every``-th pair of consecutive
    iterates ``(y_{j-1}, y_j)``; the backward recomputes each segment's ``every`` iterates from its
    pair before walking back through it, so at most ``2 ceil(d / every) + every`` blocks are held
    instead of ``d + 1``. Call as ``ChebyshevFilterFunction.apply(x, v, geometry, rows, every)`` with
    ``rows`` the host rows of :func:`chefsi.scaled_chebyshev_table`; ``geometry``'s own potential is
    replaced by ``v``."""

    @staticmethod
    def forward(ctx, x, v, geometry: StencilGeometry, rows, every: int = 4):
        geom = geometry.with_potential(v.detach().contiguous())
        table = torch.tensor(rows, dtype=torch.float64, device=x.device)
        degree = len(rows)
        every = max(1, int(every))
        y_prev, y = None, x.detach().contiguous()
        checkpoints = {}
        for j in range(degree):
            if j % every == 0:
                checkpoints[j] = (y_prev, y)
            out = torch.empty_like(y)
            ops.cheb_step(y, y_prev, geom, table, j, out)
            y_prev, y = y, out
        ctx.geometry, ctx.table, ctx.rows, ctx.every = geom, table, [tuple(map(float, r)) for r in rows], every
        ctx.checkpoints = checkpoints
        ctx.held = 2 * len(checkpoints) + every  # blocks the backward holds at most (for the record)
        return y

    @staticmethod
    def backward(ctx, grad):
        geom, table, rows, every = ctx.geometry, ctx.table, ctx.rows, ctx.every
        degree = len(rows)
        k, n = grad.shape
        ybar_next = grad.detach().contiguous().clone()  # ybar_{j+1}
        ybar_cur = torch.zeros_like(ybar_next)          # ybar_j (accumulated)
        vbar = torch.zeros(n, dtype=torch.float64, device=grad.device)
        t = torch.empty_like(ybar_next)
        tmp = torch.empty(n, dtype=torch.float64, device=grad.device)
        for start in sorted(ctx.checkpoints, reverse=True):
            stop = min(start + every, degree)
            # recompute y_start .. y_{stop-1} from the checkpoint pair
            y_prev, y = ctx.checkpoints[start]
            iterates = {start: y, start - 1: y_prev}
            for j in range(start, stop - 1):
                out = torch.empty_like(y)
                ops.cheb_step(y, y_prev, geom, table, j, out)
                y_prev, y = y, out
                iterates[j + 1] = y
            for j in range(stop - 1, start - 1, -1):
                c, a, b = rows[j]
                ops.lincomb(t.view(-1), [(ybar_next.view(-1), a)])
                at = ops.hdiv_apply_adjoint(t, geom)
                ops.lincomb(ybar_cur.view(-1), [(ybar_cur.view(-1), 1.0), (at.view(-1), 1.0), (t.view(-1), -c)])
                yj = iterates[j]
                for r in range(k):
                    ops.ewise(ops.EW_MUL, tmp, t[r], yj[r])
                    ops.ewise(ops.EW_ADD, vbar, vbar, tmp)
                # shift: ybar_{j} becomes the next "next"; ybar_{j-1} starts as -b_j ybar_{j+1}
                if j >= 1:
                    ops.lincomb(t.view(-1), [(ybar_next.view(-1), -b)])
                    ybar_next, ybar_cur, t = ybar_cur, t, ybar_next
                else:
                    ybar_next = ybar_cur
        return ybar_next, vbar, None, None, None


class KohnShamDensity(torch.autograd.Function):
    """``rho(V)`` with the Sternheimer backward: ``KohnShamDensity.apply(v, problem)``."""

    @staticmethod
    def forward(ctx, v: torch.Tensor, problem: KohnShamProblem):
        rho, state = problem.forward(v)
        ctx.problem = problem
        ctx.state = state
        return rho

    @staticmethod
    def backward(ctx, grad_rho):
        return ctx.problem.backward(ctx.state, grad_rho), None


def density_loss(rho: torch.Tensor, target: torch.Tensor, weights: torch.Tensor) -> torch.Tensor:
    """``L = sum_p weights_p (rho_p - target_p)^2``; with ``weights = omega f^2`` (the mass weights
    times ``f^2``) this is ``integral (n - n_t)^2`` of the physical densities ``n = f^2 rho`` in
    the cusp-aware measure."""
    return (weights * (rho - target) ** 2).sum()


def tight(config: EigenConfig, residual_tol: float = 1.0e-11) -> EigenConfig:
    return replace(config, residual_tol=residual_tol)
