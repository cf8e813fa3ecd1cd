"""The Kohn-Sham map and the SCF / non-interacting drivers, on the GPU.

Mirrors ``KohnShamStep`` (reference ``src/cdft/scf/step.py:111-476``), ``solve_self_consistent``
(``scf/loop.py:255-550`` with ``BoundReusePolicy`` 94-170 and ``initial_density`` 173-193),
``PulayMixer`` (``scf/mixing.py:86-162``), ``aufbau_occupations`` (``scf/occupations.py:16-48``)
and ``solve_noninteracting`` (``scf/noninteracting.py:118-330``). Every grid- or node-sized array
lives on the device and is produced by a HIP kernel; the host keeps the control flow and reads the
iteration's scalars once, as the reference does.
"""

from __future__ import annotations

import ctypes
import math
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib, ops
from .eigen import ChebyshevFilteredSubspace, DeviceArray, DeviceScalar, EigenConfig, SpectralBoundsHint
from .geometry import CuspFactor, CuspQuadrature, UniformGrid
from .hamiltonian import CuspFactoredHamiltonian, SeededDraws, prefetch_draws
from .ops import stream_of
from .poisson import CoulombCutoffPoisson

FILTER_STEPS_PER_ITERATION = 4
LANCZOS_REUSE_BACKSTOP = 8
LANCZOS_REUSE_MAX_WIDENING = 0.05
BOUND_REUSE_RITZ_FRACTION = 0.9


def aufbau(eigenvalues, n_electrons: float, capacity: float = 2.0) -> np.ndarray:
    """Fill from the bottom, fractional frontier allowed (reference ``scf/occupations.py:16-70``)."""
    occ = np.zeros(len(eigenvalues))
    remaining = float(n_electrons)
    for i in range(len(occ)):
        if remaining <= 1e-14:
            break
        take = min(capacity, remaining)
        occ[i] = take
        remaining -= take
    if remaining > 1e-10:
        raise ValueError(f"{n_electrons} electrons do not fit in {len(occ)} states of capacity {capacity}")
    return occ


def nuclear_repulsion(charges, positions) -> float:
    pos = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
    e = 0.0
    for i in range(len(charges)):
        for j in range(i + 1, len(charges)):
            e += charges[i] * charges[j] / float(np.linalg.norm(pos[i] - pos[j]))
    return e


class KohnShamStep:
    """One application of the Kohn-Sham map on the cusp-factorised path, spin-restricted LDA."""

    def __init__(self, scenario, numerics, grid: UniformGrid, factor: CuspFactor, quadrature: CuspQuadrature):
        self.scenario = scenario
        self.numerics = numerics
        xc = getattr(scenario, "xc", "lda")
        if xc not in ("lda", "pbe", "none"):
            raise ValueError(f"unsupported functional {xc!r} (lda or pbe)")
        self.gga = xc == "pbe"
        self.grid = grid
        self.factor = factor
        self.q = quadrature
        self.charges = tuple(float(z) for z in scenario.charges)
        self.positions = np.asarray(scenario.positions, dtype=np.float64).reshape(-1, 3)
        self.n_electrons = float(scenario.n_electrons)
        self.n_states = max(1, int(-(-self.n_electrons // 2.0)))
        self.ion_ion = nuclear_repulsion(self.charges, self.positions)
        self.poisson = CoulombCutoffPoisson(grid, numerics.poisson_pad_factor)
        self.bare = CuspFactoredHamiltonian(grid, factor, quadrature)
        self.draws: dict = {}  # seeded start vectors of this solve, by position in the draw sequence
        dev = grid.device
        na = len(self.charges)
        dist = [[float(np.linalg.norm(self.positions[a] - self.positions[b])) for b in range(na)] for a in range(na)]
        overlap = [[math.exp(-2.0 * zb * dist[a][b]) for b, zb in enumerate(self.charges)] for a in range(na)]
        self.overlap = torch.tensor(overlap, dtype=torch.float64, device=dev)
        np.linalg.solve(np.array(overlap), np.ones(na))  # validated once, as the reference does
        nucleus_f2 = [math.exp(-2.0 * sum(zb * dist[a][b] for b, zb in enumerate(self.charges) if b != a)) for a in range(na)]
        self.nucleus_f2 = torch.tensor(nucleus_f2, dtype=torch.float64, device=dev)
        n, m = grid.n_points, quadrature.n_nodes
        self.cusp_grid = torch.empty((na, n), dtype=torch.float64, device=dev)
        self.v1s_grid = torch.empty((na, n), dtype=torch.float64, device=dev)
        self.v1s_node = torch.empty((na, m), dtype=torch.float64, device=dev)
        s = stream_of(self.cusp_grid)
        _lib.call("cdft_hydrogenic_fields_f64", None, n, grid.c_shape, grid.spacing, grid.c_origin, na,
                  factor.c_charges, factor.c_positions, self.cusp_grid.data_ptr(), self.v1s_grid.data_ptr(), s)
        _lib.call("cdft_hydrogenic_fields_f64", quadrature.points.data_ptr(), m, grid.c_shape, grid.spacing,
                  grid.c_origin, na, factor.c_charges, factor.c_positions, None, self.v1s_node.data_ptr(), s)
        self.f2 = factor.weight
        self._alloc()

    def _alloc(self):
        g, m = self.grid, self.q.n_nodes
        f = dict(dtype=torch.float64, device=g.device)
        # the smooth Hartree potential and rho are the two rows the nodes interpolate in one launch
        self.interp_in = torch.empty((2, g.n_points), **f)
        self.v_smooth, self.rho = self.interp_in[0], self.interp_in[1]
        self.n_smooth = torch.empty(g.n_points, **f)
        # Hartree and XC potentials as two rows (grid and nodes): one launch forms both weighted rows
        self.v_grid2 = torch.empty((2, g.n_points), **f)
        self.v_h_grid, self.v_grid = self.v_grid2[0], self.v_grid2[1]
        self.interp = torch.empty((2, m), **f)
        self.v_node2 = torch.empty((2, m), **f)
        self.v_h_node, self.v_node = self.v_node2[0], self.v_node2[1]
        self.n_node = torch.empty(m, **f)
        self.e_node = torch.empty(m, **f)
        self.e_grid = torch.empty(g.n_points, **f)
        self.node_x = torch.empty((2, m), **f)
        self.grid_add = torch.empty((2, g.n_points), **f)
        self.tmp_node = torch.empty(m, **f)
        self.tmp_grid = torch.empty(g.n_points, **f)
        self.coefficients = torch.empty(len(self.charges), **f)
        self.rho_nuclei = torch.empty(len(self.charges), **f)
        self.nucleus_stencils = self.q.stencils_at(self.positions)
        # static inputs and outputs of the captured potentials / energy graphs
        self.pot_in = torch.empty(g.n_points, **f)
        self.pot_lumped = torch.empty((2, g.n_points), **f)
        self.pot_effective = torch.empty(g.n_points, **f)
        self.pot_scalars = torch.empty(3, **f)
        if self.gga:
            # the sigma path: grad rho on the grid and at the nodes, the pointwise terms at both, the
            # lifted and the grid flux (see _gga_xc)
            self.grad_rho = torch.empty((3, g.n_points), **f)
            self.node_grad = torch.empty((3, m), **f)
            self.node_rows = torch.empty((3, m), **f)   # W f2 v_H | W f2 v_n | W cross
            self.grid_rows = torch.empty((3, g.n_points), **f)  # far_w f2 v_H | far_w f2 v_n | far_w cross
            self.node_flux = torch.empty((3, m), **f)   # W flux
            self.grid_flux = torch.empty((3, g.n_points), **f)  # far flux
            self.lumped3 = torch.empty((3, g.n_points), **f)
            self.lifted_flux = torch.empty((3, g.n_points), **f)
            self.near_div = torch.empty(g.n_points, **f)
            self.far_div = torch.empty(g.n_points, **f)
            self.xc_weights = torch.empty(g.n_points, **f)
        self.graphs: dict = {}
        self.graph_record = {"captures": 0, "replays": 0, "disabled": None}
        self.use_graphs = True

    def _replay(self, key, body) -> None:
        """Run ``body`` eagerly on its first call (which allocates any lazy workspaces), capture it as
        a HIP graph on its second, and replay the graph after that. A capture failure is recorded
        and the body then runs eagerly."""
        if not self.use_graphs or self.graph_record["disabled"] is not None:
            body()
            return
        entry = self.graphs.get(key)
        if entry is None:
            self.graphs[key] = "eager"
            body()
            return
        if entry == "eager":
            try:
                graph = ops.capture_graph(body)
                self.graphs[key] = graph
                self.graph_record["captures"] += 1
            except Exception as exc:  # noqa: BLE001 - recorded; the body then runs eagerly
                self.graph_record["disabled"] = f"capture failed: {type(exc).__name__}: {exc}"
                body()
                return
            entry = graph
        entry.replay()
        self.graph_record["replays"] += 1

    # ---- potentials ----

    def potentials(self, density: torch.Tensor) -> dict:
        """Hartree (cusp split + FFT) and LDA lumped weights (reference ``scf/step.py:224-313``), one
        captured graph over static buffers after the first call.

        Returns device tensors; the three energies sit in ``scalars = [E_H, E_xc, int n v_xc]``.
        """
        self.pot_in.copy_(density.reshape(-1, self.grid.n_points)[0])
        self._replay("potentials", self._potentials_body)
        return {"effective": self.pot_effective.clone(), "scalars": self.pot_scalars.clone()}

    def _potentials_body(self) -> None:
        q, g = self.q, self.grid
        n_grid = self.pot_in
        ops.ewise(ops.EW_DIV, self.rho, n_grid, self.f2)
        q.interpolate_with(self.nucleus_stencils, self.rho, out=self.rho_nuclei)
        ops.small_solve(self.overlap, self.nucleus_f2, self.rho_nuclei, self.coefficients)
        ops.atom_combine(n_grid, self.coefficients, self.cusp_grid, -1.0, self.n_smooth)
        self.poisson.solve(self.n_smooth, self.v_smooth)
        ops.atom_combine(self.v_smooth, self.coefficients, self.v1s_grid, +1.0, self.v_h_grid)
        q.interpolate(self.interp_in, out=self.interp)
        ops.atom_combine(self.interp[0], self.coefficients, self.v1s_node, +1.0, self.v_h_node)
        if self.gga:
            self._gga_xc()
            return
        ops.ewise(ops.EW_MUL, self.n_node, q.f2, self.interp[1])
        ops.lda_vwn(n_grid, self.e_grid, self.v_grid)
        ops.lda_vwn(self.n_node, self.e_node, self.v_node)
        # node parts W P (f2 v), grid parts far_w (f2 v): Hartree in row 0, XC in row 1
        # (f2 v) w in one pass each: the same products as w (f2 v), so the same bits
        ops.ewise(ops.EW_MUL3_ROWS, self.node_x, q.f2, self.v_node2, q.weights)
        ops.ewise(ops.EW_MUL3_ROWS, self.grid_add, self.f2, self.v_grid2, q.far_weights)
        lumped = q.lift(self.node_x, add=self.grid_add, out=self.pot_lumped)
        hartree_weights, xc_weights = lumped[0], lumped[1]
        scalars = self.pot_scalars
        ops.reduce(ops.RED_DOT, q.weights, scalars, 1, b=self.e_node)
        ops.reduce_dots([(hartree_weights, self.rho, None, 0, 0.5, False),
                         (q.far_weights, self.e_grid, None, 1, 1.0, True),
                         (xc_weights, self.rho, None, 2, 1.0, False)], scalars)
        ops.ewise(ops.EW_SUMDIV, self.pot_effective, hartree_weights, xc_weights, q.mass_weights)

    def _gga_xc(self) -> None:
        """The PBE part of the potentials (reference ``scf/step.py:253-279``): sigma from the order-8 grid
        gradient of rho (interpolated to the nodes), the pointwise terms, then the XC lumped weights as
        the scalar part ``lumped(f2 v_n)`` plus the gradient part ``-div_FD(L^T (W flux)) - h^3
        div_FD(far flux)`` (``quadrature.py:640-650,718-729``) plus the cross part ``lumped(-2 flux .
        grad u)``, and the Hartree lumped weights in the same lift."""
        q, g, fac = self.q, self.grid, self.factor
        g.gradient(self.rho, out=self.grad_rho.view(1, 3, -1))
        q.interpolate(self.grad_rho, out=self.node_grad)
        # Hartree rows (f2 v_H) W at the nodes and (f2 v_H) far_w on the grid
        ops.ewise(ops.EW_MUL3, self.node_rows[0], q.f2, self.v_h_node, q.weights)
        ops.ewise(ops.EW_MUL3, self.grid_rows[0], self.f2, self.v_h_grid, q.far_weights)
        ops.gga_points(self.interp[1], self.node_grad, q.f2, q.grad_u, q.weights, q.weights, self.e_node,
                       self.node_rows[1], self.node_rows[2], self.node_flux)
        ops.gga_points(self.rho, self.grad_rho, self.f2, fac.grad_u, q.far_weights, q.far, self.e_grid,
                       self.grid_rows[1], self.grid_rows[2], self.grid_flux)
        q.lift(self.node_rows, add=self.grid_rows, out=self.lumped3)
        q.lift(self.node_flux, out=self.lifted_flux)
        ops.divergence(self.lifted_flux, g, self.near_div)
        ops.divergence(self.grid_flux, g, self.far_div)
        hartree_weights = self.lumped3[0]
        ops.gga_combine(self.lumped3[1], self.near_div, self.far_div, self.lumped3[2], g.volume_element, self.xc_weights)
        scalars = self.pot_scalars
        ops.reduce(ops.RED_DOT, q.weights, scalars, 1, b=self.e_node)
        ops.reduce_dots([(hartree_weights, self.rho, None, 0, 0.5, False),
                         (q.far_weights, self.e_grid, None, 1, 1.0, True),
                         (self.xc_weights, self.rho, None, 2, 1.0, False)], scalars)
        ops.ewise(ops.EW_SUMDIV, self.pot_effective, hartree_weights, self.xc_weights, q.mass_weights)

    def hamiltonian(self, pot: dict) -> CuspFactoredHamiltonian:
        return CuspFactoredHamiltonian(self.grid, self.factor, self.q, extra_potential=pot["effective"])

    def diagonalise(self, solver: ChebyshevFilteredSubspace, h, initial=None, max_iterations=None, bounds=None,
                    precision="auto", min_fp64_passes=0):
        gen = SeededDraws(self.numerics.seed, self.draws)
        return solver.solve(h, self.n_states, initial=initial, generator=gen, bounds=bounds, precision=precision,
                            min_fp64_passes=min_fp64_passes, max_iterations=max_iterations)

    def output_density(self, eigen, occ) -> tuple[torch.Tensor, torch.Tensor]:
        g = self.grid
        rho = g.empty()
        n = g.empty(1)
        ops.density(eigen.eigenvectors, occ, self.f2, rho, n[0])
        return n, rho

    # ---- energies ----

    def energy_scalars(self, phi: torch.Tensor, occ, rho_out: torch.Tensor, effective_in: torch.Tensor) -> torch.Tensor:
        """Device buffer ``[kinetic form (k x k, flattened) | cross | potential | dc_out]`` (reference
        ``operators/hamiltonian.py:344-377``, ``scf/step.py:398-439``), one captured graph over
        static buffers per block size after the first call."""
        k = phi.shape[0]
        key = ("energy", k, effective_in is None)
        st = self.__dict__.setdefault("_energy_static", {})
        if key not in st:
            f = dict(dtype=torch.float64, device=self.grid.device)
            st[key] = (torch.empty((k, self.grid.n_points), **f), torch.empty(self.grid.n_points, **f),
                       None if effective_in is None else torch.empty(self.grid.n_points, **f),
                       torch.empty(k * k + 3, **f))
        phi_s, rho_s, eff_s, out = st[key]
        phi_s.copy_(phi)
        rho_s.copy_(rho_out)
        if eff_s is not None:
            eff_s.copy_(effective_in)
        self._replay(key, lambda: self._energy_body(phi_s, rho_s, eff_s, out))
        return out.clone()

    def _energy_body(self, phi, rho_out, effective_in, out) -> None:
        g, q = self.grid, self.q
        k = phi.shape[0]
        kin = self.bare.kinetic_apply(phi)
        self.bare.measure.cross(phi, kin, out=out[: k * k].view(k, k))
        grad = g.gradient(rho_out)[0]  # (3, N)
        stacked = torch.empty((4, g.n_points), dtype=torch.float64, device=g.device)
        stacked[0].copy_(rho_out)
        stacked[1:].copy_(grad)
        nodes = q.interpolate(stacked)  # rho, d/dx, d/dy, d/dz at the nodes
        m = q.n_nodes
        cross_node = ops.ewise(ops.EW_DOT3, torch.empty(m, dtype=torch.float64, device=g.device), q.f2, nodes[1:], q.grad_u)
        cross_grid = ops.ewise(ops.EW_DOT3, g.empty(), self.factor.weight, grad, self.factor.grad_u)
        pot_node = ops.ewise(ops.EW_MUL3, torch.empty(m, dtype=torch.float64, device=g.device), q.f2, nodes[0],
                             q.transformed_potential)
        pot_grid = ops.ewise(ops.EW_MUL3, g.empty(), self.factor.weight, rho_out, self.factor.transformed_potential)
        # node parts first, then the grid parts accumulated onto them (the order of the separate sums)
        ops.reduce_dots([(q.weights, cross_node, None, k * k, 0.5, False),
                         (q.weights, pot_node, None, k * k + 1, 1.0, False)], out)
        grid_parts = [(q.far_weights, cross_grid, None, k * k, 0.5, True),
                      (q.far_weights, pot_grid, None, k * k + 1, 1.0, True)]
        if effective_in is not None:
            grid_parts.append((effective_in, q.mass_weights, rho_out, k * k + 2, 1.0, False))
        ops.reduce_dots(grid_parts, out)


def energies_from_host(k, occ, evals, host_scalars, e_in, e_out, ion_ion):
    """Energy breakdown from host floats (reference ``scf/step.py:398-461``). ``e_in``/``e_out`` are
    ``[E_H, E_xc, int n v_xc]`` of the input and output densities (``e_out`` None: Harris-Foulkes)."""
    form = np.asarray(host_scalars[: k * k]).reshape(k, k)
    first = float((occ[:k] * np.diag(form)).sum())
    cross, potential, dc_out = host_scalars[k * k], host_scalars[k * k + 1], host_scalars[k * k + 2]
    kinetic = first - cross - potential
    band = float((occ * evals).sum())
    kpe = band - dc_out
    harris = band - e_in[0] - e_in[2] + e_in[1] + ion_ion
    if e_out is None:
        total, hartree, xc = harris, e_in[0], e_in[1]
    else:
        hartree, xc = e_out[0], e_out[1]
        total = kpe + hartree + xc + ion_ion
    external = total - kinetic - hartree - xc - ion_ion
    return {"total": total, "kinetic": kinetic, "external": external, "hartree": hartree, "xc": xc,
            "harris_foulkes": harris, "band": band, "external_direct": cross + 2.0 * potential,
            "kinetic_plus_external": kpe, "ion_ion": ion_ion}


class PulayMixer:
    """Periodic Pulay/DIIS with Tikhonov regularisation (reference ``scf/mixing.py:86-162``), all
    on the device: the history lives in two ring buffers of ``history`` slots, the residual Gram
    matrix is one reduction over the whole ring, the bordered system is solved by a one-thread
    kernel and the combination reads its coefficients from device memory, so a mixing step never
    waits for the GPU."""

    def __init__(self, alpha, history, period, inner_weights: torch.Tensor):
        self.alpha, self.history, self.period = alpha, history, period
        if not 1 <= history <= 8:
            raise ValueError("history must be 1..8 (the combination takes at most 16 terms)")
        self.w = inner_weights
        self.inputs = self.residuals = None  # (history, N) rings
        self.order: list[int] = []  # active slots, oldest first
        self.coef = None
        self.name = f"periodic_pulay(alpha={alpha:g}, history={history}, period={period})"

    def mix(self, n_in: torch.Tensor, n_out: torch.Tensor, iteration: int) -> torch.Tensor:
        n = n_in.numel()
        if self.inputs is None:
            f = dict(dtype=torch.float64, device=n_in.device)
            self.inputs = torch.zeros((self.history, n), **f)
            self.residuals = torch.zeros((self.history, n), **f)
            self.coef = torch.zeros(2 * self.history, **f)
        slot = len(self.order) if len(self.order) < self.history else self.order.pop(0)
        self.order.append(slot)
        self.inputs[slot].copy_(n_in.reshape(-1))
        residual = self.residuals[slot]
        ops.ewise(ops.EW_SUB, residual, n_out.reshape(-1), n_in.reshape(-1))
        m = len(self.order)
        out = torch.empty_like(n_in)
        if m < 2 or iteration % self.period != 0:
            ops.lincomb(out.view(-1), [(n_in.reshape(-1), 1.0), (residual, self.alpha)])
            return out
        gram = ops.wgram(self.residuals, None, self.w)
        ops.pulay_coefficients(gram, self.order, self.alpha, self.coef)
        terms = []
        for i, s in enumerate(self.order):
            terms.append((self.inputs[s], 1.0, 1, 2 * i))
            terms.append((self.residuals[s], 1.0, 1, 2 * i + 1))
        ops.lincomb(out.view(-1), terms, self.coef)
        return out


class LinearMixer:
    def __init__(self, alpha):
        self.alpha = alpha
        self.name = f"linear(alpha={alpha:g})"

    def mix(self, n_in, n_out, iteration):
        residual = torch.empty_like(n_in)
        ops.ewise(ops.EW_SUB, residual.view(-1), n_out.reshape(-1), n_in.reshape(-1))
        out = torch.empty_like(n_in)
        ops.lincomb(out.view(-1), [(n_in.reshape(-1), 1.0), (residual.view(-1), self.alpha)])
        return out


class BoundReusePolicy:
    """When an SCF iteration may filter with the previous iteration's bounds (reference
    ``scf/loop.py:94-170``)."""

    def __init__(self):
        self.bounds = None
        self.since = 0
        self.widening = 0.0
        self.top_fraction = float("nan")
        self.history: list[str] = []

    def decide(self, residual_rose: bool, rung_changed: bool):
        if self.bounds is None:
            return "no carried interval"
        if rung_changed:
            return "fallback rung taken"
        if residual_rose:
            return "SCF residual rose"
        if self.since >= LANCZOS_REUSE_BACKSTOP:
            return f"backstop ({LANCZOS_REUSE_BACKSTOP} iterations)"
        lo, hi = self.bounds
        if self.widening > LANCZOS_REUSE_MAX_WIDENING * (hi - lo):
            return "accumulated shift"
        if not self.top_fraction < BOUND_REUSE_RITZ_FRACTION:
            return "top Ritz fraction near the upper bound"
        return None

    def observe(self, reason, info):
        self.bounds = tuple(info["spectral_bounds"])
        self.top_fraction = float(info["top_ritz_fraction"])
        if info["bounds_source"] == "reused":
            self.since += 1
            self.widening += float(info["bounds_shift"])
            self.history.append("reused")
        else:
            self.since = 0
            self.widening = 0.0
            self.history.append(f"lanczos: {reason}")


def build_geometry(scenario, numerics, device, anchor=None):
    grid = UniformGrid.from_box(numerics.box_lengths, numerics.spacing, device, anchor=anchor, fd_order=numerics.fd_order)
    factor = CuspFactor(grid, scenario.charges, scenario.positions)
    quadrature = CuspQuadrature(grid, factor)
    return grid, factor, quadrature


#: The starting densities of the SCF and their exponent scales (reference ``scf/loop.py:68,181``):
#: three genuinely different starting points for the initial-guess gate.
INITIAL_GUESSES = {"hydrogenic": 1.0, "diffuse": 0.5, "compact": 2.0}


def initial_density(step: KohnShamStep, kind: str = "hydrogenic") -> torch.Tensor:
    """Superposition of ``(zeta^3/pi) exp(-2 zeta s_a)``, ``zeta = Z_a`` times 1 (hydrogenic), 0.5
    (diffuse) or 2 (compact), normalised in the cusp-aware measure (reference ``scf/loop.py:173-193``).
    The hydrogenic fields are the step's own; the other two are evaluated for the scaled charges."""
    if kind not in INITIAL_GUESSES:
        raise ValueError(f"unknown initial guess {kind!r}; choose from {tuple(INITIAL_GUESSES)}")
    g = step.grid
    zetas = [z * INITIAL_GUESSES[kind] for z in step.charges]
    fields = step.cusp_grid
    if kind != "hydrogenic":
        na = len(zetas)
        fields = torch.empty((na, g.n_points), dtype=torch.float64, device=g.device)
        scaled = (ctypes.c_double * na)(*zetas)
        _lib.call("cdft_hydrogenic_fields_f64", None, g.n_points, g.c_shape, g.spacing, g.c_origin, na,
                  ctypes.addressof(scaled), step.factor.c_positions, fields.data_ptr(), None, stream_of(fields))
    coefs = torch.tensor([z**3 / math.pi for z in zetas], dtype=torch.float64, device=g.device)
    total = ops.atom_combine(torch.zeros(g.n_points, dtype=torch.float64, device=g.device), coefs, fields, +1.0,
                             g.empty())
    rho = ops.ewise(ops.EW_DIV, g.empty(), total, step.f2)
    charge = torch.empty(1, dtype=torch.float64, device=g.device)
    ops.reduce(ops.RED_DOT, step.q.mass_weights, charge, 0, b=rho)
    scale = step.n_electrons / float(charge.item())  # setup: one read
    return ops.ewise(ops.EW_SCALE, g.empty(), total, s=scale).reshape(1, -1)


#: Float64 filter passes that finish a mixed-precision solve.
REFINEMENT_PASSES = 2

#: Subspace solvers kept between SCF solves (least recently used first out), by grid shape, device,
#: stream, eigensolver settings and filter precision: a later solve of the same kind reuses their
#: buffers and captured graphs, and copies its own geometry into their static arrays.
SOLVER_CACHE_SIZE = 2
_solvers: dict = {}


def subspace_solver(grid, config: EigenConfig, filter_dtype, filter_hook) -> ChebyshevFilteredSubspace:
    """The SCF's subspace solver: a cached one of the same kind, started afresh, or a new one. The
    cache is keyed by the caller's stream (:func:`ops.workspace_key`), so solves that run at once on
    different streams (one host thread each) never share a solver or its graphs' scratch buffers."""
    key = (grid.shape, ops.workspace_key(grid.device), repr(config), filter_dtype)
    solver = _solvers.pop(key, None)
    if solver is None:
        solver = ChebyshevFilteredSubspace(grid, config, ritz_stop=False, filter_dtype=filter_dtype,
                                           filter_hook=filter_hook)
    else:
        solver.begin_solve(grid, config, filter_hook)
    _solvers[key] = solver
    while len(_solvers) > SOLVER_CACHE_SIZE:
        _solvers.pop(next(iter(_solvers)))
    return solver


def two_solver_check(hamiltonian, grid, config, n_states: int, seed: int, first_eigenvalues) -> dict:
    """LOBPCG on the same operator, seeded with ``seed + 1``, against the CheFSI eigenvalues
    ``first_eigenvalues`` (reference ``scf/loop.py:485-494``, ``scf/noninteracting.py:182-189``; gate
    G2.5 holds the largest difference to ``cross_check_tol``)."""
    from .lobpcg import LOBPCG

    second = LOBPCG(grid, config).solve(hamiltonian, n_states, generator=torch.Generator(device="cpu").manual_seed(seed + 1))
    first = np.asarray(first_eigenvalues, dtype=np.float64)[:n_states]
    delta = np.abs(first - second.eigenvalues[: first.shape[0]])
    return {"two_solver_max_delta": float(delta.max()), "two_solver_converged": bool(second.converged),
            "lobpcg_iterations": float(second.n_iterations), "lobpcg_eigenvalues": second.eigenvalues.tolist()}


def solve_self_consistent(scenario, numerics, device, anchor=None, geometry=None, filter_dtype=torch.float64,
                          filter_hook=None, cross_check: bool | None = None, initial_guess: str = "hydrogenic"):
    """The SCF driver (reference ``scf/loop.py:255-460``). ``filter_dtype=float32`` runs the Chebyshev
    filter of every warm SCF pass in float32 (and the cold first solve until its residual is small),
    and the final diagonalisation in float64 with at least ``REFINEMENT_PASSES`` passes; Rayleigh-Ritz,
    orthonormalisation, densities, Poisson, XC and energies stay float64 throughout. ``cross_check``
    (default ``numerics.eigen.cross_check``) also solves the final operator by LOBPCG and by a second
    CheFSI diagonalisation from the converged vectors, and reports their largest eigenvalue difference
    in ``measurements`` (reference ``scf/loop.py:485-494``). ``initial_guess`` names the starting
    density (``INITIAL_GUESSES``)."""
    if initial_guess not in INITIAL_GUESSES:
        raise ValueError(f"unknown initial guess {initial_guess!r}; choose from {tuple(INITIAL_GUESSES)}")
    started = time.perf_counter()
    store: dict = {}
    if geometry is None:
        # The start vectors a solve draws are known before the geometry exists: the cold block then
        # its Lanczos start, and in a warm solve the vectors that pad the previous eigenvectors to
        # the block then (when the bounds are not reused) the Lanczos start. Draw them on host
        # threads meanwhile.
        counts = [int(round(length / numerics.spacing)) + 1 for length in numerics.box_lengths]
        npts = counts[0] * counts[1] * counts[2]
        n_states = max(1, int(-(-float(scenario.n_electrons) // 2.0)))
        cfg = numerics.eigen
        k = min(n_states + max(cfg.n_extra_states, int(cfg.extra_states_fraction * n_states)), npts)
        warm = ((k - n_states, npts), (npts,)) if k > n_states else ((npts,),)
        prefetch_draws(store, numerics.seed, [((k, npts), (npts,)), warm])
    grid, factor, q = geometry if geometry is not None else build_geometry(scenario, numerics, device, anchor)
    step = KohnShamStep(scenario, numerics, grid, factor, q)
    step.draws = store
    g = grid
    inner_weights = ops.ewise(ops.EW_DIV, g.empty(), q.mass_weights, ops.ewise(ops.EW_MUL, g.empty(), step.f2, step.f2))
    solver = subspace_solver(grid, numerics.eigen, filter_dtype, filter_hook)
    n_in = initial_density(step, initial_guess)
    mixer = PulayMixer(numerics.alpha, numerics.history, numerics.pulay_period, inner_weights)
    ladder = [
        lambda: PulayMixer(numerics.alpha, numerics.history, numerics.pulay_period, inner_weights),
        lambda: PulayMixer(0.5 * numerics.alpha, numerics.history, numerics.pulay_period, inner_weights),
        lambda: LinearMixer(0.5 * numerics.alpha),
    ]
    rung = 0
    energies, residual_history, evals_history, events = [], [], [], []
    worsened = 0
    previous_residual = float("inf")
    initial_block = None
    policy = BoundReusePolicy()
    previous_effective = None
    rung_changed = False
    converged = False
    applications = []
    iteration = 0
    norm_buf = torch.empty(3, dtype=torch.float64, device=g.device)
    tmp = g.empty()
    pin_next = None  # the next iteration's input potentials, enqueued before this one's read
    speculated = {"used": 0, "discarded": 0}
    for iteration in range(1, numerics.max_scf + 1):
        pin = pin_next if pin_next is not None else step.potentials(n_in)
        pin_next = None
        h = step.hamiltonian(pin)
        steps = None if initial_block is None else FILTER_STEPS_PER_ITERATION
        reason = policy.decide(worsened > 0, rung_changed)
        hint = None
        if reason is None and previous_effective is not None:
            ops.reduce(ops.RED_MAXDIFF, pin["effective"], norm_buf, 2, b=previous_effective)
            hint = SpectralBoundsHint(policy.bounds[0], policy.bounds[1], DeviceScalar(norm_buf[2:3]))
        elif reason is None:
            reason = "no previous potential"
        # Mixed precision: the warm passes of an iteration filter in float32; the first (cold) solve
        # switches to float64 once its residual is small.
        eigen = step.diagonalise(solver, h, initial=initial_block, max_iterations=steps, bounds=hint,
                                 precision="auto" if steps is None else "f32")
        policy.observe(reason, eigen.info)
        previous_effective = pin["effective"]
        rung_changed = False
        applications.append(h.apply_count)
        initial_block = eigen.eigenvectors
        occ = aufbau(eigen.eigenvalues, step.n_electrons)
        n_out, rho_out = step.output_density(eigen, occ)
        pout = step.potentials(n_out)
        escal = step.energy_scalars(eigen.eigenvectors, occ, rho_out, pin["effective"])
        residual = ops.ewise(ops.EW_SUB, g.empty(), n_out[0], n_in[0])
        ops.ewise(ops.EW_DIV, tmp, residual, step.f2)
        ops.reduce(ops.RED_L1, tmp, norm_buf, 0, w=q.mass_weights)
        n_next = mixer.mix(n_in, n_out, iteration)
        ops.ewise(ops.EW_SUB, tmp, n_next[0], n_in[0])
        ops.ewise(ops.EW_DIV, tmp, tmp, step.f2)
        ops.reduce(ops.RED_L1, tmp, norm_buf, 1, w=q.mass_weights)
        read = DeviceArray(torch.cat([escal, pin["scalars"], pout["scalars"], norm_buf[:2]]))  # one read
        # The next iteration's potentials of the mixed density go on the stream before the read;
        # they are dropped if the solve stops here or the mixing rung changes.
        if iteration < numerics.max_scf:
            pin_next = step.potentials(n_next)
        host = read.numpy()
        k = eigen.eigenvectors.shape[0]
        ne = k * k + 3
        e_in, e_out = host[ne:ne + 3], host[ne + 3:ne + 6]
        terms = energies_from_host(k, occ, eigen.eigenvalues, host[:ne], e_in, e_out, step.ion_ion)
        res_norm, change = float(host[ne + 6]), float(host[ne + 7])
        energies.append(terms["total"])
        residual_history.append(res_norm)
        evals_history.append(eigen.eigenvalues.copy())
        de = abs(energies[-1] - energies[-2]) if len(energies) > 1 else float("inf")
        if iteration >= numerics.min_scf and de < numerics.energy_tol and res_norm < numerics.density_tol:
            converged = True
            speculated["discarded"] += pin_next is not None
            break
        worsened = worsened + 1 if res_norm > previous_residual else 0
        previous_residual = res_norm
        if worsened >= 3 and rung + 1 < len(ladder):
            rung += 1
            mixer = ladder[rung]()
            worsened = 0
            rung_changed = True
            events.append(f"iteration {iteration}: residual grew three times, rung {rung}")
            n_next = torch.empty_like(n_in)
            ops.lincomb(n_next.view(-1), [(n_in.reshape(-1), 1.0), (residual, 0.5 * numerics.alpha)])
            speculated["discarded"] += pin_next is not None
            pin_next = None
        speculated["used"] += pin_next is not None
        n_in = n_next
    h_final = step.hamiltonian(pin)
    # The final diagonalisation is float64 throughout; a mixed solve refines its float32-filtered
    # vectors with at least REFINEMENT_PASSES float64 filter passes.
    eigen = step.diagonalise(solver, h_final, initial=eigen.eigenvectors, precision="f64",
                             min_fp64_passes=REFINEMENT_PASSES if solver.mixed else 0)
    occ = aufbau(eigen.eigenvalues, step.n_electrons)
    n_out, rho_out = step.output_density(eigen, occ)
    pout = step.potentials(n_out)
    escal = step.energy_scalars(eigen.eigenvectors, occ, rho_out, pin["effective"])
    host = torch.cat([escal, pin["scalars"], pout["scalars"]]).cpu().numpy()
    k = eigen.eigenvectors.shape[0]
    ne = k * k + 3
    terms = energies_from_host(k, occ, eigen.eigenvalues, host[:ne], host[ne:ne + 3], host[ne + 3:ne + 6], step.ion_ion)
    vectors = eigen.eigenvectors
    measurements: dict = {}
    do_cross = getattr(numerics.eigen, "cross_check", False) if cross_check is None else cross_check
    if do_cross:
        vectors = vectors.clone()  # the cached solver's buffers are reused by the second diagonalisation
        first = step.diagonalise(solver, h_final, initial=vectors, precision="f64")
        measurements.update(two_solver_check(h_final, grid, numerics.eigen, step.n_states, numerics.seed,
                                             first.eigenvalues))
    return {
        "energies": terms,
        "eigenvalues": eigen.eigenvalues,
        "eigenvectors": vectors,
        "measurements": measurements,
        "density": n_out,
        "scf_iterations": iteration,
        "converged": converged and eigen.converged,
        "energy_history": energies,
        "residual_history": residual_history,
        "bound_history": policy.history,
        "applications": applications,
        "fallback_events": events,
        "filter_precisions": list(solver.filter_precisions),
        "graphs": dict(solver.graph_record),
        "speculative_filters": dict(solver.speculation),
        "speculative_potentials": speculated,
        "wall_seconds": time.perf_counter() - started,
        "step": step,
        "effective_potential": pin["effective"],
    }


def solve_noninteracting(scenario, numerics, device, anchor=None, geometry=None, cross_check: bool | None = None,
                         n_states: int | None = None):
    """Fixed-potential eigenproblem (reference ``scf/noninteracting.py:118-330``): CheFSI with the
    weak-form Ritz stop, then ``T_s`` and ``E_ext = band - T``. ``cross_check`` (default
    ``numerics.eigen.cross_check``) also solves the operator by LOBPCG (``noninteracting.py:182-189``).
    ``n_states`` asks for more states than the electrons occupy, never fewer (``noninteracting.py:148-152``)."""
    grid, factor, q = geometry if geometry is not None else build_geometry(scenario, numerics, device, anchor)
    h = CuspFactoredHamiltonian(grid, factor, q)
    needed = max(1, int(-(-float(scenario.n_electrons) // 2.0)))
    n_states = needed if n_states is None else max(int(n_states), needed)
    gen = torch.Generator(device="cpu").manual_seed(numerics.seed)
    solver = ChebyshevFilteredSubspace(grid, numerics.eigen, ritz_stop=True)
    result = solver.solve(h, n_states, generator=gen)
    occ = aufbau(result.eigenvalues, scenario.n_electrons)
    step = KohnShamStep.__new__(KohnShamStep)
    step.grid, step.factor, step.q, step.bare = grid, factor, q, h
    step.use_graphs, step.graph_record = False, {"captures": 0, "replays": 0, "disabled": "one evaluation"}
    rho = grid.empty()
    ops.density(result.eigenvectors, occ, None, rho, None)
    escal = KohnShamStep.energy_scalars(step, result.eigenvectors, occ, rho, None).cpu().numpy()
    k = result.eigenvectors.shape[0]
    form = escal[: k * k].reshape(k, k)
    kinetic = float((occ[:k] * np.diag(form)).sum()) - escal[k * k] - escal[k * k + 1]
    band = float((occ * result.eigenvalues).sum())
    ion = nuclear_repulsion(scenario.charges, scenario.positions)
    measurements: dict = {}
    if getattr(numerics.eigen, "cross_check", False) if cross_check is None else cross_check:
        measurements.update(two_solver_check(h, grid, numerics.eigen, n_states, numerics.seed, result.eigenvalues))
    return {"eigenvalues": result.eigenvalues, "eigenvectors": result.eigenvectors, "n_iterations": result.n_iterations,
            "converged": result.converged, "kinetic": kinetic, "band": band, "total": band + ion, "info": result.info,
            "hamiltonian": h, "measurements": measurements}


@dataclass
class Run:
    """A completed solve as the gates read it (the fields of the reference's ``RunArtifact`` that
    ``gates/tier3.py`` and ``gates/tier5.py`` use): the scenario, the numerics with the grid the run
    used, its lattice anchor, the device, the eigenvalues, the total energy and the raw result."""

    scenario: object
    numerics: object
    anchor: tuple | None
    device: torch.device
    eigenvalues: np.ndarray
    total: float
    converged: bool
    n_iterations: int
    initial_guess: str
    wall_s: float
    result: dict


def solve_scenario(scenario, numerics=None, device=None, *, derive_grid: bool = True, n_states: int | None = None,
                   cross_check: bool = False, initial_guess: str = "hydrogenic") -> Run:
    """Solve one scenario by the path it needs (reference ``scf/solve.py:32-64``): the SCF for an
    interacting one, the fixed-potential eigenproblem otherwise. ``derive_grid`` replaces the grid of
    ``numerics`` by the one the reference derives from the nuclei, anchored on the first nucleus
    (``grid.py:294-457``); ``derive_grid=False`` solves on ``numerics``' own spacing and box with the
    coordinate origin on the lattice, as the convergence gates need (``grid.py:448-457``)."""
    from .fingerprint import derived_numerics  # the reference's grid rule lives beside the scenarios

    device = torch.device("cuda") if device is None else torch.device(device)
    if derive_grid:
        numerics, anchor = derived_numerics(scenario, numerics)
    elif numerics is None:
        raise ValueError("derive_grid=False needs the numerics whose grid it solves on")
    else:
        anchor = None
    started = time.perf_counter()
    if scenario.interacting:
        if n_states is not None:
            raise ValueError("the SCF solves for the occupied states; n_states is for the fixed-potential path")
        r = solve_self_consistent(scenario, numerics, device, anchor=anchor, cross_check=cross_check,
                                  initial_guess=initial_guess)
        total, iterations = float(r["energies"]["total"]), int(r["scf_iterations"])
    else:
        if initial_guess != "hydrogenic":
            raise ValueError("a fixed-potential solve has no starting density")
        r = solve_noninteracting(scenario, numerics, device, anchor=anchor, cross_check=cross_check,
                                 n_states=n_states)
        total, iterations = float(r["total"]), int(r["n_iterations"])
    return Run(scenario, numerics, anchor, device, np.asarray(r["eigenvalues"], dtype=np.float64).copy(), total,
               bool(r["converged"]), iterations, initial_guess, time.perf_counter() - started, r)
