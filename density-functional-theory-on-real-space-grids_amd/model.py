"""The analytic model systems of the reference's registry on the device: the isotropic harmonic well
(``harmonic_w1``, omega = 1) and the 10-bohr infinite well (``box_L10``), solved with the plain
operator ``H = -1/2 lap + V`` (reference ``operators/hamiltonian.py:92-125``) by the same CheFSI as
the Coulomb scenarios (``scf/noninteracting.py:59-110,118-330`` takes the plain operator when there is
no nucleus to factor).

Settings are the reference's ``MODEL_SYSTEM_NUMERICS`` (``config.py:191-215``): h = 0.2, order 8, the
16-bohr box with a zero halo for the harmonic well (81^3 points from -8), k = n + 6 states, filter
degree 16, residual tolerance 1e-8, at most 200 passes. The infinite well's grid holds the 49^3 points
strictly inside the walls at +-5 bohr (origin -4.8), its walls an odd reflection of the halo
(``grid.py:808-827``) and its potential zero (``operators/external.py:59-63``). The uniform measure
``h^3`` is the inner product (``eigen/measure.py``). The energies are recorded as the reference's
non-interacting path records them: ``T = sum_i f_i <psi_i, -1/2 lap psi_i>``, the band energy
``sum_i f_i e_i`` as the total and the Harris-Foulkes energy, and ``E_ext = band - T``.
"""

from __future__ import annotations

import dataclasses
from dataclasses import dataclass

import numpy as np
import torch

from . import ops
from .eigen import ChebyshevFilteredSubspace, EigenConfig
from .geometry import UniformGrid
from .hamiltonian import DeviceMeasure
from .scf import aufbau


@dataclass(frozen=True)
class ModelSystem:
    scenario_id: str
    kind: str  # "harmonic" or "well"
    shape: tuple[int, int, int]
    spacing: float
    origin: tuple[float, float, float]
    boundary: int  # 0 = zero halo, 1 = odd reflection
    omega: float = 0.0
    n_electrons: float = 1.0


#: The two model systems (reference ``config.py:342-347``, grids as the reference's fingerprint records them).
MODELS: dict[str, ModelSystem] = {
    "harmonic_w1": ModelSystem("harmonic_w1", "harmonic", (81, 81, 81), 0.2, (-8.0, -8.0, -8.0), 0, omega=1.0),
    "box_L10": ModelSystem("box_L10", "well", (49, 49, 49), 0.2, (-4.8, -4.8, -4.8), 1),
}

#: ``MODEL_SYSTEM_NUMERICS.eigen`` (reference ``config.py:201-203``); the seed is the default 0.
MODEL_EIGEN = dataclasses.replace(EigenConfig(), n_extra_states=6, chebyshev_degree=16, max_iterations=200,
                                  residual_tol=1.0e-8)
FD_ORDER = 8
SEED = 0


def model_grid(system: ModelSystem, device) -> UniformGrid:
    return UniformGrid(system.shape, system.spacing, system.origin, device, fd_order=FD_ORDER, gradient_order=FD_ORDER)


def model_potential(system: ModelSystem, grid: UniformGrid) -> torch.Tensor:
    """``1/2 omega^2 r^2`` for the harmonic well (``external.py:51-53``: the squared coordinates summed
    x + y + z, then scaled), zero inside the infinite well."""
    if system.kind == "well":
        return torch.zeros(grid.n_points, dtype=torch.float64, device=grid.device)
    axes = [torch.arange(n, dtype=torch.float64) * grid.spacing + o for n, o in zip(grid.shape, grid.origin)]
    x, y, z = torch.meshgrid(*axes, indexing="ij")
    r_sq = (x * x + y * y) + z * z
    return (0.5 * system.omega**2 * r_sq).reshape(-1).to(grid.device)


class LocalHamiltonian:
    """``H = -1/2 lap + V`` on the device with the uniform measure (``operators/hamiltonian.py:92-125``)."""

    uses_divergence_form = False

    def __init__(self, grid: UniformGrid, potential: torch.Tensor, boundary: int, order: int = FD_ORDER):
        self.grid = grid
        self.potential = potential
        self.geometry = ops.LocalGeometry(grid.shape, order, boundary, ops.local_coefficients(order, grid.spacing),
                                          potential)
        self.measure = DeviceMeasure(grid, torch.full((grid.n_points,), grid.volume_element, dtype=torch.float64,
                                                      device=grid.device))
        self.apply_count = 0
        self._kinetic = None

    def apply(self, phi: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        self.apply_count += 1
        return ops.hdiv_apply(phi, self.geometry, out)

    def kinetic_apply(self, phi: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        if self._kinetic is None:
            self._kinetic = self.geometry.with_potential(torch.zeros_like(self.potential))
        return ops.hdiv_apply(phi, self._kinetic, out)

    def note_applications(self, count: int) -> None:
        self.apply_count += int(count)


def solve_model(scenario_id: str, device, n_states: int | None = None) -> dict:
    """The fixed-potential solve of one model system: eigenvalues, vectors, iteration count and the
    energy terms of the reference's record."""
    system = MODELS[scenario_id]
    grid = model_grid(system, device)
    ham = LocalHamiltonian(grid, model_potential(system, grid), system.boundary)
    needed = max(1, int(-(-system.n_electrons // 2.0)))
    n_states = needed if n_states is None else max(int(n_states), needed)
    gen = torch.Generator(device="cpu").manual_seed(SEED)
    solver = ChebyshevFilteredSubspace(grid, MODEL_EIGEN, ritz_stop=True)
    result = solver.solve(ham, n_states, generator=gen)
    evals = np.asarray(result.eigenvalues, dtype=np.float64)
    occ = aufbau(evals, system.n_electrons)
    vecs = result.eigenvectors
    forms = (vecs * ham.kinetic_apply(vecs)).sum(dim=1) * grid.volume_element
    kinetic = float((occ * forms.cpu().numpy()).sum())
    band = float((occ * evals).sum())
    return {"system": system, "grid": grid, "hamiltonian": ham, "eigenvalues": evals, "eigenvectors": vecs,
            "occupations": occ, "n_iterations": result.n_iterations, "converged": result.converged,
            "kinetic": kinetic, "band": band, "external": band - kinetic, "total": band}
