"""Configuration and scenario types of the hot path, with the reference's names and defaults.

Mirrors the subset of the reference's frozen contract that the Kohn-Sham path reads
(``contract.py:352-529``: ``GridConfig``, ``EigenConfig``, ``MixingConfig``, ``SCFConfig``,
``NumericsConfig``) and the all-electron preset ``ALL_ELECTRON_NUMERICS`` (``config.py:218-232``),
flattened into one ``Numerics`` record, plus the scenarios of ``physics_config.py`` the
benchmark configurations use.
"""

from __future__ import annotations

from dataclasses import dataclass, field, replace

from .eigen import EigenConfig


@dataclass
class Numerics:
    """How hard to try. Defaults are ``ALL_ELECTRON_NUMERICS`` with the SCF and mixing defaults of
    the contract: order-8 stencils, box domain with a zero halo, Poisson pad 2.5, CheFSI with 4 extra
    states and degree 20 to residual 1e-8, periodic Pulay (alpha 0.3, history 8, period 3), SCF
    tolerances 1e-8 Ha and 1e-6 electrons, seed 0."""

    box_lengths: tuple = (32.0, 32.0, 32.0)
    spacing: float = 1.0
    fd_order: int = 8
    poisson_pad_factor: float = 2.5
    eigen: EigenConfig = field(default_factory=EigenConfig)
    alpha: float = 0.30
    history: int = 8
    pulay_period: int = 3
    energy_tol: float = 1.0e-8
    density_tol: float = 1.0e-6
    max_scf: int = 200
    min_scf: int = 3
    seed: int = 0

    def with_grid(self, spacing: float, box: float | tuple) -> "Numerics":
        lengths = box if isinstance(box, tuple) else (box, box, box)
        return replace(self, spacing=spacing, box_lengths=lengths)


@dataclass
class Scenario:
    """Nuclei (bohr) and charges, electron count, functional name (``"none"`` or ``"lda"``)."""

    scenario_id: str
    charges: tuple
    positions: tuple
    n_electrons: float
    xc: str = "none"
    n_spin: int = 1

    @property
    def interacting(self) -> bool:
        return self.xc != "none"


ALL_ELECTRON_NUMERICS = Numerics()


def hydrogen_atom() -> Scenario:
    return Scenario("h_atom", (1.0,), ((0.0, 0.0, 0.0),), 1.0)


def helium_lda() -> Scenario:
    return Scenario("he_atom_lda", (2.0,), ((0.0, 0.0, 0.0),), 2.0, "lda")


def hydrogen_molecule_lda(bond: float = 1.4, axis: int = 0) -> Scenario:
    """H2 with the bond along ``axis`` (x for the benchmark configuration), centred at the origin."""
    a = [0.0, 0.0, 0.0]
    b = [0.0, 0.0, 0.0]
    a[axis], b[axis] = -0.5 * bond, 0.5 * bond
    return Scenario(f"h2_R{bond:g}_lda", (1.0, 1.0), (tuple(a), tuple(b)), 2.0, "lda")


def hydrogen_molecular_ion(bond: float = 2.0, axis: int = 0) -> Scenario:
    a = [0.0, 0.0, 0.0]
    b = [0.0, 0.0, 0.0]
    a[axis], b[axis] = -0.5 * bond, 0.5 * bond
    return Scenario(f"h2plus_R{bond:g}", (1.0, 1.0), (tuple(a), tuple(b)), 1.0)
