"""Oracle physics: Coulomb-cutoff Poisson, VWN LDA, the Kohn-Sham step and the SCF / non-interacting
drivers.

Test infrastructure only. Follows ``src/cdft/operators/poisson.py``, ``xc/{base,lda}.py``,
``scf/{step,loop,mixing,occupations,noninteracting}.py`` of the reference in numpy float64.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .core import CuspFactor, CuspHamiltonian, Grid, Measure
from .eigen import EigenConfig, SpectralBoundsHint, chefsi_solve, make_generator, orthonormality_error
from .quadrature import CuspQuadrature, hydrogenic_1s_potential

# --------------------------------------------------------------------------------------------
# Poisson
# --------------------------------------------------------------------------------------------


def fft_friendly(n: int) -> int:
    """Next 5-smooth integer (reference ``operators/poisson.py:110-120``)."""
    c = n
    while True:
        r = c
        for p in (2, 3, 5):
            while r % p == 0:
                r //= p
        if r == 1:
            return c
        c += 1


class CoulombCutoffPoisson:
    """Zero-padded rFFT with ``K(G) = 4 pi (1 - cos G R)/G^2``, ``K(0) = 2 pi R^2`` (reference
    ``operators/poisson.py:27-84``)."""

    def __init__(self, grid: Grid, pad_factor: float = 2.5):
        self.grid = grid
        self.padded_shape = tuple(fft_friendly(int(math.ceil(pad_factor * n))) for n in grid.shape)
        self.cutoff_radius = 0.5 * min(n * grid.spacing for n in self.padded_shape)
        h = grid.spacing
        n0, n1, n2 = self.padded_shape
        two_pi = 2.0 * math.pi
        kx = two_pi * np.fft.fftfreq(n0, d=h)
        ky = two_pi * np.fft.fftfreq(n1, d=h)
        kz = two_pi * np.fft.rfftfreq(n2, d=h)
        g_sq = kx[:, None, None] ** 2 + ky[None, :, None] ** 2 + kz[None, None, :] ** 2
        g = np.sqrt(g_sq)
        r = self.cutoff_radius
        with np.errstate(divide="ignore", invalid="ignore"):
            self.kernel = np.where(g_sq > 0.0, 4.0 * math.pi * (1.0 - np.cos(g * r)) / np.maximum(g_sq, 1e-300), 2.0 * math.pi * r**2)

    def solve(self, density):
        box = self.grid.to_box(density)
        padded = np.zeros(self.padded_shape)
        padded[: box.shape[0], : box.shape[1], : box.shape[2]] = box
        spectrum = np.fft.rfftn(padded, s=self.padded_shape)
        potential = np.fft.irfftn(spectrum * self.kernel, s=self.padded_shape)
        return potential[: box.shape[0], : box.shape[1], : box.shape[2]].reshape(-1)


# --------------------------------------------------------------------------------------------
# LDA: Slater exchange + VWN5 correlation, spin-restricted
# --------------------------------------------------------------------------------------------

DENSITY_THRESHOLD = 1.0e-14
_SLATER = 0.75 * (3.0 / math.pi) ** (1.0 / 3.0)
_VWN5_PARA = (0.0310907, -0.10498, 3.72744, 12.9352)


def _vwn_para(x):
    """VWN eq. (4.4) paramagnetic fit and its ``x``-derivative (reference ``xc/lda.py:122-134``)."""
    a, x0, b, c = _VWN5_PARA
    big_x = x * x + b * x + c
    big_x0 = x0 * x0 + b * x0 + c
    q = math.sqrt(4.0 * c - b * b)
    at = np.arctan(q / (2.0 * x + b))
    value = a * (np.log(x * x / big_x) + (2.0 * b / q) * at - (b * x0 / big_x0) * (np.log((x - x0) ** 2 / big_x) + (2.0 * (b + 2.0 * x0) / q) * at))
    deriv = a * (2.0 / x - (2.0 * x + 2.0 * b) / big_x - (b * x0 / big_x0) * (2.0 / (x - x0) - (2.0 * x + 2.0 * b + 2.0 * x0) / big_x))
    return value, deriv


def lda_vwn(density):
    """``(e_xc, v_xc)`` per volume for a spin-restricted total density ``(N,)``.

    Energy as ``SemiLocalFunctional._energy_from_contract_inputs`` + ``LDA(VWN5)`` (reference
    ``xc/base.py:171-210``, ``xc/lda.py:137-159,201-217``) at ``zeta = 0``; the potential is the
    analytic derivative (the reference differentiates by autograd; the spin terms vanish
    identically at ``zeta = 0``).
    """
    n = np.asarray(density, dtype=np.float64)
    half = 0.5 * n
    total = half + half
    empty = total < DENSITY_THRESHOLD
    floor = 0.5 * DENSITY_THRESHOLD
    ns = np.maximum(half, floor)
    factor = 2.0 ** (1.0 / 3.0) * _SLATER
    ex = -factor * (ns ** (4.0 / 3.0) + ns ** (4.0 / 3.0))
    vx = -factor * (4.0 / 3.0) * ns ** (1.0 / 3.0)
    nt = ns + ns
    rs = (3.0 / (4.0 * math.pi * nt)) ** (1.0 / 3.0)
    x = np.sqrt(rs)
    para, dpara = _vwn_para(x)
    ec = nt * para
    vc = para - (x / 6.0) * dpara
    e = np.where(empty, 0.0, ex + ec)
    v = np.where(empty, 0.0, vx + vc)
    return e, v


_KF = (3.0 * math.pi**2) ** (1.0 / 3.0)
_PBE_KAPPA = 0.804
_PBE_MU = 0.2195149727645171
_PBE_BETA = 0.06672455060314922
_PBE_GAMMA = (1.0 - math.log(2.0)) / math.pi**2
_PW92_MOD_PARA = (0.0310907, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294)
_SIGMA_THRESHOLD = 1.0e-40


def _pbe_energy(n_up, n_dn, s_uu, s_ud, s_dd, torch):
    """Spin-resolved PBE energy density per volume (reference ``xc/gga.py:30-90``, ``xc/lda.py:64-108``,
    ``xc/base.py:198-235``), written with torch operations so that autograd gives the derivatives."""
    total = torch.zeros_like(n_up)
    for n_s, sig_s in ((n_up, s_uu), (n_dn, s_dd)):  # exchange, by spin scaling
        n2 = 2.0 * n_s
        s2 = (4.0 * sig_s) / (4.0 * _KF**2 * n2 ** (8.0 / 3.0))
        fx = 1.0 + _PBE_KAPPA - _PBE_KAPPA / (1.0 + _PBE_MU * s2 / _PBE_KAPPA)
        total = total + 0.5 * (-_SLATER * n2 ** (4.0 / 3.0)) * fx
    n = n_up + n_dn
    rs = (3.0 / (4.0 * torch.pi * n)) ** (1.0 / 3.0)
    zeta = ((n_up - n_dn) / (n_up + n_dn)).clamp(-1.0 + 1.0e-15, 1.0 - 1.0e-15)
    a, alpha1, beta1, beta2, beta3, beta4 = _PW92_MOD_PARA
    sqrt_rs = rs.sqrt()
    denominator = 2.0 * a * (beta1 * sqrt_rs + beta2 * rs + beta3 * rs * sqrt_rs + beta4 * rs * rs)
    para = -2.0 * a * (1.0 + alpha1 * rs) * torch.log1p(1.0 / denominator)
    # zeta = 0 for a restricted density: f(zeta) = 0, so the ferro and stiffness terms vanish exactly
    fz = ((1.0 + zeta) ** (4.0 / 3.0) + (1.0 - zeta) ** (4.0 / 3.0) - 2.0) / (2.0 ** (4.0 / 3.0) - 2.0)
    eps = para + 0.0 * fz
    phi = 0.5 * ((1.0 + zeta) ** (2.0 / 3.0) + (1.0 - zeta) ** (2.0 / 3.0))
    sigma = s_uu + 2.0 * s_ud + s_dd
    k_f = _KF * n ** (1.0 / 3.0)
    k_s_sq = 4.0 * k_f / math.pi
    t2 = sigma / (4.0 * phi * phi * k_s_sq * n * n)
    phi3 = phi**3
    big_a = (_PBE_BETA / _PBE_GAMMA) / (torch.exp(-eps / (_PBE_GAMMA * phi3)) - 1.0)
    at2 = big_a * t2
    h = _PBE_GAMMA * phi3 * torch.log1p((_PBE_BETA / _PBE_GAMMA) * t2 * (1.0 + at2) / (1.0 + at2 + at2 * at2))
    return total + n * (eps + h)


def pbe_xc(density, sigma):
    """``(e_xc, de/dn, de/dsigma)`` per volume of spin-restricted PBE at a total density ``(N,)`` and
    ``sigma = |grad n|^2``: the reference's spin split (``n_s = n/2``, ``sigma_ss = sigma/4``), clamps
    and empty-point mask (``xc/base.py:55-71,171-189``), and its derivatives by torch autograd on the
    CPU, as ``SemiLocalFunctional.evaluate`` takes them (``xc/base.py:121-157``)."""
    import torch

    dens = torch.as_tensor(np.asarray(density, dtype=np.float64)).clone().requires_grad_(True)
    sig = torch.as_tensor(np.asarray(sigma, dtype=np.float64)).clone().requires_grad_(True)
    with torch.enable_grad():
        n_up = n_dn = 0.5 * dens
        s_uu = s_ud = s_dd = 0.25 * sig
        empty = (n_up + n_dn) < DENSITY_THRESHOLD
        floor = 0.5 * DENSITY_THRESHOLD
        e = _pbe_energy(n_up.clamp_min(floor), n_dn.clamp_min(floor), s_uu.clamp_min(_SIGMA_THRESHOLD), s_ud,
                        s_dd.clamp_min(_SIGMA_THRESHOLD), torch)
        e = torch.where(empty, torch.zeros_like(e), e)
        vn, vs = torch.autograd.grad(e.sum(), [dens, sig])
    return e.detach().numpy(), vn.numpy(), vs.numpy()


# --------------------------------------------------------------------------------------------
# Scenario and numerics
# --------------------------------------------------------------------------------------------


@dataclass
class Scenario:
    """Nuclei (bohr), charges, electron count and functional (``"none"``, ``"lda"`` or ``"pbe"``)."""

    charges: tuple
    positions: tuple
    n_electrons: float
    xc: str = "none"
    n_spin: int = 1


@dataclass
class Numerics:
    """The subset of ``ALL_ELECTRON_NUMERICS`` the hot path reads (reference ``config.py:218-232``,
    ``contract.py:352-529``)."""

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


def nuclear_repulsion(scenario: Scenario) -> float:
    pos = np.asarray(scenario.positions, dtype=np.float64).reshape(-1, 3)
    e = 0.0
    for i in range(len(scenario.charges)):
        for j in range(i + 1, len(scenario.charges)):
            e += scenario.charges[i] * scenario.charges[j] / float(np.linalg.norm(pos[i] - pos[j]))
    return e


def aufbau(eigenvalues, n_electrons, capacity=2.0):
    occ = np.zeros(len(eigenvalues))
    remaining = float(n_electrons)
    for i in range(len(occ)):
        if remaining <= 1e-14:
            break
        take = min(capacity, remaining)
        occ[i] = take
        remaining -= take
    if remaining > 1e-10:
        raise ValueError("electrons do not fit in the states")
    return occ


# --------------------------------------------------------------------------------------------
# Kohn-Sham step
# --------------------------------------------------------------------------------------------


class KohnShamStep:
    """Pure Kohn-Sham map on the cusp-factorised path (reference ``scf/step.py:111-476``)."""

    def __init__(self, scenario: Scenario, numerics: Numerics, grid: Grid, factor: CuspFactor, quadrature: CuspQuadrature):
        self.scenario = scenario
        self.numerics = numerics
        self.grid = grid
        self.factor = factor
        self.q = quadrature
        self.poisson = CoulombCutoffPoisson(grid, numerics.poisson_pad_factor)
        self.positions = np.asarray(scenario.positions, dtype=np.float64).reshape(-1, 3)
        self.charges = tuple(float(z) for z in scenario.charges)
        self.n_electrons = float(scenario.n_electrons)
        self.n_states = max(1, int(-(-self.n_electrons // 2.0)))
        self.ion_ion = nuclear_repulsion(scenario)
        self.bare = CuspHamiltonian(grid, factor, quadrature=quadrature)
        nat = len(self.charges)
        dist = [[float(np.linalg.norm(self.positions[a] - self.positions[b])) for b in range(nat)] for a in range(nat)]
        self.overlap = np.array([[math.exp(-2.0 * zb * dist[a][b]) for b, zb in enumerate(self.charges)] for a in range(nat)])
        pts = grid.points()
        self.cusp_grid = np.stack([np.exp(-2.0 * z * np.sqrt(((pts - self.positions[a]) ** 2).sum(-1))) for a, z in enumerate(self.charges)])
        self.v1s_grid = np.stack([hydrogenic_1s_potential(z, np.sqrt(((pts - self.positions[a]) ** 2).sum(-1))) for a, z in enumerate(self.charges)])
        self.v1s_node = np.stack([hydrogenic_1s_potential(z, np.sqrt(((quadrature.points - self.positions[a]) ** 2).sum(-1))) for a, z in enumerate(self.charges)])
        self.nucleus_f2 = np.array([math.exp(-2.0 * sum(zb * dist[a][b] for b, zb in enumerate(self.charges) if b != a)) for a in range(nat)])
        self.f2 = factor.weight

    def potentials(self, density):
        """Hartree (cusp split + FFT) and LDA weights (reference ``scf/step.py:224-313``)."""
        q = self.q
        n_grid = density.sum(axis=0)
        rho = n_grid / self.f2
        n_at = self.nucleus_f2 * q.interpolate_at(self.positions, rho)
        c = np.linalg.solve(self.overlap, n_at)
        n_smooth = n_grid - (c[:, None] * self.cusp_grid).sum(axis=0)
        v_smooth = self.poisson.solve(n_smooth)
        v_h_grid = (c[:, None] * self.v1s_grid).sum(axis=0) + v_smooth
        v_h_node = (c[:, None] * self.v1s_node).sum(axis=0) + q.interpolate(v_smooth)
        n_node = q.f2 * q.interpolate(rho)
        hartree_weights = q.lumped_weights(q.f2 * v_h_node, self.f2 * v_h_grid)
        hartree_energy = 0.5 * (hartree_weights * rho).sum()
        if self.scenario.xc == "pbe":
            # the sigma path (reference ``scf/step.py:253-279``)
            grad_rho = self.grid.gradient(rho)
            rho_node = q.interpolate(rho)
            grad_n_node = q.f2[None, :] * (q.interpolate(grad_rho) - 2.0 * rho_node[None, :] * q.grad_u)
            grad_n_grid = self.f2[None, :] * (grad_rho - 2.0 * rho[None, :] * self.factor.grad_u)
            sigma_grid = (grad_n_grid * grad_n_grid).sum(axis=0)
            sigma_node = (grad_n_node * grad_n_node).sum(axis=0)
            e_grid, v_grid, s_grid = pbe_xc(n_grid, sigma_grid)
            e_node, v_node, s_node = pbe_xc(n_node, sigma_node)
            xc_weights = q.lumped_weights(q.f2 * v_node, self.f2 * v_grid)
            flux_node = 2.0 * s_node[None, :] * q.f2[None, :] * grad_n_node
            flux_grid = 2.0 * s_grid[None, :] * self.f2[None, :] * grad_n_grid
            xc_weights = xc_weights + q.lumped_gradient_weights(flux_node, flux_grid)
            cross_node = -2.0 * (flux_node * q.grad_u).sum(axis=0)
            cross_grid = -2.0 * (flux_grid * self.factor.grad_u).sum(axis=0)
            xc_weights = xc_weights + q.lumped_weights(cross_node, cross_grid)
        else:
            e_grid, v_grid = lda_vwn(n_grid)
            e_node, v_node = lda_vwn(n_node)
            xc_weights = q.lumped_weights(q.f2 * v_node, self.f2 * v_grid)
        xc_energy = q.integrate(e_node, e_grid)
        dc = (xc_weights * rho).sum()
        effective = (hartree_weights + xc_weights) / q.mass_weights
        return {
            "v_hartree": v_h_grid,
            "v_xc": v_grid,
            "effective": effective,
            "hartree_energy": float(hartree_energy),
            "xc_energy": float(xc_energy),
            "xc_double_counting": float(dc),
            "coefficients": c,
        }

    def hamiltonian(self, pot):
        return CuspHamiltonian(self.grid, self.factor, extra_potential=pot["effective"], quadrature=self.q)

    def diagonalise(self, h, initial=None, max_iterations=None, bounds=None):
        gen = make_generator(self.numerics.seed)
        cfg = self.numerics.eigen
        if max_iterations is not None:
            cfg = EigenConfig(**{**cfg.__dict__, "max_iterations": max_iterations})
        return chefsi_solve(h, self.grid, cfg, self.n_states, initial=initial, generator=gen, bounds=bounds, ritz_stop=False)

    def occupy(self, evals):
        return aufbau(evals, self.n_electrons)

    def output_density(self, eigen, occ):
        rho = (occ[:, None] * eigen.eigenvectors**2).sum(axis=0)
        return (rho * self.f2)[None, :], rho

    def external_parts(self, rho):
        """``((1/2) Q[f^2 grad rho . grad u], Q[f^2 rho W])`` (reference ``hamiltonian.py:360-377``)."""
        q = self.q
        rho_node = q.interpolate(rho)
        grad_rho = self.grid.gradient(rho)
        grad_rho_node = q.interpolate(grad_rho)
        cross_node = q.f2 * (grad_rho_node * q.grad_u).sum(axis=0)
        cross_grid = self.factor.weight * (grad_rho * self.factor.grad_u).sum(axis=0)
        cross = 0.5 * q.integrate(cross_node, cross_grid)
        potential = q.integrate(q.f2 * rho_node * q.transformed_potential, self.factor.weight * rho * self.factor.transformed_potential)
        return cross, potential

    def kinetic_energy(self, phi, occ):
        """``T_s`` (reference ``operators/hamiltonian.py:344-358``)."""
        form = self.bare.measure.inner(phi, self.bare.kinetic_apply(phi))
        first = (occ * form).sum()
        rho = (occ[:, None] * phi**2).sum(axis=0)
        cross, potential = self.external_parts(rho)
        return first - cross - potential

    def energy_terms(self, eigen, occ, rho_out, pin, pout):
        """Energy breakdown (reference ``scf/step.py:398-439``)."""
        band = float((occ * eigen.eigenvalues).sum())
        kinetic = float(self.kinetic_energy(eigen.eigenvectors, occ))
        dc_out = float((pin["effective"] * self.q.mass_weights * rho_out).sum())
        kpe = band - dc_out
        cross, potential = self.external_parts(rho_out)
        external_direct = float(cross + 2.0 * potential)
        harris = band - pin["hartree_energy"] - pin["xc_double_counting"] + pin["xc_energy"] + self.ion_ion
        if pout is None:
            total = harris
            hartree, xc = pin["hartree_energy"], pin["xc_energy"]
        else:
            hartree, xc = pout["hartree_energy"], pout["xc_energy"]
            total = kpe + hartree + xc + self.ion_ion
        external = total - kinetic - hartree - xc - self.ion_ion
        return {"total": total, "kinetic": kinetic, "external": external, "hartree": hartree, "xc": xc,
                "harris_foulkes": harris, "band": band, "external_direct": external_direct, "kinetic_plus_external": kpe}


def build_geometry(scenario: Scenario, numerics: Numerics, anchor=None):
    grid = Grid.from_box(numerics.box_lengths, numerics.spacing, anchor=anchor, fd_order=numerics.fd_order)
    factor = CuspFactor(grid, scenario.charges, scenario.positions)
    quadrature = CuspQuadrature(grid, factor)
    return grid, factor, quadrature


#: The starting densities of the SCF and their exponent scales (reference ``scf/loop.py:68,181``).
INITIAL_GUESSES = {"hydrogenic": 1.0, "diffuse": 0.5, "compact": 2.0}


def initial_density(step: KohnShamStep, kind: str = "hydrogenic"):
    """Superposition of ``(zeta^3/pi) exp(-2 zeta s_a)``, ``zeta = Z_a`` times 1 (hydrogenic), 0.5
    (diffuse) or 2 (compact), normalised in the cusp-aware measure (reference ``scf/loop.py:173-193``)."""
    if kind not in INITIAL_GUESSES:
        raise ValueError(f"unknown initial guess {kind!r}; choose from {tuple(INITIAL_GUESSES)}")
    pts = step.grid.points()
    total = np.zeros(step.grid.n_points)
    for a, z in enumerate(step.charges):
        zeta = z * INITIAL_GUESSES[kind]
        d = np.sqrt(((pts - step.positions[a]) ** 2).sum(-1))
        total = total + (zeta**3 / math.pi) * np.exp(-2.0 * zeta * d)
    rho = total / step.f2
    charge = float((step.q.mass_weights * rho).sum())
    total = total * (step.n_electrons / charge)
    return total[None, :]


class PulayMixer:
    """Periodic Pulay/DIIS with Tikhonov (reference ``scf/mixing.py:86-162``)."""

    def __init__(self, alpha, history, period, inner):
        self.alpha, self.history, self.period, self.inner = alpha, history, period, inner
        self.inputs, self.residuals = [], []

    def mix(self, n_in, n_out, iteration):
        residual = n_out - n_in
        self.inputs.append(n_in.copy())
        self.residuals.append(residual.copy())
        if len(self.inputs) > self.history:
            self.inputs.pop(0)
            self.residuals.pop(0)
        m = len(self.inputs)
        if m < 2 or iteration % self.period != 0:
            return n_in + self.alpha * residual
        gram = np.zeros((m, m))
        for i in range(m):
            for j in range(i, m):
                gram[i, j] = gram[j, i] = float(self.inner(self.residuals[i].reshape(-1), self.residuals[j].reshape(-1)))
        scale = float(gram.diagonal().max()) or 1.0
        gram = gram + 1.0e-12 * scale * np.eye(m)
        system = np.zeros((m + 1, m + 1))
        system[:m, :m] = gram
        system[:m, m] = -1.0
        system[m, :m] = 1.0
        rhs = np.zeros(m + 1)
        rhs[m] = 1.0
        coeffs = np.linalg.solve(system, rhs)[:m]
        mixed = np.zeros_like(n_in)
        for c, si, sr in zip(coeffs, self.inputs, self.residuals):
            mixed = mixed + float(c) * (si + self.alpha * sr)
        return mixed


class LinearMixer:
    def __init__(self, alpha):
        self.alpha = alpha

    def mix(self, n_in, n_out, iteration):
        return n_in + self.alpha * (n_out - n_in)


LANCZOS_REUSE_BACKSTOP = 8
LANCZOS_REUSE_MAX_WIDENING = 0.05
FILTER_STEPS_PER_ITERATION = 4


class BoundReusePolicy:
    """When to reuse spectral bounds (reference ``scf/loop.py:94-170``)."""

    def __init__(self):
        self.bounds = None
        self.since = 0
        self.widening = 0.0
        self.top_fraction = float("nan")
        self.history = []

    def decide(self, residual_rose, rung_changed):
        if self.bounds is None:
            return "no carried interval"
        if rung_changed:
            return "fallback rung taken"
        if residual_rose:
            return "SCF residual rose"
        if self.since >= LANCZOS_REUSE_BACKSTOP:
            return "backstop"
        lo, hi = self.bounds
        if self.widening > LANCZOS_REUSE_MAX_WIDENING * (hi - lo):
            return "accumulated shift"
        if not self.top_fraction < 0.9:
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


def solve_self_consistent(scenario: Scenario, numerics: Numerics, anchor=None, geometry=None,
                          initial_guess: str = "hydrogenic"):
    """The SCF driver (reference ``scf/loop.py:255-460``): periodic Pulay, 4 filter steps per warm
    iteration, Weyl-widened bound reuse, fallback ladder, final full diagonalisation."""
    grid, factor, q = geometry if geometry is not None else build_geometry(scenario, numerics, anchor)
    step = KohnShamStep(scenario, numerics, grid, factor, q)
    f2 = step.f2

    def norm_l1(delta):
        return float((q.mass_weights * (np.abs(delta.sum(axis=0)) / f2)).sum())

    inner_weights = q.mass_weights / f2**2

    def inner(a, b):
        return (inner_weights * a * b).sum()

    n_in = initial_density(step, initial_guess)
    mixer = PulayMixer(numerics.alpha, numerics.history, numerics.pulay_period, inner)
    ladder = [
        lambda: PulayMixer(numerics.alpha, numerics.history, numerics.pulay_period, inner),
        lambda: PulayMixer(0.5 * numerics.alpha, numerics.history, numerics.pulay_period, inner),
        lambda: LinearMixer(0.5 * numerics.alpha),
    ]
    rung = 0
    energies, residuals, evals_hist = [], [], []
    worsened = 0
    previous_residual = float("inf")
    initial_block = None
    policy = BoundReusePolicy()
    previous_effective = None
    rung_changed = False
    converged = False
    applications = []
    iteration = 0
    for iteration in range(1, numerics.max_scf + 1):
        pin = step.potentials(n_in)
        h = step.hamiltonian(pin)
        steps = None if initial_block is None else FILTER_STEPS_PER_ITERATION
        reason = policy.decide(worsened > 0, rung_changed)
        hint = None
        if reason is None and previous_effective is not None:
            hint = SpectralBoundsHint(policy.bounds[0], policy.bounds[1], float(np.abs(pin["effective"] - previous_effective).max()))
        elif reason is None:
            reason = "no previous potential"
        eigen = step.diagonalise(h, initial=initial_block, max_iterations=steps, bounds=hint)
        policy.observe(reason, eigen.info)
        previous_effective = pin["effective"]
        rung_changed = False
        applications.append(h.apply_count)
        initial_block = eigen.eigenvectors
        occ = step.occupy(eigen.eigenvalues)
        n_out, rho_out = step.output_density(eigen, occ)
        pout = step.potentials(n_out)
        terms = step.energy_terms(eigen, occ, rho_out, pin, pout)
        residual = n_out - n_in
        res_norm = norm_l1(residual)
        n_next = mixer.mix(n_in, n_out, iteration)
        energies.append(terms["total"])
        residuals.append(res_norm)
        evals_hist.append(eigen.eigenvalues.copy())
        de = abs(energies[-1] - energies[-2]) if len(energies) > 1 else float("inf")
        if iteration >= numerics.min_scf and de < numerics.energy_tol and res_norm < numerics.density_tol:
            converged = True
            break
        worsened = worsened + 1 if res_norm > previous_residual else 0
        previous_residual = res_norm
        if worsened >= 3 and rung + 1 < len(ladder):
            rung += 1
            mixer = ladder[rung]()
            worsened = 0
            rung_changed = True
            n_next = n_in + 0.5 * numerics.alpha * residual
        n_in = n_next
    h_final = step.hamiltonian(pin)
    eigen = step.diagonalise(h_final, initial=eigen.eigenvectors)
    occ = step.occupy(eigen.eigenvalues)
    n_out, rho_out = step.output_density(eigen, occ)
    pout = step.potentials(n_out)
    terms = step.energy_terms(eigen, occ, rho_out, pin, pout)
    return {
        "energies": terms,
        "eigenvalues": eigen.eigenvalues,
        "eigenvectors": eigen.eigenvectors,
        "density": n_out,
        "scf_iterations": iteration,
        "converged": converged and eigen.converged,
        "energy_history": energies,
        "residual_history": residuals,
        "bound_history": policy.history,
        "applications": applications,
        "step": step,
        "orthonormality_error": orthonormality_error(eigen.eigenvectors, step.bare.measure),
    }


def solve_noninteracting(scenario: Scenario, numerics: Numerics, anchor=None, geometry=None):
    """Fixed-potential eigenproblem on the cusp-factorised operator (reference
    ``scf/noninteracting.py:118-330``)."""
    grid, factor, q = geometry if geometry is not None else build_geometry(scenario, numerics, anchor)
    h = CuspHamiltonian(grid, factor, quadrature=q)
    n_states = max(1, int(-(-float(scenario.n_electrons) // 2.0)))
    gen = make_generator(numerics.seed)
    result = chefsi_solve(h, grid, numerics.eigen, n_states, generator=gen, ritz_stop=True)
    occ = aufbau(result.eigenvalues, scenario.n_electrons)
    step = KohnShamStep.__new__(KohnShamStep)
    step.grid, step.factor, step.q, step.bare = grid, factor, q, h
    kinetic = float(KohnShamStep.kinetic_energy(step, result.eigenvectors, occ))
    band = float((occ * result.eigenvalues).sum())
    ion = nuclear_repulsion(scenario)
    return {
        "eigenvalues": result.eigenvalues,
        "eigenvectors": result.eigenvectors,
        "n_iterations": result.n_iterations,
        "converged": result.converged,
        "kinetic": kinetic,
        "band": band,
        "total": kinetic + (band - kinetic) + ion,
        "info": result.info,
        "hamiltonian": h,
    }
