"""Oracle eigensolver: Lanczos bounds, the scaled Chebyshev filter, Rayleigh-Ritz and CheFSI.

Test infrastructure only. Follows ``src/cdft/eigen/chefsi.py``, ``rayleigh_ritz.py`` and
``operators/hamiltonian.py:29-89`` of the reference, in numpy float64.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

BOUND_REUSE_RITZ_FRACTION = 0.9


def seeded_randn(shape, generator) -> np.ndarray:
    """Standard normal draw from a seeded CPU torch generator (reference ``precision.py:246-263``).

    The same stream the reference draws, so a solve here follows the reference's path.
    """
    import torch

    size = (shape,) if isinstance(shape, int) else tuple(shape)
    return torch.randn(size, generator=generator, dtype=torch.float64).numpy()


def make_generator(seed: int):
    import torch

    return torch.Generator(device="cpu").manual_seed(int(seed))


def lanczos_bounds(apply, start, inner, norm, steps, safety):
    """Rigorous spectral enclosure from a short Lanczos run in the operator's measure, widened by
    ``safety`` (reference ``operators/hamiltonian.py:29-89``)."""
    v = start / norm(start)
    alphas, betas = [], []
    v_prev = np.zeros_like(v)
    beta = 0.0
    for _ in range(steps):
        w = apply(v)
        alpha = inner(v, w)
        w = w - alpha * v - beta * v_prev
        w = w - inner(w, v) * v - inner(w, v_prev) * v_prev
        beta = norm(w)
        alphas.append(float(alpha))
        betas.append(float(beta))
        v_prev, v = v, w / beta
    for index, value in enumerate(betas):
        if value < 1e-12:
            del alphas[index + 1 :], betas[index + 1 :]
            break
    size = len(alphas)
    tri = np.zeros((size, size))
    for i in range(size):
        tri[i, i] = alphas[i]
        if i + 1 < size:
            tri[i, i + 1] = tri[i + 1, i] = betas[i]
    ritz = np.linalg.eigvalsh(tri)
    residual = betas[-1] if betas else 0.0
    lo = float(ritz[0]) - residual
    hi = float(ritz[-1]) + residual
    centre = 0.5 * (hi + lo)
    half = 0.5 * (hi - lo) * safety
    return centre - half, centre + half


def chebyshev_coefficients(degree, lower, upper, spectrum_min):
    """Rows ``(c, a_k, b_k)`` of the scaled recurrence [F2] (reference ``eigen/chefsi.py:116-136``)."""
    if upper <= lower:
        raise ValueError(f"filter interval is empty: lower={lower}, upper={upper}")
    e = 0.5 * (upper - lower)
    c = 0.5 * (upper + lower)
    sigma = e / (spectrum_min - c)
    sigma1 = sigma
    rows = [(c, sigma1 / e, 0.0)]
    for _ in range(2, degree + 1):
        sigma2 = 1.0 / (2.0 / sigma1 - sigma)
        rows.append((c, 2.0 * sigma2 / e, sigma * sigma2))
        sigma = sigma2
    return rows


def chebyshev_filter(hamiltonian, block, degree, lower, upper, spectrum_min):
    """Scaled Chebyshev filter damping ``[lower, upper]`` (reference ``eigen/chefsi.py:81-113``)."""
    rows = chebyshev_coefficients(degree, lower, upper, spectrum_min)
    c, a1, _ = rows[0]
    applied = hamiltonian.apply(block)
    y = (applied - c * block) * a1
    previous = block
    for c, a, b in rows[1:]:
        applied = hamiltonian.apply(y)
        y_new = (applied - c * y) * a - b * previous
        previous, y = y, y_new
    return y


def orthonormalise(block, measure):
    """Cholesky orthonormalisation in the measure; eigh fallback dropping ``lambda < lambda_max
    1e-14`` (reference ``eigen/rayleigh_ritz.py:51-72``)."""
    flat = np.asarray(block).reshape(-1, measure.grid.n_points)
    gram = measure.gram(flat)
    try:
        chol = np.linalg.cholesky(gram)
    except np.linalg.LinAlgError:
        evals, evecs = np.linalg.eigh(0.5 * (gram + gram.T))
        keep = evals > evals[-1] * 1e-14
        scaled = evecs[:, keep] / np.sqrt(evals[keep])
        return scaled.T @ flat
    import scipy.linalg

    return scipy.linalg.solve_triangular(chol, flat, lower=True)


def orthonormality_error(block, measure):
    gram = measure.gram(block)
    return float(np.abs(gram - np.eye(gram.shape[0])).max())


def rayleigh_ritz(hamiltonian, block, measure):
    """Orthonormalise, project, symmetrise, diagonalise, rotate (reference
    ``eigen/rayleigh_ritz.py:75-102``)."""
    basis = orthonormalise(block, measure)
    sub = measure.cross(basis, hamiltonian.apply(basis))
    sub = 0.5 * (sub + sub.T)
    evals, rotation = np.linalg.eigh(sub)
    vectors = rotation.T @ basis
    return evals, vectors, orthonormality_error(vectors, measure)


def residual_norms(hamiltonian, evals, vectors, measure):
    """``||A phi - eps phi||`` in the measure (reference ``eigen/rayleigh_ritz.py:105-115``)."""
    applied = hamiltonian.apply(vectors)
    return measure.norm(applied - evals[:, None] * vectors)


@dataclass
class EigenConfig:
    """Eigensolver settings (reference ``contract.py:430-462``, ``config.py:228-230``)."""

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
    eigenvalues: np.ndarray
    eigenvectors: np.ndarray
    residuals: np.ndarray
    n_iterations: int
    converged: bool
    info: dict = field(default_factory=dict)


@dataclass(frozen=True)
class SpectralBoundsHint:
    """Carried bounds widened by a Weyl shift (reference ``eigen/chefsi.py:169-182``)."""

    lower: float
    upper: float
    shift: float = 0.0


def _interval_fraction(value, low, high):
    width = high - low
    return (value - low) / width if width > 0.0 else float("inf")


def chefsi_solve(hamiltonian, grid, config: EigenConfig, n_states, initial=None, generator=None,
                 bounds: SpectralBoundsHint | None = None, ritz_stop=True) -> EigenResult:
    """Chebyshev-filtered subspace iteration (reference ``eigen/chefsi.py:442-585``)."""
    measure = hamiltonian.measure
    extra = max(config.n_extra_states, int(config.extra_states_fraction * n_states))
    size = min(n_states + extra, grid.n_points)
    if initial is None:
        block = seeded_randn((size, grid.n_points), generator)
    else:
        block = np.asarray(initial, dtype=np.float64).reshape(-1, grid.n_points)
        if block.shape[0] < size:
            pad = seeded_randn((size - block.shape[0], grid.n_points), generator)
            block = np.concatenate([block, pad], axis=0)

    def lanczos():
        return hamiltonian.spectral_bounds(steps=config.lanczos_steps, safety=config.bound_safety, generator=generator)

    source = "unset"
    shift_used = 0.0
    if bounds is None:
        spectrum_min, spectrum_max = lanczos()
        source = "lanczos"
    evals, vectors, _ = rayleigh_ritz(hamiltonian, block, measure)
    residual = residual_norms(hamiltonian, evals, vectors, measure)
    previous = evals.copy()
    stable = 0
    stop_reason = "iteration limit"
    weak_form = ritz_stop and hamiltonian.uses_divergence_form
    eigenvalue_tol = 0.1 * config.cross_check_tol
    top = min(size, evals.shape[0]) - 1
    next_lower = float(evals[top])
    if bounds is not None:
        spectrum_min, spectrum_max = bounds.lower - bounds.shift, bounds.upper + bounds.shift
        source = "reused"
        shift_used = float(bounds.shift)
    top_fraction = float("-inf")
    converged = False
    iteration = 0
    drift = float("inf")
    history = []
    for iteration in range(1, config.max_iterations + 1):
        lower = next_lower
        fraction = _interval_fraction(lower, spectrum_min, spectrum_max)
        if source == "reused" and not fraction < BOUND_REUSE_RITZ_FRACTION:
            spectrum_min, spectrum_max = lanczos()
            source = f"lanczos (reuse rejected at iteration {iteration})"
            fraction = _interval_fraction(lower, spectrum_min, spectrum_max)
        top_fraction = max(top_fraction, fraction)
        lower = min(max(lower, spectrum_min + 1e-8), spectrum_max - 1e-8)
        filtered = chebyshev_filter(hamiltonian, vectors, config.chebyshev_degree, lower, spectrum_max, spectrum_min)
        evals, vectors, orth = rayleigh_ritz(hamiltonian, filtered, measure)
        residual = residual_norms(hamiltonian, evals, vectors, measure)
        worst = float(residual[:n_states].max())
        drift = float(np.abs(evals[:n_states] - previous[:n_states]).max())
        top = min(size, evals.shape[0]) - 1
        next_lower = float(evals[top])
        previous = evals.copy()
        history.append((iteration, worst, orth, drift))
        if worst < config.residual_tol:
            converged = True
            stop_reason = "residual below residual_tol"
            break
        stable = stable + 1 if (weak_form and drift < eigenvalue_tol) else 0
        if stable >= 2:
            converged = True
            stop_reason = "Ritz values stable on the weak-form path"
            break
    info = {
        "spectral_bounds": (spectrum_min, spectrum_max),
        "bounds_source": source,
        "bounds_shift": shift_used,
        "top_ritz_fraction": top_fraction,
        "stop_reason": stop_reason,
        "eigenvalue_drift": drift,
        "history": history,
    }
    return EigenResult(evals[:n_states], vectors[:n_states], residual[:n_states], iteration, converged, info)


# ------------------------------------------------------------------------------------------------
# LOBPCG, the second eigensolver of the two-solver cross-check (reference ``eigen/lobpcg.py``)
# ------------------------------------------------------------------------------------------------


def tpa_filter(shape, spacing):
    """The Teter-Payne-Allan kinetic filter on the real-FFT half spectrum of the box (reference
    ``eigen/lobpcg.py:207-239``): ``K(x) = P(x) / (P(x) + 16 x^4)``, ``P = 27 + 18x + 12x^2 + 8x^3``,
    ``x = (G^2/2) / E_ref`` with ``E_ref`` the mean of ``G^2/2`` over the half spectrum (at least 1e-6)."""
    n0, n1, n2 = shape
    two_pi = 2.0 * np.pi
    kx = two_pi * np.fft.fftfreq(n0, d=spacing)
    ky = two_pi * np.fft.fftfreq(n1, d=spacing)
    kz = two_pi * np.fft.rfftfreq(n2, d=spacing)
    kinetic = 0.5 * (kx[:, None, None] ** 2 + ky[None, :, None] ** 2 + kz[None, None, :] ** 2)
    reference = max(float(kinetic.mean()), 1.0e-6)
    x = kinetic / reference
    poly = 27.0 + 18.0 * x + 12.0 * x**2 + 8.0 * x**3
    return poly / (poly + 16.0 * x**4)


def tpa_precondition(block, grid, cusp_factor=None):
    """The filter applied to each row of ``block`` by FFT over the box. With the cusp factor
    ``f = max(exp(-u), 1e-150)`` it acts in the wavefunction frame: ``f^-1 K (f r)`` (reference
    ``lobpcg.py:129-133``, the ``"similarity"`` frame)."""
    flat = np.asarray(block, dtype=np.float64).reshape(-1, grid.n_points)
    if cusp_factor is not None:
        flat = flat * cusp_factor
    box = flat.reshape(-1, *grid.shape)
    spectrum = np.fft.rfftn(box, axes=(-3, -2, -1)) * tpa_filter(grid.shape, grid.spacing)
    out = np.fft.irfftn(spectrum, s=grid.shape, axes=(-3, -2, -1)).reshape(-1, grid.n_points)
    return out / cusp_factor if cusp_factor is not None else out


def stable_basis(block, measure, threshold=1e-12):
    """An orthonormal basis of ``block``'s span in the measure, directions with Gram eigenvalue at
    most ``threshold`` times the largest dropped (reference ``lobpcg.py:254-271``); returns the basis
    and the number dropped."""
    flat = np.asarray(block).reshape(-1, measure.grid.n_points)
    if flat.shape[0] == 0:
        return flat, 0
    gram = measure.gram(flat)
    evals, evecs = np.linalg.eigh(0.5 * (gram + gram.T))
    largest = evals[-1]
    if largest <= 0.0:
        return flat[:0], flat.shape[0]
    keep = evals > largest * threshold
    scaled = evecs[:, keep] / np.sqrt(evals[keep])
    return scaled.T @ flat, int((~keep).sum())


def project_out(block, basis, measure):
    """``block`` with the span of the orthonormal ``basis`` removed by two classical Gram-Schmidt
    passes, then made orthonormal itself (reference ``lobpcg.py:242-251``)."""
    for _ in range(2):
        block = block - measure.cross(block, basis) @ basis
    return stable_basis(block, measure)[0]


def lobpcg_solve(hamiltonian, grid, config: EigenConfig, n_states, initial=None, generator=None,
                 frame="similarity"):
    """Block LOBPCG in the operator's measure (reference ``eigen/lobpcg.py:71-165``).

    Stops when every residual is below ``residual_tol``, or when the Ritz values move less than
    ``0.1 cross_check_tol`` for two iterations in a row. Residuals at most ``0.1 residual_tol`` are
    deflated out of the search directions. The preconditioner is the TPA filter in ``frame``
    (``"similarity"``, ``"transformed"`` or ``"none"``)."""
    measure = hamiltonian.measure
    factor = getattr(hamiltonian, "factor", None)
    cusp = None
    if frame == "similarity" and factor is not None:
        cusp = np.maximum(np.exp(-np.asarray(factor.u)), 1.0e-150)
    if initial is None:
        x = seeded_randn((n_states, grid.n_points), generator)
    else:
        x = np.asarray(initial, dtype=np.float64).reshape(-1, grid.n_points)[:n_states].copy()
    x = orthonormalise(x, measure)
    hx = hamiltonian.apply(x)
    sub = measure.cross(x, hx)
    evals, rotation = np.linalg.eigh(0.5 * (sub + sub.T))
    x = rotation.T @ x
    hx = rotation.T @ hx
    p = None
    converged = False
    previous = evals.copy()
    stable = 0
    tol = 0.1 * config.cross_check_tol
    history = []
    residual = measure.norm(hx - evals[:, None] * x)
    iteration = 0
    for iteration in range(1, config.max_iterations + 1):
        r = hx - evals[:, None] * x
        residual = measure.norm(r)
        keep = residual > 0.1 * config.residual_tol
        worst = float(residual.max())
        drift = float(np.abs(evals - previous).max()) if iteration > 1 else float("inf")
        previous = evals.copy()
        history.append((iteration, worst, drift))
        if worst < config.residual_tol:
            converged = True
            break
        stable = stable + 1 if drift < tol else 0
        if stable >= 2:
            converged = True
            break
        w = r[keep] if keep.any() else r
        if cusp is not None:
            w = tpa_precondition(w, grid, cusp)
        elif frame == "transformed":
            w = tpa_precondition(w, grid)
        w = project_out(w, x, measure)
        if w.shape[0] == 0:
            converged = True
            break
        blocks = [x, w] if p is None else [x, w, p]
        basis, _ = stable_basis(np.concatenate(blocks, axis=0), measure)
        h_basis = hamiltonian.apply(basis)
        sub = measure.cross(basis, h_basis)
        sub_evals, sub_vectors = np.linalg.eigh(0.5 * (sub + sub.T))
        coeff = sub_vectors[:, :n_states]
        x_new = coeff.T @ basis
        hx_new = coeff.T @ h_basis
        p = x_new - measure.cross(x_new, x) @ x
        x, hx = x_new, hx_new
        evals = sub_evals[:n_states]
    return EigenResult(evals[:n_states], x[:n_states], residual[:n_states], iteration, converged,
                       {"history": history, "eigenvalue_drift": history[-1][2] if history else float("inf")})
