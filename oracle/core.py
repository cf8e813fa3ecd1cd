"""Oracle core: FD weights, the uniform grid, the Kato cusp factor and the divergence-form operator.

Test infrastructure only (see the package docstring). Every routine follows the cited reference
lines operation for operation, in numpy float64, so the GPU path can be checked against it.
"""

from __future__ import annotations

import math
from concurrent.futures import ThreadPoolExecutor
from fractions import Fraction
from functools import lru_cache

import numpy as np

#: Optional host threads for the divergence (set by :func:`set_threads`): each axis's term is computed
#: in slices along another spatial axis on the threads (numpy releases the interpreter lock inside its
#: array loops). Every element gets the same operations as without threads, so the results are
#: bitwise the single-threaded ones.
_POOL: ThreadPoolExecutor | None = None
_POOL_THREADS = 1


def set_threads(threads: int) -> None:
    """Use ``threads`` host threads for block applies (1 = none)."""
    global _POOL, _POOL_THREADS
    if _POOL is not None:
        _POOL.shutdown(wait=True)
        _POOL = None
    _POOL_THREADS = max(1, int(threads))
    if _POOL_THREADS > 1:
        _POOL = ThreadPoolExecutor(max_workers=_POOL_THREADS)

# --------------------------------------------------------------------------------------------
# Finite-difference weights (exact rationals)
# --------------------------------------------------------------------------------------------


def _solve_rational(matrix: list[list[Fraction]], rhs: list[Fraction]) -> list[Fraction]:
    """Gauss-Jordan elimination over the rationals (no rounding until the final float)."""
    size = len(rhs)
    for col in range(size):
        pivot = next(r for r in range(col, size) if matrix[r][col] != 0)
        matrix[col], matrix[pivot] = matrix[pivot], matrix[col]
        rhs[col], rhs[pivot] = rhs[pivot], rhs[col]
        inv = matrix[col][col]
        matrix[col] = [v / inv for v in matrix[col]]
        rhs[col] = rhs[col] / inv
        for row in range(size):
            if row != col and matrix[row][col] != 0:
                f = matrix[row][col]
                matrix[row] = [a - f * b for a, b in zip(matrix[row], matrix[col])]
                rhs[row] = rhs[row] - f * rhs[col]
    return rhs


@lru_cache(maxsize=None)
def staggered_coefficients(accuracy: int) -> tuple[float, ...]:
    """Face-derivative weights ``c_k``, ``k = -p..p-1`` (reference ``operators/divergence.py:30-74``).

    ``sum_k c_k (k + 1/2)^m = delta_{m,1}`` solved exactly.
    """
    if accuracy % 2 or accuracy < 2:
        raise ValueError(f"accuracy must be a positive even integer, got {accuracy}")
    half = accuracy // 2
    positions = [Fraction(2 * k + 1, 2) for k in range(-half, half)]
    matrix = [[pos**m for pos in positions] for m in range(len(positions))]
    rhs = [Fraction(0)] * len(positions)
    rhs[1] = Fraction(1)
    return tuple(float(v) for v in _solve_rational(matrix, rhs))


@lru_cache(maxsize=None)
def fd_coefficients(derivative: int, accuracy: int) -> tuple[float, ...]:
    """Central FD weights, offsets ``-hw..hw`` (reference ``operators/laplacian.py:35-75``)."""
    hw = accuracy // 2 + (derivative - 1) // 2
    offsets = list(range(-hw, hw + 1))
    matrix = [[Fraction(o) ** m for o in offsets] for m in range(len(offsets))]
    rhs = [Fraction(0)] * len(offsets)
    rhs[derivative] = Fraction(math.factorial(derivative))
    return tuple(float(v) for v in _solve_rational(matrix, rhs))


@lru_cache(maxsize=None)
def gauss_legendre(order: int) -> tuple[np.ndarray, np.ndarray]:
    """GL rule on [-1, 1]; numpy's ``leggauss``, which the reference reproduces bit for bit
    (reference ``operators/gauss_legendre.py:22-50``)."""
    from numpy.polynomial.legendre import leggauss

    x, w = leggauss(int(order))
    return x, w


# --------------------------------------------------------------------------------------------
# Grid
# --------------------------------------------------------------------------------------------


class Grid:
    """Full-box uniform grid, C order with z fastest (reference ``grid.py:489-551``)."""

    def __init__(self, shape, spacing, origin, fd_order=8, gradient_order=8):
        self.shape = tuple(int(n) for n in shape)
        self.spacing = float(spacing)
        self.origin = tuple(float(o) for o in origin)
        self.fd_order = fd_order
        self.gradient_order = gradient_order
        self.n_points = self.shape[0] * self.shape[1] * self.shape[2]
        self.volume_element = self.spacing**3

    @classmethod
    def from_box(cls, box_lengths, spacing, anchor=None, fd_order=8):
        """``UniformGrid.from_config`` for ``DomainMode.BOX`` and a zero halo (reference
        ``grid.py:580-605``): counts ``round(L/h)+1``, ``h`` re-derived from the first edge, origin
        snapped so the coordinate origin (or ``anchor``) is a lattice point."""
        h = float(spacing)
        counts = tuple(max(int(round(length / h)) + 1, 1) for length in box_lengths)
        h = box_lengths[0] / (counts[0] - 1) if counts[0] > 1 else h
        if anchor is None:
            origin = tuple(-h * round(0.5 * length / h) for length in box_lengths)
        else:
            origin = tuple(float(a) - h * round((float(a) + 0.5 * length) / h) for a, length in zip(anchor, box_lengths))
        return cls(counts, h, origin, fd_order=fd_order, gradient_order=fd_order)

    def box_index(self) -> np.ndarray:
        idx = np.indices(self.shape).reshape(3, -1).T
        return idx

    def points(self) -> np.ndarray:
        """``origin + index * h``, shape ``(N, 3)`` (reference ``grid.py:664-667``)."""
        return np.asarray(self.origin) + self.box_index().astype(np.float64) * self.spacing

    def integrate(self, field):
        return (np.asarray(field, dtype=np.float64) * self.volume_element).sum(axis=-1)

    def to_box(self, field):
        return np.asarray(field).reshape(*np.shape(field)[:-1], *self.shape)

    def gradient(self, field: np.ndarray) -> np.ndarray:
        """Order-``gradient_order`` central gradient with a zero halo, ``(..., 3, N)``
        (reference ``grid.py:740-746``, ``operators/laplacian.py:111-144``)."""
        coeffs = fd_coefficients(1, self.gradient_order)
        hw = self.gradient_order // 2
        scale = 1.0 / self.spacing
        box = self.to_box(field)
        batch = box.shape[:-3]
        pad = [(0, 0)] * len(batch) + [(hw, hw)] * 3
        padded = np.pad(box, pad)
        comps = []
        n0, n1, n2 = self.shape
        for axis in range(3):
            acc = None
            for index, c in enumerate(coeffs):
                if c == 0.0:
                    continue
                sl = [slice(hw, hw + n0), slice(hw, hw + n1), slice(hw, hw + n2)]
                off = index - hw
                sl[axis] = slice(hw + off, hw + off + self.shape[axis])
                term = padded[(Ellipsis, *sl)]
                acc = term * (c * scale) if acc is None else acc + term * (c * scale)
            comps.append(acc.reshape(*batch, -1))
        return np.stack(comps, axis=-2)

    def partial_derivative(self, field, axis):
        return self.gradient(field)[..., axis, :]


# --------------------------------------------------------------------------------------------
# Kato cusp factor
# --------------------------------------------------------------------------------------------


class CuspFactor:
    """``f = exp(-sum Z_a |r-R_a|)``: ``u``, ``grad u``, ``f^2`` and the cell-averaged
    ``W = -|grad u|^2/2`` (reference ``operators/cusp.py:46-145``)."""

    average_radius_cells = 2.0
    average_quadrature = 4

    def __init__(self, grid: Grid, charges, positions):
        self.grid = grid
        self.charges = tuple(float(z) for z in charges)
        self.positions = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
        self.is_identity = len(self.charges) == 0
        pts = grid.points()
        n = grid.n_points
        u = np.zeros(n)
        grad_u = np.zeros((3, n))
        for a, z in enumerate(self.charges):
            delta = pts - self.positions[a]
            dist = np.sqrt((delta * delta).sum(axis=-1))
            u += z * dist
            safe = np.maximum(dist, 1e-30)
            on = dist <= 1e-12
            unit = np.where(on[:, None], 0.0, delta / safe[:, None])
            grad_u += z * unit.T
        self.u = u
        self.grad_u = grad_u
        self.weight = np.exp(-2.0 * u)
        self.transformed_potential = self._build_potential()

    def _build_potential(self):
        potential = -0.5 * (self.grad_u * self.grad_u).sum(axis=0)
        if self.is_identity:
            return potential
        pts = self.grid.points()
        h = self.grid.spacing
        reach = self.average_radius_cells * h
        near = np.zeros(self.grid.n_points, dtype=bool)
        for a in range(len(self.charges)):
            near |= np.sqrt(((pts - self.positions[a]) ** 2).sum(-1)) <= reach
        count = int(near.sum())
        if count == 0:
            return potential
        x, w = gauss_legendre(self.average_quadrature)
        nodes = [float(v) * 0.5 for v in x]
        weights = [float(v) * 0.5 for v in w]
        offsets = [node * h for node in nodes]
        selected = pts[near]
        mean_square = np.zeros(count)
        mean_vector = np.zeros((count, 3))
        total = 0.0
        for ix, wx in zip(offsets, weights):
            for iy, wy in zip(offsets, weights):
                for iz, wz in zip(offsets, weights):
                    weight = float(wx * wy * wz)
                    total += weight
                    shift = np.array([ix, iy, iz])
                    gradient = np.zeros_like(selected)
                    for a, z in enumerate(self.charges):
                        delta = (selected + shift) - self.positions[a]
                        dist = np.maximum(np.sqrt((delta * delta).sum(-1, keepdims=True)), 1e-300)
                        gradient = gradient + z * delta / dist
                    mean_square = mean_square + weight * (gradient * gradient).sum(-1)
                    mean_vector = mean_vector + weight * gradient
        potential = potential.copy()
        potential[near] = -0.5 * mean_square / total
        self.grad_u = self.grad_u.copy()
        self.grad_u[:, near] = (mean_vector / total).T
        self.n_cell_averaged = count
        return potential

    def face_weights(self, accuracy: int):
        """``f^2`` at the staggered faces, box extended by ``accuracy`` on axis ``d`` and indexed
        from face ``-p`` (reference ``operators/cusp.py:168-214``)."""
        g = self.grid
        half = accuracy // 2
        out = []
        for axis in range(3):
            sizes = list(g.shape)
            sizes[axis] += accuracy
            coords = []
            for d in range(3):
                index = np.arange(sizes[d], dtype=np.float64)
                if d == axis:
                    index = index - half + 0.5
                coords.append(g.origin[d] + index * g.spacing)
            mesh = np.meshgrid(*coords, indexing="ij")
            u = np.zeros(sizes)
            for a, z in enumerate(self.charges):
                sq = sum((mesh[d] - self.positions[a, d]) ** 2 for d in range(3))
                u = u + z * np.sqrt(sq)
            if u.size and float(u.max()) > 354.0:
                raise ValueError("the cusp weight exp(-2u) underflows float64 on this domain")
            out.append(np.exp(-2.0 * u))
        return tuple(out)

    def weight_integral_exact(self):
        """``integral f^2`` in closed form (reference ``operators/cusp.py:216-239``)."""
        if self.is_identity:
            return None
        if len(self.charges) == 1:
            return math.pi / self.charges[0] ** 3
        if len(self.charges) == 2 and math.isclose(self.charges[0], self.charges[1]):
            z = self.charges[0]
            r = float(np.linalg.norm(self.positions[0] - self.positions[1]))
            a = 2.0 * z * r
            return (math.pi * r**3 / 4.0) * math.exp(-a) * (2.0 * (1.0 / a + 2.0 / a**2 + 2.0 / a**3) - 2.0 / (3.0 * a))
        return None


# --------------------------------------------------------------------------------------------
# Staggered divergence form
# --------------------------------------------------------------------------------------------


def _slab(array: np.ndarray, axis: int, start: int, length: int) -> np.ndarray:
    """A view of ``length`` entries from ``start`` along ``axis``."""
    index = [slice(None)] * array.ndim
    index[axis] = slice(start, start + length)
    return array[tuple(index)]


def _axis_term(box: np.ndarray, weight: np.ndarray, ax: int, coeffs, scale: float, accuracy: int) -> np.ndarray:
    """``D_ax^T [w D_ax box]`` along array axis ``ax`` (``weight`` extended by ``accuracy`` along it)."""
    half = accuracy // 2
    extent = box.shape[ax]
    shape = list(box.shape)
    shape[ax] = extent + 2 * accuracy
    padded = np.zeros(shape)
    _slab(padded, ax, accuracy, extent)[...] = box
    length = extent + accuracy
    fshape = list(box.shape)
    fshape[ax] = length
    faces = np.empty(fshape)
    tmp = np.empty(fshape)
    for index, c in enumerate(coeffs):
        k = index - half
        term = _slab(padded, ax, accuracy - half + 1 + k, length)
        if index == 0:
            np.multiply(term, c * scale, out=faces)
        else:
            np.multiply(term, c * scale, out=tmp)
            np.add(faces, tmp, out=faces)
    np.multiply(faces, weight, out=faces)
    term_sum = np.empty(box.shape)
    part = np.empty(box.shape)
    for index, c in enumerate(coeffs):
        k = index - half
        term = _slab(faces, ax, half - 1 - k, extent)
        if index == 0:
            np.multiply(term, c * scale, out=term_sum)
        else:
            np.multiply(term, c * scale, out=part)
            np.add(term_sum, part, out=term_sum)
    return term_sum


def weighted_divergence(box: np.ndarray, face_weights, spacing: float, accuracy: int) -> np.ndarray:
    """``sum_d D_d^T [w_d D_d phi]`` with a zero halo, box-shaped ``(..., n0, n1, n2)``.

    Tap and term order of ``_weighted_divergence_eager`` (reference ``operators/divergence.py:
    84-150,172-197``): per axis pad by ``accuracy``, face ``n`` (from ``-p``) reads node
    ``n + 1 + k``; interior ``m`` reads face ``m - 1 - k``. Each sum is accumulated term by term in
    tap order (``acc = acc + term * (c / h)``, every product and sum rounded once), and the axes in
    order 0, 1, 2. With host threads (:func:`set_threads`) each axis's term is computed in slices
    along another spatial axis, which the axis's stencil does not couple: every element still gets
    the same operations in the same order.
    """
    coeffs = staggered_coefficients(accuracy)
    scale = 1.0 / spacing
    box = np.asarray(box, dtype=np.float64)
    nd = box.ndim
    out = None
    for axis in range(3):
        ax = nd - 3 + axis
        if _POOL is None:
            term = _axis_term(box, face_weights[axis], ax, coeffs, scale, accuracy)
        else:
            split = nd - 3 + (1 if axis == 0 else 0)  # a spatial axis this stencil does not run along
            wsplit = split - (nd - 3)
            bounds = np.linspace(0, box.shape[split], min(_POOL_THREADS, box.shape[split]) + 1).astype(int)
            term = np.empty(box.shape)

            def part(i, ax=ax, axis=axis, split=split, wsplit=wsplit, term=term):
                lo, hi = bounds[i], bounds[i + 1]
                _slab(term, split, lo, hi - lo)[...] = _axis_term(
                    _slab(box, split, lo, hi - lo), _slab(face_weights[axis], wsplit, lo, hi - lo), ax, coeffs, scale,
                    accuracy)

            list(_POOL.map(part, range(len(bounds) - 1)))
        if out is None:
            out = term
        else:
            np.add(out, term, out=out)
    return out


class Measure:
    """``<a,b> = sum q a b`` with ``q = weight * h^3`` (reference ``eigen/measure.py:18-96``)."""

    def __init__(self, grid: Grid, weight=None):
        self.grid = grid
        self.weight = weight
        self.quadrature = None if weight is None else np.asarray(weight, dtype=np.float64) * grid.volume_element

    def inner(self, a, b):
        product = np.asarray(a) * np.asarray(b)
        if self.quadrature is None:
            return product.sum(axis=-1) * self.grid.volume_element
        return (product * self.quadrature).sum(axis=-1)

    def norm(self, a):
        return np.sqrt(np.maximum(self.inner(a, a), 0.0))

    def gram(self, block):
        flat = np.asarray(block).reshape(-1, self.grid.n_points)
        if self.quadrature is None:
            return (flat @ flat.T) * self.grid.volume_element
        return (flat * self.quadrature) @ flat.T

    def cross(self, left, right):
        a = np.asarray(left).reshape(-1, self.grid.n_points)
        b = np.asarray(right).reshape(-1, self.grid.n_points)
        if self.quadrature is None:
            return (a @ b.T) * self.grid.volume_element
        return (a * self.quadrature) @ b.T


class CuspHamiltonian:
    """``A phi = (1/2) invw * sum_d D_d^T[f^2_face D_d phi] + V phi`` (reference
    ``operators/hamiltonian.py:178-318``).

    Without a quadrature ``invw = 1/f^2``, ``V = W`` and the measure is ``f^2``; with one,
    ``invw = h^3/omega``, ``V = W~ + extra`` and the measure is ``omega/h^3``.
    """

    def __init__(self, grid: Grid, factor: CuspFactor, extra_potential=None, quadrature=None):
        self.grid = grid
        self.factor = factor
        self.quadrature = quadrature
        if quadrature is not None:
            self.bare_potential, self.inverse_weight, self.measure = quadrature.bare_operator_terms()
        else:
            self.bare_potential = factor.transformed_potential
            self.inverse_weight = 1.0 / factor.weight
            self.measure = Measure(grid, factor.weight)
        self.potential = self.bare_potential if extra_potential is None else self.bare_potential + extra_potential
        self.face = factor.face_weights(grid.fd_order)
        self.apply_count = 0
        self.uses_divergence_form = True

    def kinetic_apply(self, phi):
        box = self.grid.to_box(phi)
        div = weighted_divergence(box, self.face, self.grid.spacing, self.grid.fd_order)
        return (div.reshape(*np.shape(phi)[:-1], -1) * 0.5) * self.inverse_weight

    def apply_with_potential(self, phi, potential):
        return self.kinetic_apply(phi) + phi * potential

    def apply(self, phi):
        self.apply_count += 1
        return self.apply_with_potential(phi, self.potential)

    def spectral_bounds(self, steps=12, safety=1.02, generator=None):
        from .eigen import lanczos_bounds, seeded_randn

        start = seeded_randn(self.grid.n_points, generator)
        return lanczos_bounds(self.apply, start, self.measure.inner, self.measure.norm, steps, safety)
