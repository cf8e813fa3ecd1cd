"""Oracle cusp-aware quadrature: sphere rules, erf partition, Becke cells, lumped far weights,
degree-7 Lagrange interpolation ``L`` and its transpose, mass and lumped weights.

Test infrastructure only. Follows ``src/cdft/operators/quadrature.py`` of the reference; the sparse
operator is a scipy CSR matrix with the same rows (``(degree+1)^3`` tensor-product weights per node).
"""

from __future__ import annotations

import math

import numpy as np
import scipy.sparse as sparse
from scipy.special import erfc

from .core import Grid, CuspFactor, Measure, gauss_legendre


def lagrange_weights(t: np.ndarray, degree: int) -> np.ndarray:
    """Lagrange basis at fractional offset ``t`` on nodes ``-half..-half+degree``
    (reference ``operators/quadrature.py:56-93``), shape ``(n, degree+1)``."""
    half = degree // 2
    offsets = np.arange(-half, -half + degree + 1, dtype=np.float64)
    x = t[:, None] - offsets[None, :]
    weights = np.ones_like(x)
    for j in range(degree + 1):
        for m in range(degree + 1):
            if m == j:
                continue
            weights[:, j] = weights[:, j] * x[:, m] / float(j - m)
    return weights


class SphereRule:
    def __init__(self, atom, centre, radius_partition, radius_sphere, points, weights, n_radial, n_angular):
        self.atom = atom
        self.centre = centre
        self.radius_partition = radius_partition
        self.radius_sphere = radius_sphere
        self.points = points
        self.weights = weights
        self.n_radial = n_radial
        self.n_angular = n_angular


def sphere_rule(atom, centre, radius_partition, radius_sphere, spacing, radial_order, node_spacing):
    """Composite GL radius x GL(cos theta) x uniform phi rule (reference ``quadrature.py:129-180``)."""
    n_panels = max(2, int(math.ceil(radius_sphere / spacing)))
    edges = np.linspace(0.0, radius_sphere, n_panels + 1)
    gl_nodes, gl_weights = gauss_legendre(radial_order)
    points, weights = [], []
    n_ang = 0
    n_rad = 0
    for lo, hi in zip(edges[:-1], edges[1:]):
        mid, half = 0.5 * (hi + lo), 0.5 * (hi - lo)
        for node, weight in zip(gl_nodes, gl_weights):
            r = mid + half * node
            wr = half * weight * r * r
            n_theta = max(4, int(math.ceil(math.sqrt(2.0 * math.pi) * r / (node_spacing * spacing))))
            n_phi = 2 * n_theta
            cos_nodes, cos_weights = gauss_legendre(n_theta)
            phis = 2.0 * math.pi * (np.arange(n_phi) + 0.5) / n_phi
            sin_theta = np.sqrt(np.clip(1.0 - cos_nodes**2, 0.0, 1.0))
            directions = np.stack(
                [np.outer(sin_theta, np.cos(phis)).ravel(), np.outer(sin_theta, np.sin(phis)).ravel(), np.repeat(cos_nodes, n_phi)],
                axis=1,
            )
            w_ang = np.repeat(cos_weights, n_phi) * (2.0 * math.pi / n_phi)
            points.append(r * directions)
            weights.append(wr * w_ang)
            n_ang = max(n_ang, directions.shape[0])
            n_rad += 1
    pts = np.concatenate(points) + centre
    return SphereRule(atom, centre, radius_partition, radius_sphere, pts, np.concatenate(weights), n_rad, n_ang)


class CuspQuadrature:
    """The cusp-aware rule for one grid and one set of nuclei (reference ``quadrature.py:183-497``)."""

    degree = 7
    taper_width_in_spacings = 1.3
    taper_margin = 4.5
    min_centre_in_spacings = 3.5
    max_radius_in_spacings = 11.5
    radial_order = 6
    node_spacing = 1.0
    node_spacing_overlapping = 0.7
    far_refine = 2
    far_edge_margin_in_spacings = 6.0
    far_edge_width_in_spacings = 1.5
    cell_width_in_spacings = 1.0

    def __init__(self, grid: Grid, factor: CuspFactor):
        self.grid = grid
        self.factor = factor
        h = grid.spacing
        positions = factor.positions
        charges = factor.charges
        n_atoms = len(charges)
        origin = np.asarray(grid.origin)
        upper = origin + np.array([(n - 1) * h for n in grid.shape])
        half = self.degree // 2
        margin = (half + 1.5) * h
        self.taper = []
        self.overlapping = []
        self.spheres = []
        for atom in range(n_atoms):
            centre = positions[atom]
            edge_distance = float(np.minimum(centre - origin, upper - centre).min())
            radius_sphere = min(edge_distance - margin, self.max_radius_in_spacings * h)
            if radius_sphere <= 2.0 * h:
                raise ValueError(f"nucleus {atom} is too close to the box edge for the cusp quadrature")
            width = self.taper_width_in_spacings * h
            taper_centre = radius_sphere - self.taper_margin * width
            if taper_centre < self.min_centre_in_spacings * h:
                scale = radius_sphere / ((self.min_centre_in_spacings + self.taper_margin * self.taper_width_in_spacings) * h)
                width = self.taper_width_in_spacings * h * scale
                taper_centre = self.min_centre_in_spacings * h * scale
            self.taper.append((taper_centre, width))
            overlapping = False
            if n_atoms > 1:
                others = np.concatenate([positions[:atom], positions[atom + 1 :]])
                nearest = float(np.sqrt(((others - centre) ** 2).sum(-1)).min())
                overlapping = nearest < 2.0 * radius_sphere
            self.overlapping.append(overlapping)
            self.spheres.append(
                sphere_rule(atom, centre, taper_centre, radius_sphere, h, self.radial_order,
                            self.node_spacing_overlapping if overlapping else self.node_spacing)
            )
        self.points = np.concatenate([s.points for s in self.spheres])
        self.owner = np.concatenate([np.full(s.points.shape[0], s.atom) for s in self.spheres])
        base_weights = np.concatenate([s.weights for s in self.spheres])
        self.n_nodes = self.points.shape[0]
        self.partition = self._partition_weights(self.points, self.owner)
        self.weights = base_weights * self.partition
        u = np.zeros(self.n_nodes)
        grad_u = np.zeros((3, self.n_nodes))
        for atom, z in enumerate(charges):
            delta = self.points - positions[atom]
            dist = np.maximum(np.sqrt((delta * delta).sum(-1)), 1e-300)
            u = u + z * dist
            grad_u = grad_u + z * (delta / dist[:, None]).T
        self.f2 = np.exp(-2.0 * u)
        self.grad_u = grad_u
        self.grad_u_squared = (grad_u * grad_u).sum(axis=0)
        self.transformed_potential = -0.5 * self.grad_u_squared
        self._build_stencils()
        grid_points = grid.points()
        self.far = self._far_factor(grid_points)
        self.far_weights = grid.volume_element * self.far
        self.plain_far = self.far_weights * factor.weight
        self.far_lumped = any(self.overlapping)
        self.lumped_far = self._lumped_far_weights() if self.far_lumped else self.plain_far
        self.mass_weights = self.lift(self.weights * self.f2) + self.lumped_far
        if (self.mass_weights <= 0.0).any():
            raise ValueError("the cusp-aware mass weights are not all positive; refine the grid")
        self._bare = None

    # --- partition ---

    def _bump(self, points, atom):
        sphere = self.spheres[atom]
        centre, width = self.taper[atom]
        s = np.sqrt(((points - sphere.centre) ** 2).sum(-1))
        value = 0.5 * erfc((s - centre) / width)
        return np.where(s < sphere.radius_sphere, value, 0.0)

    def _far_factor(self, points):
        far = np.ones(points.shape[0])
        for atom in range(len(self.spheres)):
            far = far * (1.0 - self._bump(points, atom))
        return far

    def _becke_cells(self, points):
        n_atoms = len(self.spheres)
        if n_atoms == 1:
            return np.ones((1, points.shape[0]))
        dist = np.stack([np.sqrt(((points - s.centre) ** 2).sum(-1)) for s in self.spheres])
        width = self.cell_width_in_spacings * self.grid.spacing * math.sqrt(2.0)
        cells = np.ones((n_atoms, points.shape[0]))
        for a in range(n_atoms):
            for b in range(n_atoms):
                if a != b:
                    cells[a] = cells[a] * 0.5 * erfc((dist[a] - dist[b]) / width)
        return cells / np.maximum(cells.sum(axis=0, keepdims=True), 1e-300)

    def _partition_weights(self, points, owner):
        """``P_a = Q (beta_a + sum_b beta_b [s_b >= R_b])`` (reference ``quadrature.py:384-419``)."""
        near = 1.0 - self._far_factor(points)
        cells = self._becke_cells(points)
        result = np.zeros(points.shape[0])
        for atom in range(len(self.spheres)):
            mask = owner == atom
            share = cells[atom]
            for other in range(len(self.spheres)):
                if other == atom:
                    continue
                sphere = self.spheres[other]
                outside = np.sqrt(((points - sphere.centre) ** 2).sum(-1)) >= sphere.radius_sphere
                share = share + cells[other] * outside.astype(np.float64)
            result[mask] = (share * near)[mask]
        return result

    def _lumped_far_weights(self):
        """Lumped far mass weights on a twice-refined lattice with the face blend (reference
        ``quadrature.py:421-497``)."""
        g = self.grid
        h = g.spacing
        refine = self.far_refine
        shape = g.shape
        k = self.degree + 1
        half = self.degree // 2
        matrices, fine_axes = [], []
        for axis in range(3):
            n = shape[axis]
            n_fine = n * refine
            fine = g.origin[axis] - 0.5 * h + h * (np.arange(n_fine, dtype=np.float64) + 0.5) / refine
            fine_axes.append(fine)
            fractional = (fine - g.origin[axis]) / h
            first = np.clip(np.floor(fractional).astype(np.int64) - half, 0, n - k)
            t = fractional - first.astype(np.float64)
            x = t[:, None] - np.arange(k, dtype=np.float64)[None, :]
            w = np.ones_like(x)
            for j in range(k):
                for m in range(k):
                    if m != j:
                        w[:, j] = w[:, j] * x[:, m] / float(j - m)
            mat = np.zeros((n, n_fine))
            rows = (first[:, None] + np.arange(k)[None, :]).reshape(-1)
            cols = np.repeat(np.arange(n_fine), k)
            np.add.at(mat, (rows, cols), w.reshape(-1))
            matrices.append(mat)
        n_x, n_y, n_z = shape[0] * refine, shape[1] * refine, shape[2] * refine
        xs, ys = np.meshgrid(fine_axes[0], fine_axes[1], indexing="ij")
        plane = np.stack([xs.reshape(-1), ys.reshape(-1)], axis=1)
        collapsed = np.zeros((n_x, n_y, shape[2]))
        volume = (h / refine) ** 3
        charges = self.factor.charges
        for iz in range(n_z):
            pts = np.concatenate([plane, np.full((plane.shape[0], 1), fine_axes[2][iz])], axis=1)
            far = np.ones(pts.shape[0])
            u = np.zeros(pts.shape[0])
            for atom, s in enumerate(self.spheres):
                centre_t, width = self.taper[atom]
                d = np.sqrt(((pts - s.centre) ** 2).sum(-1))
                bump = np.where(d < s.radius_sphere, 0.5 * erfc((d - centre_t) / width), 0.0)
                far = far * (1.0 - bump)
                u = u + charges[atom] * d
            values = (volume * far * np.exp(-2.0 * u)).reshape(n_x, n_y)
            collapsed = collapsed + values[:, :, None] * matrices[2][:, iz][None, None, :]
        lumped = np.einsum("ax,by,xyc->abc", matrices[0], matrices[1], collapsed).reshape(-1)
        pts = g.points()
        origin = np.asarray(g.origin)
        upper = origin + np.array([(n - 1) * h for n in shape])
        face = np.minimum((pts - origin).min(axis=1), (upper - pts).min(axis=1))
        margin = self.far_edge_margin_in_spacings * h
        width = self.far_edge_width_in_spacings * h
        interior = 0.5 * erfc((margin - face) / width)
        return self.plain_far + interior * (lumped - self.plain_far)

    # --- interpolation ---

    def _stencil(self, points):
        g = self.grid
        h = g.spacing
        origin = np.asarray(g.origin)
        fractional = (points - origin) / h
        base = np.floor(fractional)
        t = fractional - base
        half = self.degree // 2
        k = self.degree + 1
        ws, idx = [], []
        for axis in range(3):
            w = lagrange_weights(t[:, axis], self.degree)
            index = base[:, axis].astype(np.int64)[:, None] + np.arange(-half, -half + k)[None, :]
            if (index < 0).any() or (index >= g.shape[axis]).any():
                raise ValueError("an interpolation stencil reaches outside the grid")
            ws.append(w)
            idx.append(index)
        sx, sy, sz = g.shape[1] * g.shape[2], g.shape[2], 1
        cols = (idx[0][:, :, None, None] * sx + idx[1][:, None, :, None] * sy + idx[2][:, None, None, :] * sz).reshape(len(points), -1)
        vals = (ws[0][:, :, None, None] * ws[1][:, None, :, None] * ws[2][:, None, None, :]).reshape(len(points), -1)
        return cols, vals

    def _build_stencils(self):
        cols, vals = self._stencil(self.points)
        k3 = (self.degree + 1) ** 3
        crow = np.arange(0, self.n_nodes * k3 + 1, k3)
        self.operator = sparse.csr_matrix((vals.reshape(-1), cols.reshape(-1), crow), shape=(self.n_nodes, self.grid.n_points))
        self.operator_t = self.operator.T.tocsr()

    def interpolate(self, field):
        """Degree-7 interpolant at the nodes (reference ``quadrature.py:613-620``)."""
        f = np.asarray(field, dtype=np.float64)
        if f.ndim == 1:
            return self.operator @ f
        return (self.operator @ f.reshape(-1, self.grid.n_points).T).T.reshape(*f.shape[:-1], self.n_nodes)

    def interpolate_gradient(self, field):
        return self.interpolate(self.grid.gradient(field))

    def lift(self, node_values):
        """``L^T x`` (reference ``quadrature.py:628-638``)."""
        x = np.asarray(node_values, dtype=np.float64)
        if x.ndim == 1:
            return self.operator_t @ x
        return (self.operator_t @ x.reshape(-1, self.n_nodes).T).T.reshape(*x.shape[:-1], self.grid.n_points)

    def interpolate_at(self, points, field):
        cols, vals = self._stencil(np.asarray(points, dtype=np.float64))
        return (np.asarray(field)[cols] * vals).sum(axis=-1)

    def integrate(self, node_values, grid_values):
        """``sum W P g_node + sum h^3 far g_grid`` (reference ``quadrature.py:701-709``)."""
        return (self.weights * node_values).sum(axis=-1) + (self.far_weights * grid_values).sum(axis=-1)

    def lumped_weights(self, node_values, grid_values):
        return self.lift(self.weights * node_values) + self.far_weights * grid_values

    def lift_gradient(self, node_vectors):
        """``-div_FD (L^T X)`` for a ``(3, m)`` node field (reference ``quadrature.py:640-650``)."""
        lifted = self.lift(node_vectors)
        divergence = np.zeros(self.grid.n_points)
        for d in range(3):
            divergence = divergence + self.grid.partial_derivative(lifted[d], d)
        return -divergence

    def lumped_gradient_weights(self, node_vectors, grid_vectors):
        """``lift_gradient(W X_node) - h^3 div_FD(far X_grid)`` (reference ``quadrature.py:718-729``)."""
        near = self.lift_gradient(self.weights * node_vectors)
        weighted = self.far[None, :] * grid_vectors
        divergence = np.zeros(self.grid.n_points)
        for d in range(3):
            divergence = divergence + self.grid.gradient(weighted[d])[d]
        return near - self.grid.volume_element * divergence

    def effective_potential(self, node_values, grid_values):
        tilde = self.lumped_weights(node_values * self.f2, grid_values * self.factor.weight)
        return tilde / self.mass_weights

    def bare_operator_terms(self):
        """``(W~/omega... , h^3/omega, measure)`` (reference ``quadrature.py:731-745``)."""
        if self._bare is None:
            potential = self.effective_potential(self.transformed_potential, self.factor.transformed_potential)
            inverse_weight = self.grid.volume_element / self.mass_weights
            measure = Measure(self.grid, self.mass_weights / self.grid.volume_element)
            self._bare = (potential, inverse_weight, measure)
        return self._bare

    def density_gradient_at_nodes(self, rho):
        rho_node = self.interpolate(rho)
        grad_rho = self.interpolate_gradient(rho)
        n = self.f2 * rho_node
        grad_n = self.f2[None, :] * (grad_rho - 2.0 * rho_node[None, :] * self.grad_u)
        return n, grad_n


def hydrogenic_1s_potential(charge, distance):
    """Potential of ``exp(-2 Z s)`` with its small-``s`` series (reference ``quadrature.py:805-820``)."""
    z = float(charge)
    distance = np.asarray(distance, dtype=np.float64)
    x = z * distance
    small = x < 1.0e-3
    safe = np.maximum(distance, 1e-300)
    full = (math.pi / z**3) * (1.0 - np.exp(-2.0 * x) * (1.0 + x)) / safe
    series = (math.pi / z**2) * (1.0 - (2.0 / 3.0) * x**2 + (2.0 / 3.0) * x**3 - 0.4 * x**4 + (8.0 / 45.0) * x**5)
    return np.where(small, series, full)
