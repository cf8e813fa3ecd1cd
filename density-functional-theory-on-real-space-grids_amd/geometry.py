"""Grid, Kato cusp factor and cusp-aware quadrature, with every per-point field built on the GPU.

Mirrors the reference's ``UniformGrid`` (``src/cdft/grid.py:489-638``, box domain), ``CuspFactor``
(``src/cdft/operators/cusp.py:46-214``) and ``CuspQuadrature`` (``src/cdft/operators/quadrature.py``).
The only host work is the geometry-sized bookkeeping the reference also does on the host: the grid
rules, the Gauss-Legendre tables, the sphere-rule node coordinates and the binning of nodes by cell
that makes the transposed interpolation a deterministic gather. Interpolation ``L`` and its
transpose are matrix-free kernels (no CSR operator is stored).
"""

from __future__ import annotations

import ctypes
import functools
import math
from dataclasses import dataclass

import numpy as np
import torch
from numpy.polynomial.legendre import leggauss

from . import _lib, ops
from .ops import host_doubles, stream_of
from .stencil import central_coefficients, staggered_coefficients


@functools.lru_cache(maxsize=None)
def gauss_legendre(order: int) -> tuple[np.ndarray, np.ndarray]:
    """The ``order``-point Gauss-Legendre rule on [-1, 1] (numpy's Golub-Welsch + Newton polish),
    computed once per order (read-only arrays)."""
    x, w = leggauss(int(order))
    x.setflags(write=False)
    w.setflags(write=False)
    return x, w


def _ints(values) -> ctypes.Array:
    return (ctypes.c_int * len(values))(*[int(v) for v in values])


class UniformGrid:
    """Full-box uniform grid, flat C order with z fastest (reference ``grid.py:489-551``)."""

    def __init__(self, shape, spacing, origin, device, fd_order=8, gradient_order=8):
        self.shape = tuple(int(n) for n in shape)
        self.spacing = float(spacing)
        self.origin = tuple(float(o) for o in origin)
        self.device = torch.device(device)
        self.fd_order = int(fd_order)
        self.gradient_order = int(gradient_order)
        self.n_points = self.shape[0] * self.shape[1] * self.shape[2]
        self.volume_element = self.spacing**3
        self._shape_c = _ints(self.shape)
        self._origin_c = host_doubles(self.origin)
        taps = central_coefficients(1, self.gradient_order)
        self._grad_taps = torch.tensor([c * (1.0 / self.spacing) for c in taps], dtype=torch.float64,
                                       device=self.device)

    @classmethod
    def from_box(cls, box_lengths, spacing, device, anchor=None, fd_order=8):
        """``UniformGrid.from_config`` for a box domain with a zero halo (reference ``grid.py:580-605``):
        ``round(L/h) + 1`` points per edge, ``h`` re-derived from the first edge, the origin snapped
        so the coordinate origin (or ``anchor``) is a lattice point."""
        h = float(spacing)
        counts = tuple(max(int(round(length / h)) + 1, 1) for length in box_lengths)
        h = box_lengths[0] / (counts[0] - 1) if counts[0] > 1 else h
        if anchor is None:
            origin = tuple(-h * round(0.5 * length / h) for length in box_lengths)
        else:
            origin = tuple(float(a) - h * round((float(a) + 0.5 * length) / h) for a, length in zip(anchor, box_lengths))
        return cls(counts, h, origin, device, fd_order=fd_order, gradient_order=fd_order)

    @property
    def c_shape(self):
        return ctypes.addressof(self._shape_c)

    @property
    def c_origin(self):
        return ctypes.addressof(self._origin_c)

    def empty(self, *lead: int) -> torch.Tensor:
        return torch.empty((*lead, self.n_points), dtype=torch.float64, device=self.device)

    def zeros(self, *lead: int) -> torch.Tensor:
        return torch.zeros((*lead, self.n_points), dtype=torch.float64, device=self.device)

    @property
    def gradient_taps(self) -> torch.Tensor:
        """The central first-derivative taps divided by h, on the device."""
        return self._grad_taps

    def gradient(self, field: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """Order-``gradient_order`` central gradient, ``(nf, 3, N)`` for ``(nf, N)`` input."""
        f = field.reshape(-1, self.n_points)
        nf = f.shape[0]
        out = torch.empty((nf, 3, self.n_points), dtype=torch.float64, device=self.device) if out is None else out
        _lib.call("cdft_gradient_f64", f.data_ptr(), nf, self.c_shape, self.spacing, self.c_origin,
                  self._grad_taps.data_ptr(), self.gradient_order, out.data_ptr(), stream_of(f))
        return out

    def describe(self) -> dict:
        return {"shape": self.shape, "spacing_bohr": self.spacing, "origin_bohr": self.origin,
                "n_points": self.n_points, "fd_order": self.fd_order}


class CuspFactor:
    """The Kato factor on the grid: ``u``, ``grad u`` (3, N), ``f^2`` and the cell-averaged ``W``,
    plus the staggered face weights, all computed by HIP kernels (reference ``cusp.py:46-214``)."""

    average_radius_cells = 2.0
    average_quadrature = 4

    def __init__(self, grid: UniformGrid, charges, positions):
        self.grid = grid
        self.charges = tuple(float(z) for z in charges)
        self.positions = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
        if len(self.charges) > 8:
            raise ValueError("at most 8 nuclei")
        self.is_identity = len(self.charges) == 0
        self._charges_c = host_doubles(self.charges or [0.0])
        self._positions_c = host_doubles(self.positions.reshape(-1).tolist() or [0.0])
        n = grid.n_points
        dev = grid.device
        self.u = torch.empty(n, dtype=torch.float64, device=dev)
        self.grad_u = torch.empty((3, n), dtype=torch.float64, device=dev)
        self.weight = torch.empty(n, dtype=torch.float64, device=dev)
        self.transformed_potential = torch.empty(n, dtype=torch.float64, device=dev)
        x, w = gauss_legendre(self.average_quadrature)
        nodes = [float(v) * 0.5 for v in x]
        weights = [float(v) * 0.5 for v in w]
        offsets = [node * grid.spacing for node in nodes]
        products, total = [], 0.0
        for wx in weights:
            for wy in weights:
                for wz in weights:
                    weight = float(wx * wy * wz)
                    total += weight
                    products.append(weight)
        offsets_c = host_doubles(offsets)  # held until the call returns
        products_c = host_doubles(products)
        _lib.call("cdft_cusp_grid_fields_f64", len(self.charges), ctypes.addressof(self._charges_c),
                  ctypes.addressof(self._positions_c), grid.c_shape, grid.spacing, grid.c_origin,
                  self.average_radius_cells * grid.spacing, ctypes.addressof(offsets_c),
                  ctypes.addressof(products_c), total, self.u.data_ptr(), self.grad_u.data_ptr(),
                  self.weight.data_ptr(), self.transformed_potential.data_ptr(), stream_of(self.u))
        self._faces: dict[int, tuple] = {}

    @property
    def c_charges(self):
        return ctypes.addressof(self._charges_c)

    @property
    def c_positions(self):
        return ctypes.addressof(self._positions_c)

    def face_weights(self, order: int):
        """``f^2`` at the staggered faces; raises if ``exp(-2u)`` would underflow (``u > 354``)."""
        if order in self._faces:
            return self._faces[order]
        g = self.grid
        n0, n1, n2 = g.shape
        dev = g.device
        w0 = torch.empty((n0 + order, n1, n2), dtype=torch.float64, device=dev)
        w1 = torch.empty((n0, n1 + order, n2), dtype=torch.float64, device=dev)
        w2 = torch.empty((n0, n1, n2 + order), dtype=torch.float64, device=dev)
        umax = torch.zeros(1, dtype=torch.int64, device=dev)
        _lib.call("cdft_face_weights_f64", len(self.charges), self.c_charges, self.c_positions, g.c_shape, g.spacing,
                  g.c_origin, order, w0.data_ptr(), w1.data_ptr(), w2.data_ptr(), umax.data_ptr(), stream_of(w0))
        largest = float(umax.cpu().view(torch.float64)[0])  # geometry build: one read
        if largest > 354.0:
            raise ValueError(
                f"the cusp weight exp(-2u) underflows float64 on this domain: max u = {largest:.1f} at the "
                "box corner, and the limit is 354. Reduce the box or the nuclear charge.")
        self._faces[order] = (w0, w1, w2)
        return self._faces[order]

    def weight_integral_exact(self):
        """``integral f^2`` in closed form (reference ``cusp.py:216-239``)."""
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


# ------------------------------------------------------------------------------------------------
# Cusp-aware quadrature
# ------------------------------------------------------------------------------------------------


@functools.lru_cache(maxsize=None)
def sphere_directions(n_theta: int) -> tuple[np.ndarray, np.ndarray]:
    """Unit directions (GL in cos theta x 2 n_theta uniform phi) and their solid-angle weights,
    computed once per ``n_theta`` (read-only arrays)."""
    n_phi = 2 * n_theta
    cos_nodes, cos_weights = gauss_legendre(n_theta)
    phis = 2.0 * math.pi * (np.arange(n_phi) + 0.5) / n_phi
    sin_theta = np.sqrt(np.clip(1.0 - cos_nodes**2, 0.0, 1.0))
    directions = np.stack([np.outer(sin_theta, np.cos(phis)).ravel(), np.outer(sin_theta, np.sin(phis)).ravel(),
                           np.repeat(cos_nodes, n_phi)], axis=1)
    weights = np.repeat(cos_weights, n_phi) * (2.0 * math.pi / n_phi)
    directions.setflags(write=False)
    weights.setflags(write=False)
    return directions, weights


@functools.lru_cache(maxsize=None)
def _direction_rows(n_theta: int) -> np.ndarray:
    """:func:`sphere_directions` as three contiguous rows ``(3, n)`` (read-only)."""
    rows = np.ascontiguousarray(sphere_directions(n_theta)[0].T)
    rows.setflags(write=False)
    return rows


def sphere_node_rows(centre, radius_sphere, spacing, radial_order, node_spacing):
    """Composite GL radius x GL(cos theta) x uniform phi nodes and bare weights ``r^2 dr dOmega``
    (reference ``quadrature.py:129-180``), the coordinates as three rows ``(3, N)``. All shells are
    built at once with the same IEEE operations per shell as a loop over panels and GL nodes; rows
    keep numpy's loops long (an ``(N, 3)`` broadcast runs a 3-element inner loop per node)."""
    n_panels = max(2, int(math.ceil(radius_sphere / spacing)))
    edges = np.linspace(0.0, radius_sphere, n_panels + 1)
    gl_nodes, gl_weights = gauss_legendre(radial_order)
    mid = (0.5 * (edges[1:] + edges[:-1]))[:, None]
    half = (0.5 * (edges[1:] - edges[:-1]))[:, None]
    r = (mid + half * gl_nodes[None, :]).ravel()
    wr = (half * gl_weights[None, :]).ravel() * r * r
    n_theta = np.maximum(4, np.ceil(math.sqrt(2.0 * math.pi) * r / (node_spacing * spacing)).astype(np.int64)).tolist()
    counts = np.fromiter((2 * n * n for n in n_theta), dtype=np.int64, count=len(n_theta))
    rows = np.concatenate([_direction_rows(n) for n in n_theta], axis=1) * np.repeat(r, counts)
    rows += np.asarray(centre, dtype=np.float64)[:, None]
    weights = np.repeat(wr, counts) * np.concatenate([sphere_directions(n)[1] for n in n_theta])
    return rows, weights


def sphere_nodes(centre, radius_sphere, spacing, radial_order, node_spacing):
    """:func:`sphere_node_rows` with the coordinates as ``(N, 3)`` (a transposed view)."""
    rows, weights = sphere_node_rows(centre, radius_sphere, spacing, radial_order, node_spacing)
    return rows.T, weights


_dir_table: dict = {"max": 0}


def _direction_table(max_theta: int) -> dict:
    """The unit directions and solid-angle weights of every direction count 4..``max_theta`` (those of
    :func:`sphere_directions`), concatenated as rows x, y, z, w with the start of each count at
    ``offset[n]``; built once per process and grown when a larger count is asked for."""
    if max_theta > _dir_table["max"]:
        top = max(max_theta, 2 * _dir_table["max"], 48)
        offset = np.zeros(top + 1, dtype=np.int64)
        parts, weights = [], []
        at = 0
        for n in range(4, top + 1):
            offset[n] = at
            parts.append(_direction_rows(n))
            weights.append(sphere_directions(n)[1])
            at += 2 * n * n
        rows = np.ascontiguousarray(np.concatenate(parts, axis=1))
        _dir_table.update(max=top, offset=offset, x=rows[0].copy(), y=rows[1].copy(), z=rows[2].copy(),
                          w=np.ascontiguousarray(np.concatenate(weights)))
    return _dir_table


def node_table(centres, radii, spacing, radial_order, node_spacings, origin) -> dict:
    """The quadrature's nodes, numbered in the order of their interpolation base cell (stable), so
    the nodes a grid point gathers from in the transpose are contiguous in memory (only the order of
    the node sums changes): coordinate rows ``(3, m)``, owners, bare weights, base-cell rows ``(3,
    m)`` (``floor((p - origin) / h)``), the permutation from the sphere-by-sphere enumeration, the
    cell box (``lo``, ``ext``) and the node count per cell of the box.

    The shells (radius, radial weight and direction count of each composite Gauss-Legendre node, as
    :func:`sphere_node_rows` computes them) are computed here; ``cdft_node_table`` places the nodes
    and sorts them on the host, with the same IEEE operations as :func:`sphere_node_rows`."""
    r_all, wr_all, nt_all, atom_all = [], [], [], []
    gl_nodes, gl_weights = gauss_legendre(radial_order)
    for atom, centre in enumerate(centres):
        radius_sphere, node_spacing = radii[atom], node_spacings[atom]
        n_panels = max(2, int(math.ceil(radius_sphere / spacing)))
        edges = np.linspace(0.0, radius_sphere, n_panels + 1)
        mid = (0.5 * (edges[1:] + edges[:-1]))[:, None]
        half = (0.5 * (edges[1:] - edges[:-1]))[:, None]
        r = (mid + half * gl_nodes[None, :]).ravel()
        wr = (half * gl_weights[None, :]).ravel() * r * r
        n_theta = np.maximum(4, np.ceil(math.sqrt(2.0 * math.pi) * r / (node_spacing * spacing)).astype(np.int64))
        r_all.append(r)
        wr_all.append(wr)
        nt_all.append(n_theta)
        atom_all.append(np.full(r.shape[0], atom, dtype=np.int32))
    r = np.ascontiguousarray(np.concatenate(r_all))
    wr = np.ascontiguousarray(np.concatenate(wr_all))
    n_theta = np.ascontiguousarray(np.concatenate(nt_all).astype(np.int32))
    atoms = np.ascontiguousarray(np.concatenate(atom_all))
    m = int((2 * n_theta.astype(np.int64) ** 2).sum())
    table = _direction_table(int(n_theta.max()))
    cen = np.ascontiguousarray(np.asarray(centres, dtype=np.float64).reshape(-1, 3))
    org = np.ascontiguousarray(np.asarray(origin, dtype=np.float64))
    points = np.empty((3, m), dtype=np.float64)
    owners = np.empty(m, dtype=np.int32)
    bare = np.empty(m, dtype=np.float64)
    base = np.empty((3, m), dtype=np.int32)
    perm = np.empty(m, dtype=np.int64)
    keys = np.empty(m, dtype=np.int64)
    box = np.empty(6, dtype=np.int32)
    ptr = lambda a: a.ctypes.data  # noqa: E731
    _lib.call("cdft_node_table", r.shape[0], ptr(r), ptr(wr), ptr(n_theta), ptr(atoms), ptr(cen), float(spacing),
              ptr(org), int(table["max"]), ptr(table["offset"]), ptr(table["x"]), ptr(table["y"]), ptr(table["z"]),
              ptr(table["w"]), m, ptr(points), ptr(owners), ptr(bare), ptr(base), ptr(perm), ptr(keys), ptr(box))
    lo, ext = box[:3].copy(), box[3:].copy()
    return {"point_rows": points, "owners": owners, "bare_weights": bare, "base_rows": base, "perm": perm,
            "lo": lo, "ext": ext, "counts": np.bincount(keys, minlength=int(ext.prod()))}


class CuspQuadrature:
    """The cusp-aware rule on the GPU (reference ``quadrature.py:183-497``).

    Device arrays: node coordinates, ``weights = W P`` (bare weight times partition), analytic
    ``f2``, ``grad_u``, ``transformed_potential`` at the nodes, per-node stencils (base cell and
    three Lagrange rows), the cell-binned node list for the transpose, ``far``, ``far_weights``,
    ``mass_weights``.
    """

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

    def __init__(self, grid: UniformGrid, factor: CuspFactor):
        if factor.is_identity:
            raise ValueError("the cusp-aware quadrature needs at least one nucleus")
        self.grid = grid
        self.factor = factor
        h = grid.spacing
        dev = grid.device
        positions = factor.positions
        n_atoms = len(factor.charges)
        origin = np.asarray(grid.origin)
        upper = origin + np.array([(n - 1) * h for n in grid.shape])
        margin = (self.degree // 2 + 1.5) * h
        self.taper, self.overlapping, spheres = [], [], []
        radii, node_spacings = [], []
        for atom in range(n_atoms):
            centre = positions[atom]
            edge_distance = float(np.minimum(centre - origin, upper - centre).min())
            radius = min(edge_distance - margin, self.max_radius_in_spacings * h)
            if radius <= 2.0 * h:
                raise ValueError(f"nucleus {atom} is {edge_distance:.3g} bohr from the box edge; the cusp quadrature "
                                 f"needs at least {2.0 * h + margin:.3g} bohr. Enlarge the box.")
            width = self.taper_width_in_spacings * h
            tc = radius - self.taper_margin * width
            if tc < self.min_centre_in_spacings * h:
                scale = radius / ((self.min_centre_in_spacings + self.taper_margin * self.taper_width_in_spacings) * h)
                width = self.taper_width_in_spacings * h * scale
                tc = self.min_centre_in_spacings * h * scale
            self.taper.append((tc, width))
            overlapping = False
            if n_atoms > 1:
                others = np.concatenate([positions[:atom], positions[atom + 1:]])
                overlapping = float(np.sqrt(((others - centre) ** 2).sum(-1)).min()) < 2.0 * radius
            self.overlapping.append(overlapping)
            radii.append(radius)
            node_spacings.append(self.node_spacing_overlapping if overlapping else self.node_spacing)
            spheres.append((*centre, radius, tc, width))
        self.spheres = spheres
        self._spheres_c = host_doubles([v for s in spheres for v in s])
        table = node_table([positions[a] for a in range(n_atoms)], radii, h, self.radial_order, node_spacings, origin)
        perm, lo, ext = table["perm"], table["lo"], table["ext"]
        self.n_nodes = m = perm.shape[0]
        as_dev = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt).to(dev)  # noqa: E731
        self.points = as_dev(table["point_rows"]).t().contiguous()  # (m, 3) on the device
        owner = as_dev(table["owners"], torch.int32)
        base_w = as_dev(table["bare_weights"])
        self.weights = torch.empty(m, dtype=torch.float64, device=dev)
        self.f2 = torch.empty(m, dtype=torch.float64, device=dev)
        self.grad_u = torch.empty((3, m), dtype=torch.float64, device=dev)
        self.transformed_potential = torch.empty(m, dtype=torch.float64, device=dev)
        self.partition = torch.empty(m, dtype=torch.float64, device=dev)
        cell_width = self.cell_width_in_spacings * h * math.sqrt(2.0)
        s = stream_of(self.points)
        _lib.call("cdft_node_fields_f64", self.points.data_ptr(), owner.data_ptr(), base_w.data_ptr(), m, n_atoms,
                  factor.c_charges, factor.c_positions, ctypes.addressof(self._spheres_c), cell_width,
                  self.weights.data_ptr(), self.f2.data_ptr(), self.grad_u.data_ptr(),
                  self.transformed_potential.data_ptr(), self.partition.data_ptr(), s)
        # Interpolation stencils and the cell binning of the nodes.
        self.base = torch.empty((m, 3), dtype=torch.int32, device=dev)
        self.lw = torch.empty((m, 24), dtype=torch.float64, device=dev)
        bad = torch.zeros(1, dtype=torch.int32, device=dev)
        _lib.call("cdft_node_stencils_f64", self.points.data_ptr(), m, grid.c_shape, h, grid.c_origin,
                  self.base.data_ptr(), self.lw.data_ptr(), bad.data_ptr(), s)
        # The kernel's base cells are floor((p - origin) / h) in the same IEEE operations as the host's
        # ordering key, so the binning is built from the host copy without a device read.
        dims = np.asarray(grid.shape)
        if (lo - 3 < 0).any() or (lo + ext - 1 + 4 >= dims).any():
            raise ValueError("an interpolation stencil reaches outside the grid; the sphere radius bound is wrong")
        self._base_rows = table["base_rows"]
        counts = table["counts"]  # nodes per cell of the cell box
        start = np.zeros(counts.size + 1, dtype=np.int32)
        np.cumsum(counts, out=start[1:])
        self.cell_start = as_dev(start, torch.int32)
        self.node_order = perm  # GPU node i is node perm[i] of the sphere-by-sphere enumeration
        self._cell_box = _ints([lo[0], lo[1], lo[2], ext[0], ext[1], ext[2]])
        self._build_lift_chunks(self._base_rows[2], start, lo, ext, as_dev)
        self._build_interp_chunks(counts, start, as_dev)
        # Far part of the rule.
        self.far = grid.empty()
        _lib.call("cdft_far_factor_f64", grid.c_shape, h, grid.c_origin, n_atoms, ctypes.addressof(self._spheres_c),
                  self.far.data_ptr(), s)
        self.far_weights = ops.ewise(ops.EW_SCALE, grid.empty(), self.far, s=grid.volume_element)
        self.plain_far = ops.ewise(ops.EW_MUL, grid.empty(), self.far_weights, factor.weight)
        self.far_lumped = any(self.overlapping)
        if self.far_lumped:
            from .lumped import lumped_far_weights

            self.lumped_far = lumped_far_weights(self)
        else:
            self.lumped_far = self.plain_far
        wf2 = ops.ewise(ops.EW_MUL, torch.empty(m, dtype=torch.float64, device=dev), self.weights, self.f2)
        self.mass_weights = self.lift(wf2, add=self.lumped_far)
        self._bare = None

    # --- matrix-free interpolation and its transpose ---

    #: Nodes per chunk of the transpose's column walk (one LDS staging of node records).
    lift_chunk = 64

    @property
    def _base_host(self) -> np.ndarray:
        """The nodes' base cells (``(m, 3)`` int32, node order), as the stencil kernel computes them."""
        return self._base_rows.T

    def _build_lift_chunks(self, base_z, cell_start, lo, ext, as_dev) -> None:
        """Split every z-column of the cell box into chunks of at most ``lift_chunk`` consecutive
        nodes, with each chunk's z-cell range and the offset of its (64 x height) slab."""
        c2 = int(ext[2])
        ncols = int(ext[0] * ext[1])
        col_start = np.asarray(cell_start, dtype=np.int64)[np.arange(ncols + 1) * c2]
        per_col = -(-(col_start[1:] - col_start[:-1]) // self.lift_chunk)  # chunks per column
        col_chunk = np.zeros(ncols + 1, dtype=np.int64)
        np.cumsum(per_col, out=col_chunk[1:])
        n_chunks = int(col_chunk[-1])
        # chunk j of column c starts at node col_start[c] + j * lift_chunk
        col_of = np.repeat(np.arange(ncols), per_col)
        first = col_start[col_of] + (np.arange(n_chunks) - col_chunk[col_of]) * self.lift_chunk
        last = np.minimum(first + self.lift_chunk, col_start[col_of + 1]) - 1
        b2 = base_z.astype(np.int64) - int(lo[2])
        zlo = b2[first]
        hz = b2[last] - zlo + 8
        node = np.append(first, self.n_nodes)
        off = np.zeros(n_chunks + 1, dtype=np.int64)
        np.cumsum(64 * hz, out=off[1:])
        if off[-1] >= 2**31:
            raise ValueError("the transpose's workspace exceeds 32-bit offsets")
        self._lift_chunks = n_chunks
        self._chunk_node = as_dev(node.astype(np.int32), torch.int32)
        self._chunk_zlo = as_dev(zlo.astype(np.int32), torch.int32)
        self._chunk_hz = as_dev(hz.astype(np.int32), torch.int32)
        self._chunk_off = as_dev(off.astype(np.int32), torch.int32)
        self._col_chunk = as_dev(col_chunk.astype(np.int32), torch.int32)
        self._lift_per_field = int(off[-1])

    #: Nodes per chunk of the interpolation (nodes of one base cell that share one load of its 8^3
    #: stencil block; the kernel's own bound).
    interp_chunk = 16

    def _build_interp_chunks(self, counts, cell_start, as_dev) -> None:
        """Split every non-empty cell's nodes (consecutive in node order) into chunks of at most
        ``interp_chunk``: the chunk boundaries as node offsets, ending with the node count."""
        cells = np.flatnonzero(counts)
        first = np.asarray(cell_start, dtype=np.int64)[cells]
        per_cell = -(-np.asarray(counts, dtype=np.int64)[cells] // self.interp_chunk)
        total = int(per_cell.sum())
        before = np.repeat(np.cumsum(per_cell) - per_cell, per_cell)
        starts = np.repeat(first, per_cell) + (np.arange(total) - before) * self.interp_chunk
        self._interp_chunks = total
        self._interp_node = as_dev(np.append(starts, self.n_nodes).astype(np.int32), torch.int32)

    def interpolate(self, field: torch.Tensor, out: torch.Tensor | None = None, per_node: bool = False) -> torch.Tensor:
        """Degree-7 interpolant at the nodes, ``(nf, m)`` for ``(nf, N)`` (``nf <= 4``): one load of
        each base cell's stencil block per chunk of its nodes (``per_node``: one per node, the same
        values bitwise)."""
        f = field.reshape(-1, self.grid.n_points)
        nf = f.shape[0]
        out = torch.empty((nf, self.n_nodes), dtype=torch.float64, device=self.grid.device) if out is None else out
        if per_node:
            _lib.call("cdft_interpolate_f64", f.data_ptr(), nf, self.base.data_ptr(), self.lw.data_ptr(), self.n_nodes,
                      self.grid.c_shape, self.grid.spacing, self.grid.c_origin, out.data_ptr(), stream_of(f))
        else:
            _lib.call("cdft_interpolate_chunks_f64", f.data_ptr(), nf, self._interp_node.data_ptr(), self._interp_chunks,
                      self.base.data_ptr(), self.lw.data_ptr(), self.n_nodes, self.grid.c_shape, self.grid.spacing,
                      self.grid.c_origin, out.data_ptr(), stream_of(f))
        return out.reshape(-1) if field.dim() == 1 else out

    def lift(self, node_values: torch.Tensor, add: torch.Tensor | None = None, out: torch.Tensor | None = None):
        """``L^T x`` (``+ add``), ``(nf, N)`` for ``(nf, m)`` (``nf <= 4``)."""
        x = node_values.reshape(-1, self.n_nodes)
        nf = x.shape[0]
        out = torch.empty((nf, self.grid.n_points), dtype=torch.float64, device=self.grid.device) if out is None else out
        if getattr(self, "_lift_work", None) is None:
            box = self._cell_box
            columns = int(box[3]) * int(box[4]) * 64 * (int(box[5]) + 7)
            self._lift_work = torch.empty(4 * (self._lift_per_field + columns), dtype=torch.float64,
                                          device=self.grid.device)
        _lib.call("cdft_lift_f64", x.data_ptr(), nf, self.n_nodes, self.base.data_ptr(), self.lw.data_ptr(),
                  self.grid.c_shape, self.grid.spacing, self.grid.c_origin, ctypes.addressof(self._cell_box),
                  self._lift_chunks, self._chunk_node.data_ptr(), self._chunk_zlo.data_ptr(),
                  self._chunk_hz.data_ptr(), self._chunk_off.data_ptr(), self._col_chunk.data_ptr(),
                  self._lift_per_field, _ptr_or_none(add), self._lift_work.data_ptr(), self._lift_work.numel(),
                  out.data_ptr(), stream_of(x))
        return out.reshape(-1) if node_values.dim() == 1 else out

    def stencils_at(self, points: np.ndarray) -> tuple[torch.Tensor, torch.Tensor]:
        """The degree-7 stencils (base cells, Lagrange rows) of arbitrary points, e.g. the nuclei."""
        pts = torch.as_tensor(np.ascontiguousarray(points, dtype=np.float64)).to(self.grid.device)
        p = pts.shape[0]
        base = torch.empty((p, 3), dtype=torch.int32, device=self.grid.device)
        lw = torch.empty((p, 24), dtype=torch.float64, device=self.grid.device)
        bad = torch.zeros(1, dtype=torch.int32, device=self.grid.device)
        _lib.call("cdft_node_stencils_f64", pts.data_ptr(), p, self.grid.c_shape, self.grid.spacing, self.grid.c_origin,
                  base.data_ptr(), lw.data_ptr(), bad.data_ptr(), stream_of(pts))
        if int(bad.item()):
            raise ValueError("a point's interpolation stencil leaves the grid")
        return base, lw

    def interpolate_with(self, stencils: tuple[torch.Tensor, torch.Tensor], field: torch.Tensor,
                         out: torch.Tensor | None = None) -> torch.Tensor:
        """``field`` at the points whose stencils are given (from :meth:`stencils_at`), ``(p,)``."""
        base, lw = stencils
        p = base.shape[0]
        out = torch.empty(p, dtype=torch.float64, device=self.grid.device) if out is None else out
        _lib.call("cdft_interpolate_f64", field.data_ptr(), 1, base.data_ptr(), lw.data_ptr(), p, self.grid.c_shape,
                  self.grid.spacing, self.grid.c_origin, out.data_ptr(), stream_of(field))
        return out

    def interpolate_at(self, points: np.ndarray, field: torch.Tensor) -> torch.Tensor:
        """The same degree-7 stencil at arbitrary points (the nuclei), ``(p,)``."""
        return self.interpolate_with(self.stencils_at(points), field)

    def bare_operator_terms(self):
        """``(W~, h^3/omega, q)`` of the bare transformed operator, built once (reference
        ``quadrature.py:731-745``); ``q = (omega/h^3) h^3`` is the measure's quadrature weight."""
        if self._bare is None:
            g = self.grid
            m = self.n_nodes
            dev = g.device
            # node_values * f2 first (W f2), then weights * (W f2): the reference's pairing.
            wf = ops.ewise(ops.EW_MUL, torch.empty(m, dtype=torch.float64, device=dev), self.transformed_potential, self.f2)
            node = ops.ewise(ops.EW_MUL, torch.empty(m, dtype=torch.float64, device=dev), self.weights, wf)
            gw = ops.ewise(ops.EW_MUL, g.empty(), self.factor.transformed_potential, self.factor.weight)
            grid_term = ops.ewise(ops.EW_MUL, g.empty(), self.far_weights, gw)
            tilde = self.lift(node, add=grid_term)
            potential = ops.ewise(ops.EW_DIV, g.empty(), tilde, self.mass_weights)
            inverse_weight = ops.ewise(ops.EW_SDIV, g.empty(), self.mass_weights, s=g.volume_element)
            q = ops.ewise(ops.EW_ROUNDTRIP, g.empty(), self.mass_weights, s=g.volume_element)
            self._bare = (potential, inverse_weight, q)
        return self._bare

    def describe(self) -> dict:
        return {"degree": self.degree, "n_nodes": self.n_nodes, "overlapping": self.overlapping,
                "far_rule": "mass lumped-2x, fields plain" if self.far_lumped else "plain", "taper": self.taper}


def _ptr_or_none(t):
    return None if t is None else t.data_ptr()
