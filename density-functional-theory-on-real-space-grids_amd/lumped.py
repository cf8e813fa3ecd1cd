"""Lumped far mass weights of an overlapping geometry (reference
``src/cdft/operators/quadrature.py:421-497``).

The midpoint rule of ``far f^2`` on the lattice refined ``far_refine`` times per axis is evaluated on
the device. It is moved to the grid by three one-dimensional transposed degree-7 Lagrange
interpolations (z, then y, then x, each summing its fine points in increasing order), then blended
back to the plain weights ``h^3 far f^2`` within ``far_edge_margin_in_spacings`` of a box face. The
one-dimensional Lagrange tables (a few thousand numbers) are built on the host with the reference's
formula and operation order.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from .ops import stream_of


def axis_table(origin: float, h: float, n: int, refine: int, degree: int):
    """Fine coordinates, first coarse index, the (n_fine, 8) Lagrange weights and the contiguous fine
    range of every coarse index, for one axis (reference ``quadrature.py:447-466``)."""
    k = degree + 1
    half = degree // 2
    n_fine = n * refine
    fine = origin - 0.5 * h + h * (np.arange(n_fine, dtype=np.float64) + 0.5) / refine
    fractional = (fine - origin) / h
    first = np.clip(np.floor(fractional).astype(np.int64) - half, 0, n - k)
    t = fractional - first.astype(np.float64)
    x = t[:, None] - np.arange(k, dtype=np.float64)[None, :]
    w = np.ones_like(x)
    for j in range(k):
        for m in range(k):
            if m != j:
                w[:, j] = w[:, j] * x[:, m] / float(j - m)
    coarse = np.arange(n)
    lo = np.searchsorted(first, coarse - (k - 1), side="left")
    hi = np.searchsorted(first, coarse, side="right")
    ranges = np.stack([lo, hi], axis=1).astype(np.int32)
    return fine, first.astype(np.int32), w, ranges


def lumped_far_weights(quadrature) -> torch.Tensor:
    """``plain + interior (lumped - plain)`` on the grid of ``quadrature`` (a device
    ``CuspQuadrature`` whose spheres overlap)."""
    grid = quadrature.grid
    h = grid.spacing
    refine = int(quadrature.far_refine)
    shape = tuple(int(n) for n in grid.shape)
    dev = grid.device
    s = stream_of(quadrature.plain_far)
    tables = [axis_table(float(grid.origin[a]), h, shape[a], refine, quadrature.degree) for a in range(3)]
    as_dev = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt).to(dev)  # noqa: E731
    fine = [as_dev(t[0], torch.float64) for t in tables]
    first = [as_dev(t[1], torch.int32) for t in tables]
    weights = [as_dev(t[2], torch.float64) for t in tables]
    ranges = [as_dev(t[3], torch.int32) for t in tables]
    fshape = [n * refine for n in shape]
    values = torch.empty(fshape, dtype=torch.float64, device=dev)
    fine_shape_c = (ctypes.c_int * 3)(*fshape)
    n_spheres = len(quadrature.spheres)
    _lib.call("cdft_far_fine_f64", n_spheres, quadrature.factor.c_charges, ctypes.addressof(quadrature._spheres_c),
              ctypes.addressof(fine_shape_c), fine[0].data_ptr(), fine[1].data_ptr(), fine[2].data_ptr(),
              (h / refine) ** 3, values.data_ptr(), s)
    current, dims = values, list(fshape)
    for axis in (2, 1, 0):
        nxt_dims = list(dims)
        nxt_dims[axis] = shape[axis]
        out = torch.empty(nxt_dims, dtype=torch.float64, device=dev)
        dims_c = (ctypes.c_int * 3)(*dims)
        _lib.call("cdft_interp_transpose_axis_f64", current.data_ptr(), ctypes.addressof(dims_c), axis, shape[axis],
                  ranges[axis].data_ptr(), first[axis].data_ptr(), weights[axis].data_ptr(), out.data_ptr(), s)
        current, dims = out, nxt_dims
    lumped = current.reshape(-1)
    result = torch.empty_like(lumped)
    margin = quadrature.far_edge_margin_in_spacings * h
    width = quadrature.far_edge_width_in_spacings * h
    _lib.call("cdft_far_blend_f64", grid.c_shape, h, grid.c_origin, margin, width, quadrature.plain_far.data_ptr(),
              lumped.data_ptr(), result.data_ptr(), s)
    return result
