"""Open-boundary Hartree solver: the Coulomb-cutoff kernel on a zero-padded box, rocFFT on the GPU.

Mirrors ``CoulombCutoffPoisson`` (reference ``src/cdft/operators/poisson.py:27-120``): padded shape
``fft_friendly(ceil(pad * n))`` per axis, cutoff radius ``R = min(P_d h)/2``, kernel
``K(G) = 4 pi (1 - cos G R)/G^2`` with ``K(0) = 2 pi R^2``. The plan and the Green's-function table
are built once per grid and spacing (and shared by later solvers of the same grid); a solve is
pad-scatter, (y, z) D2Z on the occupied x-slabs, x transform, multiply, x transform (the three as one
hand-written column kernel), (y, z) Z2D on the kept x-slabs, gather (K5).
"""

from __future__ import annotations

import ctypes
import math

import torch

from . import _lib, ops
from .ops import stream_of, workspace_key


def fft_friendly(n: int) -> int:
    """Next integer whose only prime factors are 2, 3 and 5."""
    c = n
    while True:
        r = c
        for p in (2, 3, 5):
            while r % p == 0:
                r //= p
        if r == 1:
            return c
        c += 1


def set_x_path(which: int) -> None:
    """0: the x transforms and the multiply as one column kernel (default); 1: rocFFT's strided
    transforms around a separate multiply. Both give the same potential to rounding."""
    _lib.call("cdft_poisson_set_x_path", int(which))


class _Plan:
    """hipFFT plans, padded buffers, spectrum and Green's-function table of one (grid, padding,
    spacing) on one device."""

    def __init__(self, shape, padded_shape, spacing, radius, device):
        p0, p1, p2 = padded_shape
        n0, n1, n2 = shape
        handle = ctypes.c_void_p()
        _lib.call("cdft_poisson_create", n0, n1, n2, p0, p1, p2, ctypes.addressof(handle))
        self.handle = handle
        # the transforms are pruned to the n0 occupied x-slabs (see cdft_poisson_solve_f64)
        self.padded_in = torch.zeros((n0, p1, p2), dtype=torch.float64, device=device)  # padding stays zero
        self.padded_out = torch.empty((n0, p1, p2), dtype=torch.float64, device=device)
        self.slabs = torch.zeros((p0, p1, p2 // 2 + 1, 2), dtype=torch.float64, device=device)  # slabs >= n0 stay zero
        self.spectrum = torch.empty((p0, p1, p2 // 2 + 1, 2), dtype=torch.float64, device=device)
        self.kernel = torch.empty((p0, p1, p2 // 2 + 1), dtype=torch.float64, device=device)
        _lib.call("cdft_poisson_kernel_f64", handle, spacing, radius, self.kernel.data_ptr(), stream_of(self.kernel))

    def __del__(self, _release_later=ops.release_later, _lib=_lib):
        handle = getattr(self, "handle", None)
        if handle is not None and handle.value:
            # destroying the plans frees device memory: deferred to a moment no capture is recorded
            # (the helpers are bound as defaults: module globals are gone at interpreter shutdown)
            _release_later(lambda h=handle: _lib.call("cdft_poisson_destroy", h))
            self.handle = None


#: Plans are built once per (shape, padded shape, spacing, device, stream) and shared by the solvers
#: of that key, as FFT libraries cache plans. Solvers sharing a plan share its scratch buffers; their
#: solves are ordered by the one stream they run on (:func:`ops.workspace_key`). At most
#: ``PLAN_CACHE_SIZE`` plans are cached, least recently used first out: a scan whose box changes with
#: the geometry does not keep one per step. An evicted plan is destroyed once no solver holds it.
PLAN_CACHE_SIZE = 4
_plans: dict[tuple, _Plan] = {}


class CoulombCutoffPoisson:
    def __init__(self, grid, pad_factor: float = 2.5):
        if pad_factor < 2.0:
            raise ValueError("pad_factor must be at least 2.0 for the convolution to be linear")
        self.grid = grid
        self.padded_shape = tuple(fft_friendly(int(math.ceil(pad_factor * n))) for n in grid.shape)
        self.cutoff_radius = 0.5 * min(n * grid.spacing for n in self.padded_shape)
        key = (tuple(grid.shape), self.padded_shape, float(grid.spacing), *workspace_key(grid.device))
        plan = _plans.pop(key, None)
        if plan is None:
            plan = _Plan(grid.shape, self.padded_shape, grid.spacing, self.cutoff_radius, grid.device)
        _plans[key] = plan
        while len(_plans) > PLAN_CACHE_SIZE:
            _plans.pop(next(iter(_plans)))
        self._plan = plan
        self._handle = plan.handle
        self.padded_in = plan.padded_in
        self.padded_out = plan.padded_out
        self.slabs = plan.slabs
        self.spectrum = plan.spectrum
        self.kernel = plan.kernel

    def solve(self, density: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """``v_H`` on the grid for a total density ``(N,)``."""
        if density.dim() != 1 or density.numel() != self.grid.n_points:
            raise ValueError("expected a total density of shape (n_points,)")
        out = torch.empty_like(density) if out is None else out
        _lib.call("cdft_poisson_solve_f64", self._handle, density.data_ptr(), self.padded_in.data_ptr(),
                  self.slabs.data_ptr(), self.spectrum.data_ptr(), self.padded_out.data_ptr(), self.kernel.data_ptr(),
                  out.data_ptr(), stream_of(density))
        return out

