"""cdft_amd: an MI355X-native Kohn-Sham hot path for the cusp-factorised real-space DFT solver.

Hand-written HIP kernels for gfx950 behind a C ABI (``libcdft_hip.so``, ``include/cdft_hip.h``),
driven from Python through ctypes with torch tensors as device buffers. Nothing here runs on the
CPU in place of a kernel: without the library every compute entry point raises.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib, ops
from .ops import StencilGeometry, host_doubles
from .stencil import staggered_coefficients

__all__ = ["load_library", "ops", "StencilGeometry", "geometry_from_arrays", "chebyshev_filter_device",
           "staggered_coefficients"]


def load_library():
    """Load ``libcdft_hip.so``; raises ``ImportError`` if it has not been built."""
    return _lib.load()


def geometry_from_arrays(shape, order, spacing, face_weights, inverse_weight, potential, device,
                         dtype=torch.float64) -> StencilGeometry:
    """Upload host arrays (face weights, ``invw``, ``V``) into a :class:`StencilGeometry`."""
    to = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64).to(device=device, dtype=dtype).contiguous()  # noqa: E731
    scale = 1.0 / spacing
    taps = host_doubles([c * scale for c in staggered_coefficients(order)])
    geom = StencilGeometry(tuple(shape), order, to(face_weights[0]), to(face_weights[1]), to(face_weights[2]),
                           to(inverse_weight).reshape(-1), to(potential).reshape(-1), taps)
    geom.validate()
    return geom


def chebyshev_filter_device(block, geom, degree, lower, upper, spectrum_min):
    """Scaled Chebyshev filter with the coefficient table built on the host and uploaded."""
    from .chefsi import scaled_chebyshev_table

    rows = scaled_chebyshev_table(degree, lower, upper, spectrum_min)
    table = torch.tensor(rows, dtype=torch.float64).to(geom.device)
    return ops.cheb_filter(block.to(geom.dtype).contiguous(), geom, table, degree)
