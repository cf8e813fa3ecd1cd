"""Torch-facing wrappers of the HIP entry points: shape/dtype/device validation, then one C call.

Torch owns every buffer (caching allocator); the wrappers pass raw device pointers and the current
HIP stream of the tensors' device, so every kernel is ordered with the caller's torch work.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import torch

from . import _lib


def stream_of(tensor: torch.Tensor) -> int:
    return torch.cuda.current_stream(tensor.device).cuda_stream


def _ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()


def require(t: torch.Tensor, shape: tuple, dtype: torch.dtype, device: torch.device, name: str) -> None:
    if t.device != device:
        raise ValueError(f"{name} is on {t.device}, expected {device}")
    if t.dtype != dtype:
        raise ValueError(f"{name} has dtype {t.dtype}, expected {dtype}")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {tuple(shape)}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")


def host_doubles(values) -> ctypes.Array:
    arr = (ctypes.c_double * len(values))(*[float(v) for v in values])
    return arr


@dataclass
class StencilGeometry:
    """Device arrays the divergence-form apply reads, all in one dtype.

    ``w0, w1, w2``: face weights on the box extended by ``order`` along axis d (indexed from face
    ``-order/2``); ``invw``: ``h^3/omega`` (or ``1/f^2``); ``potential``: the diagonal ``V``;
    ``taps``: the staggered weights ``c_k / h`` as host doubles.
    """

    shape: tuple[int, int, int]
    order: int
    w0: torch.Tensor
    w1: torch.Tensor
    w2: torch.Tensor
    invw: torch.Tensor
    potential: torch.Tensor
    taps: ctypes.Array

    @property
    def dtype(self) -> torch.dtype:
        return self.w0.dtype

    @property
    def device(self) -> torch.device:
        return self.w0.device

    @property
    def n_points(self) -> int:
        n0, n1, n2 = self.shape
        return n0 * n1 * n2

    def validate(self) -> None:
        n0, n1, n2 = self.shape
        o = self.order
        require(self.w0, (n0 + o, n1, n2), self.dtype, self.device, "w0")
        require(self.w1, (n0, n1 + o, n2), self.dtype, self.device, "w1")
        require(self.w2, (n0, n1, n2 + o), self.dtype, self.device, "w2")
        require(self.invw, (n0 * n1 * n2,), self.dtype, self.device, "invw")
        require(self.potential, (n0 * n1 * n2,), self.dtype, self.device, "potential")

    def with_potential(self, potential: torch.Tensor) -> "StencilGeometry":
        return StencilGeometry(self.shape, self.order, self.w0, self.w1, self.w2, self.invw, potential, self.taps)

    def to(self, dtype: torch.dtype) -> "StencilGeometry":
        if dtype == self.dtype:
            return self
        cast = lambda t: t.to(dtype).contiguous()  # noqa: E731
        return StencilGeometry(self.shape, self.order, cast(self.w0), cast(self.w1), cast(self.w2),
                               cast(self.invw), cast(self.potential), self.taps)


def _suffix(dtype: torch.dtype) -> str:
    if dtype == torch.float64:
        return "f64"
    if dtype == torch.float32:
        return "f32"
    raise ValueError(f"unsupported dtype {dtype}; the stencil runs in float64 or float32")


def hdiv_apply(phi: torch.Tensor, geom: StencilGeometry, out: torch.Tensor | None = None) -> torch.Tensor:
    """``A phi`` for a block ``(k, N)`` (or ``(N,)``): K1 in one launch."""
    squeeze = phi.dim() == 1
    block = phi.reshape(1, -1) if squeeze else phi
    k = block.shape[0]
    require(block, (k, geom.n_points), geom.dtype, geom.device, "phi")
    if out is None:
        out = torch.empty_like(block)
    require(out, (k, geom.n_points), geom.dtype, geom.device, "out")
    n0, n1, n2 = geom.shape
    _lib.call(f"cdft_hdiv_apply_{_suffix(geom.dtype)}", block.data_ptr(), geom.w0.data_ptr(), geom.w1.data_ptr(),
              geom.w2.data_ptr(), geom.invw.data_ptr(), geom.potential.data_ptr(), out.data_ptr(), k, n0, n1, n2,
              ctypes.addressof(geom.taps), geom.order, stream_of(block))
    return out.reshape(-1) if squeeze else out


def cheb_step(y: torch.Tensor, prev: torch.Tensor | None, geom: StencilGeometry, table: torch.Tensor, step: int,
              out: torch.Tensor) -> torch.Tensor:
    """One fused recurrence step ``(A y - c y) a - b prev`` with ``(c, a, b) = table[step]``."""
    k = y.shape[0]
    n0, n1, n2 = geom.shape
    require(table, (table.shape[0], 3), torch.float64, geom.device, "table")
    _lib.call(f"cdft_cheb_step_{_suffix(geom.dtype)}", y.data_ptr(), _ptr(prev), geom.w0.data_ptr(), geom.w1.data_ptr(),
              geom.w2.data_ptr(), geom.invw.data_ptr(), geom.potential.data_ptr(), out.data_ptr(), k, n0, n1, n2,
              ctypes.addressof(geom.taps), geom.order, table.data_ptr(), int(step), stream_of(y))
    return out


def cheb_filter(x: torch.Tensor, geom: StencilGeometry, table: torch.Tensor, degree: int,
                out: torch.Tensor | None = None, work: tuple[torch.Tensor, torch.Tensor] | None = None) -> torch.Tensor:
    """The degree-``degree`` scaled Chebyshev filter (K2): ``degree`` fused launches, no host sync."""
    k = x.shape[0]
    require(x, (k, geom.n_points), geom.dtype, geom.device, "x")
    require(table, (degree, 3), torch.float64, geom.device, "table")
    out = torch.empty_like(x) if out is None else out
    if work is None:
        work = (torch.empty_like(x), torch.empty_like(x))
    n0, n1, n2 = geom.shape
    _lib.call(f"cdft_cheb_filter_{_suffix(geom.dtype)}", x.data_ptr(), out.data_ptr(), work[0].data_ptr(),
              work[1].data_ptr(), geom.w0.data_ptr(), geom.w1.data_ptr(), geom.w2.data_ptr(), geom.invw.data_ptr(),
              geom.potential.data_ptr(), k, n0, n1, n2, ctypes.addressof(geom.taps), geom.order, table.data_ptr(),
              int(degree), stream_of(x))
    return out
