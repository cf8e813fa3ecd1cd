"""Torch-facing wrappers of the HIP entry points: shape/dtype/device validation, then one C call.

Torch owns every buffer (caching allocator); the wrappers pass raw device pointers and the current
HIP stream of the tensors' device, so every kernel is ordered with the caller's torch work.
"""

from __future__ import annotations

import ctypes
import threading
from dataclasses import dataclass

import torch

from . import _lib


_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)


def _device_index(device: torch.device | None) -> int:
    if device is None or device.index is None:
        return torch.cuda.current_device()
    return device.index


def stream_of(tensor: torch.Tensor) -> int:
    """The raw ``hipStream_t`` torch currently issues work on for the tensor's device."""
    if _raw_stream is not None:
        return _raw_stream(_device_index(tensor.device))
    return torch.cuda.current_stream(tensor.device).cuda_stream


def _ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()


#: The stream a graph capture of this thread was started from (``None`` outside a capture).
_capture_origin = threading.local()


def workspace_key(device: torch.device) -> tuple:
    """The key of the scratch buffers (reduction partials, Poisson plans, cached solvers) of work
    issued now on ``device``: the device and the caller's stream. Work on different streams, which may
    run at once, never shares scratch memory; work on one stream is ordered by the stream. Inside a
    capture (:func:`capture_graph`) the stream is the one the capture was started from, where the
    graph is replayed."""
    index = _device_index(device)
    origin = getattr(_capture_origin, "stream", None)
    if origin is None:
        origin = _raw_stream(index) if _raw_stream is not None else torch.cuda.current_stream(index).cuda_stream
    return (index, int(origin))


_capture_streams: dict = {}

# Captures and graph destruction exclude each other process-wide: destroying a graph (its executable
# and private memory pool) while another host thread records a capture aborts the process. Graphs
# that are dropped go to `_released` and are destroyed under the lock, when a capture starts or ends.
# Other device-side releases (FFT plans) go there too, as callables: a finaliser runs wherever the
# last reference goes, including inside a capture, where a free or a synchronisation invalidates it.
_graph_lock = threading.RLock()
_released: list = []


def release_later(release, _released=_released) -> None:
    """Run ``release()`` (a device-side free) at the next capture-free moment (:func:`capture_graph`
    or :func:`release_graphs`) instead of now. Safe to call from a finaliser, also while the
    interpreter shuts down (the list is bound at definition, not looked up in the module)."""
    _released.append(release)


class CapturedGraph:
    """A captured HIP graph. Dropping it never destroys the graph at once: the graph is handed to
    the next capture (in any thread), which destroys it while no capture is being recorded."""

    __slots__ = ("_graph",)

    def __init__(self, graph: torch.cuda.CUDAGraph):
        self._graph = graph

    def replay(self) -> None:
        self._graph.replay()

    def __del__(self, _released=_released):  # bound here: module globals are gone at interpreter exit
        graph = getattr(self, "_graph", None)
        if graph is not None:
            _released.append(graph)  # a single list append: atomic under the interpreter lock
            self._graph = None


def _destroy_released() -> None:
    """Destroy the dropped graphs and run the deferred releases; the caller holds ``_graph_lock`` and
    records no capture."""
    while _released:
        item = _released.pop()
        if callable(item):
            try:
                item()
            except Exception:  # noqa: BLE001 - a failed free must not stop the others
                pass
        del item


def release_graphs() -> None:
    """Destroy every dropped graph and run every deferred release now (waits for a capture in another
    thread to finish)."""
    with _graph_lock:
        _destroy_released()


def capture_graph(body) -> CapturedGraph:
    """``body``'s launches captured into a HIP graph, to be replayed on the current stream.

    Captures on a side stream (one per device and origin stream) without draining the device first
    (torch's ``torch.cuda.graph`` synchronises the device and empties the allocator cache before every
    capture, which leaves the GPU idle while the host records), in thread-local mode, so work that
    host threads enqueue meanwhile (pinned draws) cannot invalidate it. The graph's allocations come
    from its own private pool, as with ``torch.cuda.graph``. Scratch buffers the captured calls take
    are those of the origin stream (:func:`workspace_key`). Captures from several host threads take
    turns, and no graph is destroyed while one is recorded (:class:`CapturedGraph`)."""
    device = torch.cuda.current_device()
    key = workspace_key(torch.device("cuda", device))
    with _graph_lock:
        _destroy_released()
        stream = _capture_streams.get(key)
        if stream is not None:
            with torch.cuda.stream(stream):
                stuck = torch.cuda.is_current_stream_capturing()
            if stuck:  # a capture that failed was left open on it: never begin another there
                stream = None
        if stream is None:
            stream = _capture_streams[key] = torch.cuda.Stream(device)
        graph = torch.cuda.CUDAGraph()
        outer = getattr(_capture_origin, "stream", None)
        _capture_origin.stream = key[1]
        try:
            with torch.cuda.stream(stream):
                graph.capture_begin(capture_error_mode="thread_local")
                try:
                    body()
                finally:
                    graph.capture_end()
        finally:
            _capture_origin.stream = outer
        _destroy_released()  # graphs the body dropped wait until the capture has ended
    return CapturedGraph(graph)


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
    packed: torch.Tensor | None = None  # the geometry records of the stencil kernels (built on creation)

    def __post_init__(self) -> None:
        if self.packed is None:
            self.packed = pack_records(self)

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
        require(self.packed, (int(_lib.load().cdft_hdiv_pack_len(n0, n1, n2, o)),), self.dtype, self.device, "packed")

    def with_potential(self, potential: torch.Tensor) -> "StencilGeometry":
        """The same geometry with another diagonal potential (the records do not involve it)."""
        return StencilGeometry(self.shape, self.order, self.w0, self.w1, self.w2, self.invw, potential, self.taps,
                               self.packed)

    def to(self, dtype: torch.dtype) -> "StencilGeometry":
        if dtype == self.dtype:
            return self
        cast = lambda t: t.to(dtype).contiguous()  # noqa: E731
        return StencilGeometry(self.shape, self.order, cast(self.w0), cast(self.w1), cast(self.w2),
                               cast(self.invw), cast(self.potential), self.taps)


@dataclass
class LocalGeometry:
    """The plain operator ``-1/2 lap + V`` of the analytic model systems (reference
    ``operators/hamiltonian.py:92-125``): central second-derivative weights ``c_j / h^2`` of order
    ``order`` as host doubles (``coef``), the halo's ``boundary`` (0 = zero, 1 = odd reflection about
    walls one spacing outside the outermost points) and the diagonal ``potential``. Float64 only; the
    solver's calls that take a :class:`StencilGeometry` take this one too."""

    shape: tuple[int, int, int]
    order: int
    boundary: int
    coef: ctypes.Array
    potential: torch.Tensor

    dtype = torch.float64

    @property
    def device(self) -> torch.device:
        return self.potential.device

    @property
    def n_points(self) -> int:
        n0, n1, n2 = self.shape
        return n0 * n1 * n2

    def with_potential(self, potential: torch.Tensor) -> "LocalGeometry":
        return LocalGeometry(self.shape, self.order, self.boundary, self.coef, potential)


def local_coefficients(order: int, spacing: float) -> ctypes.Array:
    """``c_j * (1 / h^2)`` for the order-``order`` central second derivative (the reference scales each
    weight by ``1 / (h h)`` the same way, ``laplacian.py:96-104``)."""
    from .stencil import central_coefficients

    scale = 1.0 / (spacing * spacing)
    taps = central_coefficients(2, order)
    return (ctypes.c_double * len(taps))(*[c * scale for c in taps])


def _local_partials(geom: LocalGeometry, k: int) -> torch.Tensor:
    n0, n1, n2 = geom.shape
    need = int(_lib.load().cdft_local_partials_len(n0, n1, n2, k))
    key = ("local",) + workspace_key(geom.device)
    ws = _stencil_partials.get(key)
    if ws is None or ws.numel() < need:
        if ws is not None:
            _retired.append(ws)  # a captured graph may still address it
        ws = torch.empty(max(need, 4096), dtype=torch.float64, device=geom.device)
        _stencil_partials[key] = ws
    return ws


def pack_records(geom: "StencilGeometry") -> torch.Tensor:
    """The per-point geometry records (w0 at the plane's newest x-face, w1, w2, invw) over the
    face-extended (y, z) rows that the stencil kernels read each plane (``cdft_hdiv_pack_*``)."""
    n0, n1, n2 = geom.shape
    length = int(_lib.load().cdft_hdiv_pack_len(n0, n1, n2, geom.order))
    out = torch.empty(length, dtype=geom.w0.dtype, device=geom.w0.device)
    for name, t in (("w0", geom.w0), ("w1", geom.w1), ("w2", geom.w2), ("invw", geom.invw)):
        if t.dtype != out.dtype or not t.is_contiguous():
            raise ValueError(f"StencilGeometry: {name} must be contiguous and of the geometry's dtype")
    _lib.call(f"cdft_hdiv_pack_{_suffix(out.dtype)}", geom.w0.data_ptr(), geom.w1.data_ptr(), geom.w2.data_ptr(),
              geom.invw.data_ptr(), n0, n1, n2, geom.order, out.data_ptr(), stream_of(out))
    return out


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
    if isinstance(geom, LocalGeometry):
        _lib.call("cdft_local_apply_f64", block.data_ptr(), geom.potential.data_ptr(), out.data_ptr(), k, n0, n1, n2,
                  ctypes.addressof(geom.coef), geom.order, geom.boundary, stream_of(block))
        return out.reshape(-1) if squeeze else out
    _lib.call(f"cdft_hdiv_apply_{_suffix(geom.dtype)}", block.data_ptr(), geom.w0.data_ptr(), geom.w1.data_ptr(),
              geom.w2.data_ptr(), geom.invw.data_ptr(), geom.potential.data_ptr(), geom.packed.data_ptr(), out.data_ptr(), k, n0, n1, n2,
              ctypes.addressof(geom.taps), geom.order, stream_of(block))
    return out.reshape(-1) if squeeze else out


def hdiv_apply_adjoint(x: torch.Tensor, geom: StencilGeometry, out: torch.Tensor | None = None,
                       work: torch.Tensor | None = None) -> torch.Tensor:
    """``A^T x`` (the Euclidean transpose of K1: ``invw`` applied on the input) for a block ``(k, N)``."""
    squeeze = x.dim() == 1
    block = x.reshape(1, -1) if squeeze else x
    k = block.shape[0]
    require(block, (k, geom.n_points), geom.dtype, geom.device, "x")
    out = torch.empty_like(block) if out is None else out
    work = torch.empty_like(block) if work is None else work
    require(out, (k, geom.n_points), geom.dtype, geom.device, "out")
    require(work, (k, geom.n_points), geom.dtype, geom.device, "work")
    n0, n1, n2 = geom.shape
    _lib.call(f"cdft_hdiv_apply_adjoint_{_suffix(geom.dtype)}", block.data_ptr(), geom.w0.data_ptr(), geom.w1.data_ptr(),
              geom.w2.data_ptr(), geom.invw.data_ptr(), geom.potential.data_ptr(), geom.packed.data_ptr(), out.data_ptr(),
              work.data_ptr(), k,
              n0, n1, n2, ctypes.addressof(geom.taps), geom.order, stream_of(block))
    return out.reshape(-1) if squeeze else out


def hdiv_set_segment(planes: int) -> None:
    """Planes per x-segment of the marching stencil kernels (0 restores the automatic choice)."""
    _lib.call("cdft_hdiv_set_segment", int(planes))


def hdiv_set_tile(tile_y: int = 0, tile_z: int = 0) -> None:
    """Tile extents of the marching stencil kernels: the largest along y (1..16) and the width along
    z (13 or 16 for the column kernel; 16, 20, 22, 24 or 28 select the vector kernel of an order-8
    batch plan); 0 restores the automatic choice. Outputs do not depend on them."""
    _lib.call("cdft_hdiv_set_tile", int(tile_y), int(tile_z))


def hdiv_launch_shape(shape, k: int, order: int = 8) -> dict:
    """The launch the column stencil kernel uses for a ``(k, *shape)`` block (host-side, no device
    work): the transpose, the residual-norm epilogue, batched launches and orders other than 8."""
    buf = (ctypes.c_int * 6)()
    _lib.call("cdft_hdiv_launch_shape", int(shape[0]), int(shape[1]), int(shape[2]), int(k), int(order),
              ctypes.addressof(buf))
    return dict(zip(("tile_y", "tile_z", "segment", "segments", "threads", "blocks"), list(buf)))


def hdiv_set_kernel(which: int) -> None:
    """Kernel of the order-8 apply and Chebyshev steps: 0 = the vector kernel (default), 1 = the
    column kernel. Both give bitwise the same outputs."""
    _lib.call("cdft_hdiv_set_kernel", int(which))


def _elem_bytes(dtype) -> int:
    if dtype == torch.float64:
        return 8
    if dtype == torch.float32:
        return 4
    raise ValueError(f"stencil kernels take float64 or float32, not {dtype}")


def hdiv_vec_set(shape, k: int, dtype, tile_y: int, tile_z: int, segment: int) -> None:
    """Launch settings of the vector kernel for ``(k, *shape)`` blocks of ``dtype`` (``tile_y = 0``
    removes them). Outputs do not depend on them."""
    _lib.call("cdft_hdiv_vec_set", int(shape[0]), int(shape[1]), int(shape[2]), int(k), _elem_bytes(dtype),
              int(tile_y), int(tile_z), int(segment))


def hdiv_vec_shape(shape, k: int, dtype) -> dict:
    """The vector kernel's launch for ``(k, *shape)`` blocks of ``dtype`` (host-side)."""
    buf = (ctypes.c_int * 6)()
    _lib.call("cdft_hdiv_vec_shape", int(shape[0]), int(shape[1]), int(shape[2]), int(k), _elem_bytes(dtype),
              ctypes.addressof(buf))
    return dict(zip(("tile_y", "tile_z", "segment", "segments", "threads", "blocks"), list(buf)))


_stencil_partials: dict[tuple, torch.Tensor] = {}


def hdiv_resnorm(x: torch.Tensor, geom: StencilGeometry, evals: torch.Tensor, q: torch.Tensor,
                 out: torch.Tensor, tail: tuple | None = None) -> torch.Tensor:
    """``out[v] = ||A x_v - evals[v] x_v||_q`` in one stencil launch plus a small combine (float64).

    ``tail = (index, bounds, degree, table, gram, orth)`` also builds the next filter's Chebyshev
    table (as :func:`cheb_table` from ``evals[index]`` and ``bounds``) and writes ``max|gram - I|``
    to ``orth`` (as :func:`orth_error`) in the combine's launch, bitwise the separate calls."""
    k = x.shape[0]
    if geom.dtype != torch.float64:
        raise ValueError("hdiv_resnorm: float64 only")
    require(x, (k, geom.n_points), torch.float64, geom.device, "x")
    require(q, (geom.n_points,), torch.float64, geom.device, "q")
    require(out, (k,), torch.float64, geom.device, "norms")
    if evals.dtype != torch.float64 or evals.numel() < k or not evals.is_contiguous():
        raise ValueError("hdiv_resnorm: evals must hold k contiguous float64 values")
    if isinstance(geom, LocalGeometry):
        return _local_resnorm(x, geom, evals, q, out, tail)
    blocks = hdiv_launch_shape(geom.shape, k, geom.order)["blocks"]
    key = workspace_key(geom.device)
    ws = _stencil_partials.get(key)
    if ws is None or ws.numel() < blocks:
        if ws is not None:
            _retired.append(ws)  # a captured graph may still address it
        ws = torch.empty(max(blocks, 4096), dtype=torch.float64, device=geom.device)
        _stencil_partials[key] = ws
    n0, n1, n2 = geom.shape
    args = (x.data_ptr(), geom.w0.data_ptr(), geom.w1.data_ptr(), geom.w2.data_ptr(), geom.invw.data_ptr(),
            geom.potential.data_ptr(), geom.packed.data_ptr(), evals.data_ptr(), q.data_ptr(), k, n0, n1, n2,
            ctypes.addressof(geom.taps), geom.order, ws.data_ptr(), ws.numel(), out.data_ptr())
    if tail is None:
        _lib.call("cdft_hdiv_resnorm_f64", *args, stream_of(x))
    else:
        index, bounds, degree, table, gram, orth = tail
        require(bounds, (2,), torch.float64, geom.device, "bounds")
        require(table, (degree, 3), torch.float64, geom.device, "table")
        require(gram, (k, k), torch.float64, geom.device, "gram")
        require(orth, (1,), torch.float64, geom.device, "orth")
        if not 0 <= index < k:
            raise ValueError("hdiv_resnorm: the table's eigenvalue index must be below k")
        _lib.call("cdft_hdiv_resnorm_tail_f64", *args, int(index), bounds.data_ptr(), int(degree), table.data_ptr(),
                  gram.data_ptr(), orth.data_ptr(), stream_of(x))
    return out


def _local_resnorm(x, geom: LocalGeometry, evals, q, out, tail):
    k = x.shape[0]
    ws = _local_partials(geom, k)
    n0, n1, n2 = geom.shape
    if tail is None:
        index, bounds, degree, table, gram, orth = -1, None, 0, None, None, None
    else:
        index, bounds, degree, table, gram, orth = tail
        require(bounds, (2,), torch.float64, geom.device, "bounds")
        require(table, (degree, 3), torch.float64, geom.device, "table")
        require(gram, (k, k), torch.float64, geom.device, "gram")
        require(orth, (1,), torch.float64, geom.device, "orth")
        if not 0 <= index < k:
            raise ValueError("hdiv_resnorm: the table's eigenvalue index must be below k")
    _lib.call("cdft_local_resnorm_f64", x.data_ptr(), geom.potential.data_ptr(), evals.data_ptr(), q.data_ptr(), k,
              n0, n1, n2, ctypes.addressof(geom.coef), geom.order, geom.boundary, ws.data_ptr(), ws.numel(),
              out.data_ptr(), int(index), _ptr(bounds), int(degree), _ptr(table), _ptr(gram), _ptr(orth), stream_of(x))
    return out


def cheb_step(y: torch.Tensor, prev: torch.Tensor | None, geom: StencilGeometry, table: torch.Tensor, step: int,
              out: torch.Tensor) -> torch.Tensor:
    """One fused recurrence step ``(A y - c y) a - b prev`` with ``(c, a, b) = table[step]``."""
    k = y.shape[0]
    n0, n1, n2 = geom.shape
    require(table, (table.shape[0], 3), torch.float64, geom.device, "table")
    _lib.call(f"cdft_cheb_step_{_suffix(geom.dtype)}", y.data_ptr(), _ptr(prev), geom.w0.data_ptr(), geom.w1.data_ptr(),
              geom.w2.data_ptr(), geom.invw.data_ptr(), geom.potential.data_ptr(), geom.packed.data_ptr(), out.data_ptr(), k, n0, n1, n2,
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
    if isinstance(geom, LocalGeometry):
        _lib.call("cdft_local_filter_f64", x.data_ptr(), out.data_ptr(), work[0].data_ptr(), work[1].data_ptr(),
                  geom.potential.data_ptr(), k, n0, n1, n2, ctypes.addressof(geom.coef), geom.order, geom.boundary,
                  table.data_ptr(), int(degree), stream_of(x))
        return out
    _lib.call(f"cdft_cheb_filter_{_suffix(geom.dtype)}", x.data_ptr(), out.data_ptr(), work[0].data_ptr(),
              work[1].data_ptr(), geom.w0.data_ptr(), geom.w1.data_ptr(), geom.w2.data_ptr(), geom.invw.data_ptr(),
              geom.potential.data_ptr(), geom.packed.data_ptr(), k, n0, n1, n2, ctypes.addressof(geom.taps), geom.order, table.data_ptr(),
              int(degree), stream_of(x))
    return out


# ------------------------------------------------------------------------------------------------
# Reductions, rotations and small dense kernels (float64)
# ------------------------------------------------------------------------------------------------

_workspaces: dict[tuple, torch.Tensor] = {}
_retired: list[torch.Tensor] = []


def partials_for(n: int, device: torch.device) -> torch.Tensor:
    """The block-partials workspace of the deterministic reductions for length ``n`` (one per device
    and stream, :func:`workspace_key`; grown on demand through torch). A grown workspace's
    predecessor is kept alive: captured graphs may still address it."""
    need = int(_lib.load().cdft_reduction_workspace_doubles(int(n)))
    key = ("partials", *workspace_key(device))
    ws = _workspaces.get(key)
    if ws is None or ws.numel() < need:
        if ws is not None:
            _retired.append(ws)
        ws = torch.empty(max(need, 64 * 64), dtype=torch.float64, device=device)
        _workspaces[key] = ws
    return ws


def wgram(a: torch.Tensor, b: torch.Tensor | None = None, q: torch.Tensor | None = None, scale: float = 1.0,
          out: torch.Tensor | None = None, sqrt: bool = False) -> torch.Tensor:
    """``scale * sum_p q_p a_ip b_jp`` as a ``(ka, kb)`` float64 device matrix (K3)."""
    a2 = a.reshape(-1, a.shape[-1])
    b2 = a2 if b is None else b.reshape(-1, b.shape[-1])
    n = a2.shape[1]
    if b2.shape[1] != n or a2.dtype != torch.float64 or b2.dtype != torch.float64:
        raise ValueError("wgram: blocks must be float64 with equal point counts")
    if not (a2.is_contiguous() and b2.is_contiguous()):
        raise ValueError("wgram: blocks must be contiguous")
    if q is not None:
        require(q, (n,), torch.float64, a2.device, "q")
    ka, kb = a2.shape[0], b2.shape[0]
    if out is None:
        out = torch.empty((ka, kb), dtype=torch.float64, device=a2.device)
    require(out, (ka, kb), torch.float64, a2.device, "out")
    ws = partials_for(n, a2.device)
    _lib.call("cdft_wgram_f64", a2.data_ptr(), ka, b2.data_ptr(), kb, _ptr(q), n, float(scale), 1 if sqrt else 0,
              ws.data_ptr(), ws.numel(), out.data_ptr(), stream_of(a2))
    return out


def rotate(r: torch.Tensor, x: torch.Tensor, transpose: bool = False, out: torch.Tensor | None = None) -> torch.Tensor:
    """``R x`` (or ``R^T x``) for a small float64 device matrix ``R`` and a block ``x`` (K4)."""
    kin = x.shape[0]
    kout = r.shape[1] if transpose else r.shape[0]
    if (r.shape[0] if transpose else r.shape[1]) != kin:
        raise ValueError(f"rotate: R {tuple(r.shape)} does not match a block of {kin} vectors")
    n = x.shape[1]
    if out is None:
        out = torch.empty((kout, n), dtype=torch.float64, device=x.device)
    require(out, (kout, n), torch.float64, x.device, "out")
    require(x, (kin, n), torch.float64, x.device, "x")
    rr = r.contiguous()
    _lib.call("cdft_rotate_f64", rr.data_ptr(), kin, kout, 1 if transpose else 0, x.data_ptr(), out.data_ptr(), n,
              stream_of(x))
    return out


def rotate_gram(r: torch.Tensor, x: torch.Tensor, y: torch.Tensor, q: torch.Tensor | None,
                gram: torch.Tensor) -> torch.Tensor:
    """``y = r^T x`` and ``gram = sum_p q_p y_ip y_jp`` in one pass over ``y`` (bitwise ``rotate`` with
    ``transpose=True`` then ``wgram(y, None, q)``); returns ``y``."""
    k, n = x.shape
    require(r, (k, k), torch.float64, x.device, "r")
    require(y, (k, n), torch.float64, x.device, "y")
    require(gram, (k, k), torch.float64, x.device, "gram")
    if q is not None:
        require(q, (n,), torch.float64, x.device, "q")
    ws = partials_for(n, x.device)
    _lib.call("cdft_rotate_gram_f64", r.data_ptr(), k, x.data_ptr(), y.data_ptr(), _ptr(q), n, ws.data_ptr(), ws.numel(),
              gram.data_ptr(), stream_of(x))
    return y


def lincomb(out: torch.Tensor, terms: list, dev: torch.Tensor | None = None) -> torch.Tensor:
    """``out = sum_t c_t x_t``; each term is ``(x, c)`` or ``(x, c, mode, idx)`` with mode 1 multiplying
    and mode 2 dividing by the device scalar ``dev[idx]``. ``out`` may alias a term."""
    n = out.numel()
    m = len(terms)
    ptrs = (ctypes.c_void_p * m)()
    coefs = (ctypes.c_double * m)()
    modes = (ctypes.c_int * m)()
    idxs = (ctypes.c_int * m)()
    for i, term in enumerate(terms):
        x, c = term[0], term[1]
        if x.numel() != n or x.dtype != torch.float64 or not x.is_contiguous():
            raise ValueError("lincomb: every term must be a contiguous float64 array of the output's size")
        ptrs[i] = x.data_ptr()
        coefs[i] = float(c)
        modes[i] = term[2] if len(term) > 2 else 0
        idxs[i] = term[3] if len(term) > 3 else 0
    _lib.call("cdft_lincomb_f64", out.data_ptr(), m, ctypes.addressof(ptrs), ctypes.addressof(coefs),
              ctypes.addressof(modes), ctypes.addressof(idxs), _ptr(dev), n, stream_of(out))
    return out


def lincomb_dots(out: torch.Tensor, terms: list, dev: torch.Tensor, q: torch.Tensor | None, others: list,
                 dots: torch.Tensor, sqrt: bool = False) -> torch.Tensor:
    """:func:`lincomb` into ``out`` followed by ``dots[j] = sum q out others[j]`` (``others[j]`` may be
    ``out`` itself: the squared norm, ``sqrt`` takes the root) for one or two operands, in one pass;
    bitwise the separate ``lincomb`` and ``wgram`` calls. ``dots`` holds ``len(others)`` adjacent
    doubles."""
    n = out.numel()
    m = len(terms)
    ptrs = (ctypes.c_void_p * m)()
    coefs = (ctypes.c_double * m)()
    modes = (ctypes.c_int * m)()
    idxs = (ctypes.c_int * m)()
    for i, term in enumerate(terms):
        x, c = term[0], term[1]
        if x.numel() != n or x.dtype != torch.float64 or not x.is_contiguous():
            raise ValueError("lincomb_dots: every term must be a contiguous float64 array of the output's size")
        ptrs[i] = x.data_ptr()
        coefs[i] = float(c)
        modes[i] = term[2] if len(term) > 2 else 0
        idxs[i] = term[3] if len(term) > 3 else 0
    nd = len(others)
    if not 1 <= nd <= 2 or dots.numel() < nd:
        raise ValueError("lincomb_dots: one or two dot operands, with room for their results")
    for b in others:
        if b.numel() != n or b.dtype != torch.float64 or not b.is_contiguous():
            raise ValueError("lincomb_dots: dot operands must be contiguous float64 arrays of the output's size")
    if q is not None:
        require(q, (n,), torch.float64, out.device, "q")
    bs = (ctypes.c_void_p * nd)(*[b.data_ptr() for b in others])
    ws = partials_for(n, out.device)
    _lib.call("cdft_lincomb_dots_f64", out.data_ptr(), m, ctypes.addressof(ptrs), ctypes.addressof(coefs),
              ctypes.addressof(modes), ctypes.addressof(idxs), _ptr(dev), n, _ptr(q), nd, ctypes.addressof(bs),
              1 if sqrt else 0, ws.data_ptr(), ws.numel(), dots.data_ptr(), stream_of(out))
    return out


def resnorm(av: torch.Tensor, v: torch.Tensor, evals: torch.Tensor, q: torch.Tensor | None,
            out: torch.Tensor | None = None) -> torch.Tensor:
    """``sqrt(sum q (Av - e v)^2)`` per vector."""
    k, n = v.shape
    out = torch.empty(k, dtype=torch.float64, device=v.device) if out is None else out
    ws = partials_for(n, v.device)
    _lib.call("cdft_resnorm_f64", av.data_ptr(), v.data_ptr(), evals.data_ptr(), _ptr(q), k, n, ws.data_ptr(),
              ws.numel(), out.data_ptr(), stream_of(v))
    return out


def chol_inv(s: torch.Tensor, t: torch.Tensor, flag: torch.Tensor) -> None:
    k = s.shape[0]
    _lib.call("cdft_chol_inv_f64", s.data_ptr(), k, t.data_ptr(), flag.data_ptr(), stream_of(s))


def eigh(m: torch.Tensor, evals: torch.Tensor, vecs: torch.Tensor) -> None:
    k = m.shape[0]
    _lib.call("cdft_eigh_f64", m.data_ptr(), k, evals.data_ptr(), vecs.data_ptr(), stream_of(m))


def orth_error(m: torch.Tensor, out: torch.Tensor) -> None:
    _lib.call("cdft_orth_error_f64", m.data_ptr(), m.shape[0], out.data_ptr(), stream_of(m))


# ------------------------------------------------------------------------------------------------
# Pointwise kernels and scalar reductions of the Kohn-Sham step (float64)
# ------------------------------------------------------------------------------------------------

EW_MUL, EW_MUL3, EW_SDIV, EW_DIV, EW_SCALE, EW_SUMDIV, EW_SUB, EW_ROUNDTRIP, EW_FILL, EW_ADD, EW_DOT3, EW_MUL3_ROWS = range(12)


def ewise(op: int, out: torch.Tensor, a: torch.Tensor | None = None, b: torch.Tensor | None = None,
          c: torch.Tensor | None = None, s: float = 0.0) -> torch.Tensor:
    """Pointwise expression (op codes above); ``EW_DOT3`` is ``a * (b . c)`` with ``b, c`` of shape
    ``(3, len)``; ``EW_MUL3_ROWS`` is ``(a b_r) c`` for both rows of ``b`` and ``out`` (shape ``(2, len)``)."""
    rows = 2 if op == EW_MUL3_ROWS else 1
    if out.numel() % rows:
        raise ValueError("ewise: the output must hold two rows")
    length = out.numel() // rows
    for i, t in enumerate((a, b, c)):
        size = length * (3 if (op == EW_DOT3 and i > 0) else (rows if i == 1 else 1))
        if t is not None and (t.numel() != size or t.dtype != torch.float64 or not t.is_contiguous()):
            raise ValueError("ewise: operands must be contiguous float64 arrays of the output's size")
    if not out.is_contiguous():
        raise ValueError("ewise: the output must be contiguous")
    _lib.call("cdft_ewise_f64", op, _ptr(a), _ptr(b), _ptr(c), float(s), out.data_ptr(), length, stream_of(out))
    return out


def pbe(n: torch.Tensor, sigma: torch.Tensor, e: torch.Tensor | None = None, vn: torch.Tensor | None = None,
        vs: torch.Tensor | None = None):
    """Spin-unpolarised PBE at (n, sigma): ``(e, de/dn, de/dsigma)``."""
    e = torch.empty_like(n) if e is None else e
    vn = torch.empty_like(n) if vn is None else vn
    vs = torch.empty_like(n) if vs is None else vs
    for t, name in ((sigma, "sigma"), (e, "e"), (vn, "vn"), (vs, "vs")):
        require(t, tuple(n.shape), torch.float64, n.device, name)
    _lib.call("cdft_pbe_f64", n.data_ptr(), sigma.data_ptr(), n.numel(), e.data_ptr(), vn.data_ptr(), vs.data_ptr(),
              stream_of(n))
    return e, vn, vs


def gga_points(rho: torch.Tensor, grad_rho: torch.Tensor, f2: torch.Tensor, grad_u: torch.Tensor,
               w_s: torch.Tensor, w_f: torch.Tensor, e: torch.Tensor, aw: torch.Tensor, crossw: torch.Tensor,
               fluxw: torch.Tensor) -> None:
    """The pointwise sigma path (``cdft_gga_points_f64``) at ``len(rho)`` points."""
    n = rho.numel()
    for t, shape, name in ((rho, (n,), "rho"), (grad_rho, (3, n), "grad_rho"), (f2, (n,), "f2"),
                           (grad_u, (3, n), "grad_u"), (w_s, (n,), "w_s"), (w_f, (n,), "w_f"), (e, (n,), "e"),
                           (aw, (n,), "aw"), (crossw, (n,), "crossw"), (fluxw, (3, n), "fluxw")):
        require(t, shape, torch.float64, rho.device, name)
    _lib.call("cdft_gga_points_f64", rho.data_ptr(), grad_rho.data_ptr(), f2.data_ptr(), grad_u.data_ptr(),
              w_s.data_ptr(), w_f.data_ptr(), n, e.data_ptr(), aw.data_ptr(), crossw.data_ptr(), fluxw.data_ptr(),
              stream_of(rho))


def divergence(field: torch.Tensor, grid, out: torch.Tensor) -> torch.Tensor:
    """``d0 F0 + d1 F1 + d2 F2`` of a ``(3, N)`` field with the grid's gradient taps."""
    require(field, (3, grid.n_points), torch.float64, field.device, "field")
    require(out, (grid.n_points,), torch.float64, field.device, "out")
    _lib.call("cdft_divergence_f64", field.data_ptr(), grid.c_shape, grid.gradient_taps.data_ptr(),
              grid.gradient_order, out.data_ptr(), stream_of(field))
    return out


def gga_combine(a: torch.Tensor, near_div: torch.Tensor, far_div: torch.Tensor, cross: torch.Tensor, h3: float,
                out: torch.Tensor) -> torch.Tensor:
    n = out.numel()
    for t, name in ((a, "a"), (near_div, "near_div"), (far_div, "far_div"), (cross, "cross")):
        require(t, (n,), torch.float64, out.device, name)
    _lib.call("cdft_gga_combine_f64", a.data_ptr(), near_div.data_ptr(), far_div.data_ptr(), cross.data_ptr(),
              float(h3), n, out.data_ptr(), stream_of(out))
    return out


def lda_vwn(n: torch.Tensor, e: torch.Tensor | None = None, v: torch.Tensor | None = None):
    e = torch.empty_like(n) if e is None else e
    v = torch.empty_like(n) if v is None else v
    _lib.call("cdft_lda_vwn_f64", n.data_ptr(), n.numel(), e.data_ptr(), v.data_ptr(), stream_of(n))
    return e, v


def density(phi: torch.Tensor, occupations, f2: torch.Tensor | None, rho: torch.Tensor, n: torch.Tensor | None) -> None:
    k = phi.shape[0]
    occ = host_doubles(list(occupations)[:k])
    _lib.call("cdft_density_f64", phi.data_ptr(), k, phi.shape[1], _ptr(f2), ctypes.addressof(occ), rho.data_ptr(),
              _ptr(n), stream_of(phi))


def atom_combine(base: torch.Tensor, c: torch.Tensor, fields: torch.Tensor, sign: float, out: torch.Tensor) -> torch.Tensor:
    _lib.call("cdft_atom_combine_f64", base.data_ptr(), c.data_ptr(), fields.data_ptr(), fields.shape[0],
              base.numel(), float(sign), out.data_ptr(), stream_of(out))
    return out


def cheb_table(evals: torch.Tensor, index: int, bounds: torch.Tensor, degree: int, table: torch.Tensor) -> torch.Tensor:
    """The filter's ``(degree, 3)`` Chebyshev rows on the device from ``evals[index]`` (the lower
    edge, clamped into the bounds) and ``bounds = [spectrum_min, upper]`` (bitwise the host rows)."""
    require(table, (degree, 3), torch.float64, evals.device, "table")
    require(bounds, (2,), torch.float64, evals.device, "bounds")
    if evals.dtype != torch.float64 or not 0 <= index < evals.numel():
        raise ValueError("cheb_table: evals must be float64 and hold the index")
    _lib.call("cdft_cheb_table_f64", evals.data_ptr(), int(index), bounds.data_ptr(), int(degree), table.data_ptr(),
              stream_of(evals))
    return table


def pulay_coefficients(gram: torch.Tensor, slots, alpha: float, coef: torch.Tensor) -> torch.Tensor:
    """Device Pulay/DIIS coefficients from the history Gram matrix ``gram`` (``(H, H)``) for the
    active ``slots`` (oldest first): ``coef[2i] = c_i``, ``coef[2i+1] = alpha c_i``."""
    h = gram.shape[0]
    require(gram, (h, h), torch.float64, gram.device, "gram")
    if coef.numel() < 2 * len(slots):
        raise ValueError("pulay_coefficients: coef too small")
    arr = (ctypes.c_int * len(slots))(*[int(v) for v in slots])
    _lib.call("cdft_pulay_coefficients_f64", gram.data_ptr(), h, ctypes.addressof(arr), len(slots), float(alpha),
              coef.data_ptr(), stream_of(gram))
    return coef


def small_solve(a: torch.Tensor, scale: torch.Tensor, b: torch.Tensor, x: torch.Tensor) -> torch.Tensor:
    _lib.call("cdft_small_solve_f64", a.data_ptr(), scale.data_ptr(), b.data_ptr(), a.shape[0], x.data_ptr(),
              stream_of(x))
    return x


RED_DOT, RED_L1, RED_MAXDIFF = 0, 1, 2


def reduce_dots(products: list, out: torch.Tensor) -> torch.Tensor:
    """Up to four ``reduce(RED_DOT, ...)`` calls over arrays of one length in one pass: each product is
    ``(a, b, w, slot, scale, accumulate)`` (``b``, ``w`` may be None); bitwise the separate calls."""
    nd = len(products)
    if not 1 <= nd <= 4:
        raise ValueError("reduce_dots: one to four products")
    n = products[0][0].numel()
    for a, b, w, *_ in products:
        for t in (a, b, w):
            if t is not None and (t.numel() != n or t.dtype != torch.float64 or not t.is_contiguous()):
                raise ValueError("reduce_dots: operands must be contiguous float64 arrays of one length")
    ws = partials_for(max(n, 1), products[0][0].device)  # >= 4096 doubles >= nd x blocks(n)
    arr = lambda kind, vals: (kind * nd)(*vals)  # noqa: E731
    _lib.call("cdft_reduce_dots_f64", nd, arr(ctypes.c_void_p, [p[0].data_ptr() for p in products]),
              arr(ctypes.c_void_p, [_ptr(p[1]) for p in products]), arr(ctypes.c_void_p, [_ptr(p[2]) for p in products]),
              n, arr(ctypes.c_double, [float(p[4]) for p in products]), arr(ctypes.c_int, [int(bool(p[5])) for p in products]),
              arr(ctypes.c_int, [int(p[3]) for p in products]), ws.data_ptr(), ws.numel(), out.data_ptr(),
              stream_of(products[0][0]))
    return out


def reduce(mode: int, a: torch.Tensor, out: torch.Tensor, slot: int = 0, b: torch.Tensor | None = None,
           w: torch.Tensor | None = None, scale: float = 1.0, accumulate: bool = False) -> torch.Tensor:
    """``out[slot] = scale * reduction`` (``+=`` with ``accumulate``): 0 = sum a b w, 1 = sum w |a|,
    2 = max |a - b|. Deterministic two-pass, no atomics."""
    n = a.numel()
    need = int(_lib.load().cdft_reduce_workspace_doubles(n))
    ws = partials_for(max(n, 1), a.device)
    if ws.numel() < need:
        raise RuntimeError("reduction workspace too small")
    _lib.call("cdft_reduce_f64", mode, a.data_ptr(), _ptr(b), _ptr(w), n, float(scale), 1 if accumulate else 0,
              ws.data_ptr(), ws.numel(), out.data_ptr(), int(slot), stream_of(a))
    return out
