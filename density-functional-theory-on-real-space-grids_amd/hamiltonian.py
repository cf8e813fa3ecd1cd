"""The cusp-factorised Hamiltonian on the GPU, exposed only through its action on a block.

Mirrors ``CuspFactoredHamiltonian`` (reference ``src/cdft/operators/hamiltonian.py:178-396``):
``A phi = (1/2)(h^3/omega) sum_d D_d^T[f^2_face D_d phi] + (W~ + v~) phi``, self-adjoint in the
measure ``q = omega``. ``apply`` is one K1 launch; the Lanczos bounds run on the device with all
recurrence scalars kept in one device buffer and read once (reference ``_lanczos_bounds`` 29-89).
"""

from __future__ import annotations

import numpy as np
import torch

from . import ops
from .ops import StencilGeometry, host_doubles
from .stencil import staggered_coefficients


def seeded_randn(shape, generator, device) -> torch.Tensor:
    """Standard-normal draw on the generator's own (CPU) device, then moved (reference
    ``precision.py:246-263``), so the GPU solve starts from the reference's vectors. ``generator``
    may be a :class:`SeededDraws`, which serves repeated draw sequences from device memory."""
    size = (shape,) if isinstance(shape, int) else tuple(shape)
    if isinstance(generator, SeededDraws):
        return generator.draw(size, device)
    draw = torch.randn(size, generator=generator, dtype=torch.float64)
    return draw.to(device)


class SeededDraws:
    """A freshly seeded CPU generator whose draws are memoised by their position in the draw
    sequence. Every diagonalisation of one self-consistent solve re-seeds the same generator, so
    its draws repeat exactly; the memo (shared by the solve through ``store``) turns the repeats
    into device-to-device copies of the first draw instead of fresh CPU draws and uploads, and
    :func:`prefetch_draws` can fill it on a host thread ahead of use. The values are bitwise those
    of ``torch.randn`` on ``torch.Generator().manual_seed(seed)``."""

    def __init__(self, seed: int, store: dict):
        self.seed = int(seed)
        self.store = store
        self.sequence: tuple = ()
        self._gen = None

    def draw(self, size: tuple, device) -> torch.Tensor:
        key = (self.seed, self.sequence + (size,))
        hit = self.store.get(key)
        if isinstance(hit, tuple):  # prefetched on a host thread: (future, position)
            future, position = hit
            hit = future.result()[position].to(device, non_blocking=True)
            self.store[key] = hit
        if hit is None:
            if self._gen is None:
                self._gen = torch.Generator(device="cpu").manual_seed(self.seed)
                for earlier in self.sequence:  # replay the positions already served from the memo
                    torch.randn(earlier, generator=self._gen, dtype=torch.float64)
            hit = torch.randn(size, generator=self._gen, dtype=torch.float64).to(device)
            self.store[key] = hit
        elif self._gen is not None:
            torch.randn(size, generator=self._gen, dtype=torch.float64)
        self.sequence = self.sequence + (size,)
        return hit.clone()


_draw_pool = None
_chunk_pool = None
#: Host threads drawing chunks: the CPUs this process may run on, less two (the solve's own thread
#: and one for the rest), at most 14 (a GPU box's CPU share is 16).
_DRAW_THREADS = None
_CHUNK_VALUES = 1 << 16  # values per chunk; a multiple of 16
_state_layout_ok = None


def _pools():
    global _draw_pool, _chunk_pool
    if _draw_pool is None:
        import concurrent.futures
        import os

        try:
            cpus = len(os.sched_getaffinity(0))
        except (AttributeError, OSError):
            cpus = os.cpu_count() or 4
        threads = _DRAW_THREADS or max(2, min(14, cpus - 2))
        # One sequence at a time: the first one a solve needs is drawn before the later ones start.
        _draw_pool = concurrent.futures.ThreadPoolExecutor(max_workers=1, thread_name_prefix="seeded-draws")
        _chunk_pool = concurrent.futures.ThreadPoolExecutor(max_workers=threads, thread_name_prefix="draw-chunks")
    return _draw_pool, _chunk_pool


def _generator_at(seed: int, key: np.ndarray, index: int) -> torch.Generator:
    """A CPU generator whose Mersenne-Twister stream sits at the given state (the 624-word array
    and the index of its next output), in the layout ``torch.Generator.get_state`` uses: initial
    seed, then the remaining-count and seeded flag, the next index, and the words as 64-bit slots."""
    state = torch.Generator().manual_seed(int(seed)).get_state()
    q = state.numpy()[: 8 * (3 + 624)].view(np.uint64)
    q[1] = (1 << 32) | (625 - int(index))
    q[2] = int(index)
    q[3:] = key
    gen = torch.Generator()
    gen.set_state(state)
    return gen


def _states(seed: int, positions: list[int]) -> tuple[np.ndarray, np.ndarray]:
    import ctypes

    from . import _lib

    n = len(positions)
    pos = (ctypes.c_int64 * n)(*positions)
    key = np.empty((n, 624), dtype=np.uint32)
    index = np.empty(n, dtype=np.int32)
    _lib.call("cdft_mt19937_states", int(seed) & 0xFFFFFFFFFFFFFFFF, n, ctypes.addressof(pos), key.ctypes.data,
              index.ctypes.data)
    return key, index


def _layout_verified() -> bool:
    """Whether a generator placed by :func:`_generator_at` continues exactly as torch's own does
    (checked once: the state mid-stream, across a twist, and at a fresh seed)."""
    global _state_layout_ok
    if _state_layout_ok is None:
        ok = True
        for seed, skip in ((7, 0), (7, 5), (2**32 + 11, 311), (123, 312)):
            ref = torch.Generator().manual_seed(seed)
            if skip:
                torch.rand(skip, generator=ref, dtype=torch.float64)
            key, index = _states(seed, [2 * skip])
            mine = _generator_at(seed, key[0], int(index[0]))
            ok &= torch.equal(torch.randn(40, generator=ref, dtype=torch.float64),
                              torch.randn(40, generator=mine, dtype=torch.float64))
        _state_layout_ok = bool(ok)
    return _state_layout_ok


def randn_at(seed: int, jobs) -> list[torch.Tensor]:
    """For each job ``(u0, n)`` (``n >= 16``): the ``n`` float64 normals ``torch.randn(n)`` gives from a
    generator seeded ``seed`` whose next uniform is the ``u0``-th of its stream, bitwise, drawn in
    chunks on host threads into pinned memory.

    torch fills ``n >= 16`` float64 normals by drawing ``n`` uniforms (two 32-bit outputs each),
    transforming every full group of 16 on its own, and, when 16 does not divide ``n``, drawing 16
    more uniforms for the last 16 values. A chunk of whole groups drawn from the generator state at
    its first uniform is therefore bitwise that part of the serial draw; the states come from
    ``cdft_mt19937_states``. The caller checks :func:`_layout_verified` first."""
    _, chunk_pool = _pools()
    pinned = torch.cuda.is_available()
    plan = []  # (job, lo, hi, uniform offset of lo) per chunk, then the tails
    tails = []
    counts = [int(n) for _, n in jobs]
    for d, (u0, n) in enumerate(jobs):
        if n < 16:
            raise ValueError("randn_at draws 16 values or more a job")
        full = n - n % 16
        for lo in range(0, full, _CHUNK_VALUES):
            hi = min(full, lo + _CHUNK_VALUES)
            plan.append((d, lo, hi, int(u0) + lo))
        if n % 16:
            tails.append((d, int(u0) + n))
    positions = sorted({2 * u for *_, u in plan} | {2 * u for _, u in tails})
    key, index = _states(seed, positions)
    where = {p: i for i, p in enumerate(positions)}
    outs = [torch.empty(n, dtype=torch.float64, pin_memory=pinned) for n in counts]

    def chunk(d, lo, hi, u):
        i = where[2 * u]
        torch.randn(hi - lo, generator=_generator_at(seed, key[i], int(index[i])), dtype=torch.float64,
                    out=outs[d][lo:hi])

    futures = [chunk_pool.submit(chunk, *job) for job in plan]
    tail_values = []
    for d, u in tails:
        i = where[2 * u]
        tail_values.append((d, torch.randn(16, generator=_generator_at(seed, key[i], int(index[i])),
                                           dtype=torch.float64)))
    for f in futures:
        f.result()
    for d, values in tail_values:
        outs[d][counts[d] - 16:] = values
    return outs


def _consumed(n: int) -> int:
    """Uniforms a ``torch.randn(n)`` float64 draw (``n >= 16``) takes from the generator."""
    return n + (16 if n % 16 else 0)


def parallel_randn(seed: int, sizes) -> list[torch.Tensor]:
    """``[torch.randn(s, generator=g, dtype=float64) for s in sizes]`` with ``g`` freshly seeded
    ``seed``, bitwise, drawn in chunks on host threads into pinned memory (:func:`randn_at`). Sizes
    below 16 (a different torch path) or a generator layout this code does not recognise fall back
    to the serial draw."""
    sizes = [tuple(s) if not isinstance(s, int) else (s,) for s in sizes]
    counts = [int(np.prod(s)) for s in sizes]
    if min(counts, default=16) < 16 or not _layout_verified():
        gen = torch.Generator().manual_seed(int(seed))
        return [torch.randn(s, generator=gen, dtype=torch.float64) for s in sizes]
    jobs, uniform = [], 0
    for n in counts:
        jobs.append((uniform, n))
        uniform += _consumed(n)
    return [o.view(s) for o, s in zip(randn_at(seed, jobs), sizes)]


def seeded_starts(seed: int, k: int, sizes, device) -> list[tuple[torch.Tensor, torch.Tensor]]:
    """For each grid size ``n`` of ``sizes``: the ``(k, n)`` start block and the ``(n,)`` Lanczos start a
    solve draws from a fresh generator seeded ``seed`` (``torch.randn((k, n))`` then
    ``torch.randn(n)``), on ``device``, bitwise.

    Every grid's block is a draw from the start of the same stream, so all blocks share its full
    groups of 16: that common prefix (as long as the longest block's full groups) is drawn and
    uploaded once, and each grid's block copies its part on the device and adds its own tail group.
    Only the Lanczos starts, which begin at a position that depends on the size, are drawn per
    grid. Sizes whose draws fall below 16 values take the serial draw."""
    sizes = [int(n) for n in sizes]
    if not sizes:
        return []
    if min(sizes) < 16 or not _layout_verified():
        out = []
        for n in sizes:
            gen = torch.Generator().manual_seed(int(seed))
            block = torch.randn((k, n), generator=gen, dtype=torch.float64)
            start = torch.randn((n,), generator=gen, dtype=torch.float64)
            out.append((block.to(device), start.to(device)))
        return out
    blocks = [k * n for n in sizes]
    prefix = max(b - b % 16 for b in blocks)
    jobs = [(0, prefix)] if prefix >= 16 else []
    tail_of, start_of = {}, {}
    for i, (n, b) in enumerate(zip(sizes, blocks)):
        if b % 16:
            tail_of[i] = len(jobs)
            jobs.append((b, 16))
        start_of[i] = len(jobs)
        jobs.append((_consumed(b), n))
    drawn = randn_at(seed, jobs)
    common = drawn[0].to(device, non_blocking=True) if prefix >= 16 else None
    out = []
    for i, (n, b) in enumerate(zip(sizes, blocks)):
        block = torch.empty(b, dtype=torch.float64, device=device)
        full = b - b % 16
        if full:
            block[:full].copy_(common[:full])
        if i in tail_of:
            block[b - 16:].copy_(drawn[tail_of[i]].to(device, non_blocking=True))
        out.append((block.view(k, n), drawn[start_of[i]].to(device, non_blocking=True)))
    return out


def prefetch_draws(store: dict, seed: int, sequences) -> None:
    """Start drawing each sequence of sizes (from a fresh generator seeded ``seed``) on host
    threads, and register every prefix in ``store`` for :class:`SeededDraws` to pick up."""
    draw_pool, _ = _pools()
    for seq in sequences:
        seq = tuple(tuple(s) for s in seq)
        future = draw_pool.submit(parallel_randn, int(seed), seq)
        for i in range(len(seq)):
            store.setdefault((int(seed), seq[: i + 1]), (future, i))


class DeviceMeasure:
    """``<a, b> = sum_p q_p a_p b_p`` on the device (reference ``eigen/measure.py:18-96``)."""

    def __init__(self, grid, q: torch.Tensor):
        self.grid = grid
        self.q = q

    def gram(self, block: torch.Tensor, out=None) -> torch.Tensor:
        return ops.wgram(block, None, self.q, out=out)

    def cross(self, left: torch.Tensor, right: torch.Tensor, out=None) -> torch.Tensor:
        return ops.wgram(left, right, self.q, out=out)

    def inner_into(self, a: torch.Tensor, b: torch.Tensor, out: torch.Tensor) -> torch.Tensor:
        return ops.wgram(a.reshape(1, -1), b.reshape(1, -1), self.q, out=out.view(1, 1))

    def norm_into(self, a: torch.Tensor, out: torch.Tensor) -> torch.Tensor:
        return ops.wgram(a.reshape(1, -1), None, self.q, out=out.view(1, 1), sqrt=True)

    def norms(self, block: torch.Tensor) -> torch.Tensor:
        zeros = torch.zeros(block.shape[0], dtype=torch.float64, device=block.device)
        return ops.resnorm(block, block, zeros, self.q)


class CuspFactoredHamiltonian:
    """``A = f^-1 H f`` in staggered divergence form with the cusp-aware measure."""

    uses_divergence_form = True

    def __init__(self, grid, factor, quadrature, extra_potential: torch.Tensor | None = None):
        self.grid = grid
        self.factor = factor
        self.quadrature = quadrature
        bare, invw, q = quadrature.bare_operator_terms()
        self.bare_potential = bare
        self.inverse_weight = invw
        if extra_potential is None:
            potential = bare
        else:
            potential = ops.ewise(ops.EW_ADD, grid.empty(), bare, extra_potential)
        self.potential = potential
        order = grid.fd_order
        faces = factor.face_weights(order)
        taps = host_doubles([c * (1.0 / grid.spacing) for c in staggered_coefficients(order)])
        # the records depend on the faces and invw only: the operators of one quadrature share them
        records = quadrature.__dict__.setdefault("_stencil_records", {})
        held = records.get(order)
        packed = held[2] if held is not None and held[0] is faces and held[1] is invw else None
        self.geometry = StencilGeometry(grid.shape, order, faces[0], faces[1], faces[2], invw, potential, taps,
                                        packed)
        records[order] = (faces, invw, self.geometry.packed)
        self.measure = DeviceMeasure(grid, q)
        self.apply_count = 0
        self._kinetic_geometry = None

    def apply(self, phi: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        self.apply_count += 1
        return ops.hdiv_apply(phi, self.geometry, out)

    def geometry_in(self, dtype: torch.dtype) -> StencilGeometry:
        """The stencil geometry cast to ``dtype`` (cached): the float32 filter reads float32 face
        weights, ``invw`` and ``V``, halving its geometry traffic."""
        cache = self.__dict__.setdefault("_cast", {})
        if dtype not in cache:
            cache[dtype] = self.geometry.to(dtype)
        return cache[dtype]

    def note_applications(self, count: int) -> None:
        self.apply_count += int(count)

    def kinetic_apply(self, phi: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """``(1/2)(h^3/omega) sum_d D_d^T[f^2 D_d phi]`` alone (zero diagonal potential)."""
        if self._kinetic_geometry is None:
            self._kinetic_geometry = self.geometry.with_potential(torch.zeros_like(self.potential))
        return ops.hdiv_apply(phi, self._kinetic_geometry, out)

    def spectral_bounds(self, steps: int = 12, safety: float = 1.02, generator=None) -> tuple[float, float]:
        """Rigorous Lanczos enclosure in the operator's measure, widened by ``safety``."""
        g = self.grid
        start = seeded_randn(g.n_points, generator, g.device)
        return lanczos_bounds(self, start, steps, safety)


def lanczos_bounds(hamiltonian, start: torch.Tensor, steps: int, safety: float) -> tuple[float, float]:
    """Short Lanczos run with one reorthogonalisation pass against the previous two vectors; the
    alphas and betas stay on the device and are read in one transfer (reference
    ``operators/hamiltonian.py:29-89``)."""
    return lanczos_finish(lanczos_launch(hamiltonian, start, steps).cpu().numpy(), steps, safety)


def lanczos_launch(hamiltonian, start: torch.Tensor, steps: int) -> torch.Tensor:
    """The device part of :func:`lanczos_bounds`: enqueue the run and return the device array whose
    first ``2 steps`` entries are the alphas and betas (read it when the bounds are needed, so
    several runs can share one transfer)."""
    return lanczos_enqueue(lambda x, y: hamiltonian.apply(x, y), hamiltonian.measure, start, steps)


def lanczos_enqueue(apply, m, start: torch.Tensor, steps: int, scal: torch.Tensor | None = None) -> torch.Tensor:
    """The Lanczos recurrence of :func:`lanczos_launch` for an operator given by ``apply(x, y)``
    (``(1, N)`` blocks) and a measure ``m``, into ``scal`` (``2 steps + 3`` doubles; every entry is
    written before it is read). Only device work: it can be captured in a graph."""
    dev = start.device
    n = start.numel()
    if scal is None:
        scal = torch.zeros(2 * steps + 3, dtype=torch.float64, device=dev)
    a_slot = lambda i: i  # noqa: E731
    b_slot = lambda i: steps + i  # noqa: E731
    s1, s2, s0 = 2 * steps, 2 * steps + 1, 2 * steps + 2
    m.norm_into(start, scal[s0:s0 + 1])
    v = ops.lincomb(torch.empty(n, dtype=torch.float64, device=dev), [(start, 1.0, 2, s0)], scal)
    v_prev = torch.zeros_like(v)
    w = torch.empty_like(v)
    for i in range(steps):
        apply(v.view(1, -1), w.view(1, -1))
        m.inner_into(v, w, scal[a_slot(i):a_slot(i) + 1])
        terms = [(w, 1.0), (v, -1.0, 1, a_slot(i))]
        if i > 0:
            terms.append((v_prev, -1.0, 1, b_slot(i - 1)))
        # each update fused with the products that follow it (bitwise the separate lincomb and
        # inner/norm calls): <w, v> and <w, v_prev> into the adjacent slots s1, s2, then ||w||
        ops.lincomb_dots(w, terms, scal, m.q, [v, v_prev], scal[s1:s2 + 1])
        ops.lincomb_dots(w, [(w, 1.0), (v, -1.0, 1, s1), (v_prev, -1.0, 1, s2)], scal, m.q, [w],
                         scal[b_slot(i):b_slot(i) + 1], sqrt=True)
        v_new = ops.lincomb(torch.empty_like(v), [(w, 1.0, 2, b_slot(i))], scal)
        v_prev, v = v, v_new
    return scal


def lanczos_finish(values: np.ndarray, steps: int, safety: float) -> tuple[float, float]:
    """The host part of :func:`lanczos_bounds`: the tridiagonal's extreme Ritz values widened by the
    last beta and by ``safety``."""
    alphas = list(values[:steps])
    betas = list(values[steps:2 * steps])
    for index, value in enumerate(betas):
        if value < 1e-12:
            del alphas[index + 1:], betas[index + 1:]
            break
    size = len(alphas)
    tri = np.zeros((size, size))
    for i in range(size):
        tri[i, i] = alphas[i]
        if i + 1 < size:
            tri[i, i + 1] = tri[i + 1, i] = betas[i]
    ritz = np.linalg.eigvalsh(tri)  # a 12 x 12 tridiagonal on the host, as the reference does
    residual = betas[-1] if betas else 0.0
    lo = float(ritz[0]) - residual
    hi = float(ritz[-1]) + residual
    centre = 0.5 * (hi + lo)
    half = 0.5 * (hi - lo) * safety
    return centre - half, centre + half
