"""Small autotuners for the marching stencil's launch.

:func:`tune` serves the batched launch: it times candidate x-segment lengths (the knob that trades
blocks against the 4P-2 plane warm-up of each segment), then candidate tile extents (y rows and z
width of a block's tile, which trade halo reads against block count) at the best segment length,
on the real buffers, and keeps the fastest.

:func:`tune_stencil` serves single-grid solves: it times the vector kernel's Chebyshev step (or the
apply, for single vectors) under candidate (tile rows, tile width, segment length) settings and puts
the fastest in force for that problem shape (``ops.hdiv_vec_set``). A launch's time depends on how
its blocks fall on the 256 compute units, which no simple rule captures to within the 10-20%
between neighbouring settings.

Results are cached per problem signature for the life of the process. Nothing here changes
arithmetic: every segment length and every tiling computes each output with the same operations in
the same order, so tuning never moves a result bit.
"""

from __future__ import annotations

import math

import torch

from . import ops

#: Planes per segment tried (0 = the built-in heuristic).
CANDIDATES = (0, 4, 6, 8, 11, 16, 22, 33)

#: (largest tile rows along y, tile width along z) tried at the best segment length: (0, 0) is the
#: built-in choice (for order 8 the vector kernel at the width with the fewest idle columns over the
#: batch, at most 17 rows); widths 16-28 are the vector kernel's, 13 the column kernel's.
TILES = ((0, 0), (16, 16), (12, 16), (0, 20), (0, 22), (11, 22), (0, 24), (0, 28), (9, 28), (16, 13), (13, 13))

_cache: dict[tuple, dict] = {}


def _time(run, repeats: int) -> float:
    run()  # warm-up (and plan build)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(repeats):
        run()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) * 1e3 / repeats


def tune(signature: tuple, run, candidates=CANDIDATES, tiles=TILES, repeats: int = 2) -> dict:
    """Time ``run()`` (which must build whatever plan it needs, so the launch settings in force at
    call time apply) under each candidate segment length with the automatic tiling, then under each
    candidate tiling at the fastest segment length; return ``{"best", "tile", "us", "tile_us"}``
    and leave the library's settings at their automatic defaults. ``signature`` keys the cache."""
    if signature in _cache:
        return _cache[signature]
    timings, tile_timings = {}, {}
    try:
        for c in candidates:
            ops.hdiv_set_segment(c)
            timings[c] = _time(run, repeats)
        best = min(timings, key=timings.get)
        ops.hdiv_set_segment(best)
        for t in tiles:
            ops.hdiv_set_tile(*t)
            tile_timings[t] = _time(run, repeats)
    finally:
        ops.hdiv_set_segment(0)
        ops.hdiv_set_tile(0, 0)
    tile = min(tile_timings, key=tile_timings.get) if tile_timings else (0, 0)
    _cache[signature] = {"best": best, "tile": tile, "us": timings,
                         "tile_us": {f"{ty}x{tz}": v for (ty, tz), v in tile_timings.items()}}
    return _cache[signature]


def apply(tuning: dict | None) -> None:
    """Put a tuning's launch settings in force (``None``: the automatic ones)."""
    if tuning is None:
        ops.hdiv_set_segment(0)
        ops.hdiv_set_tile(0, 0)
    else:
        ops.hdiv_set_segment(tuning["best"])
        ops.hdiv_set_tile(*tuning.get("tile", (0, 0)))


# ---- single-grid launches of the vector kernel ----

#: Tile widths of the vector kernel by element type (the widths it is compiled for).
VEC_WIDTHS = {torch.float64: (16, 20, 22, 24, 28), torch.float32: (16, 20, 24, 28)}
#: Largest tile rows tried (each is rounded to the even split of n1) and segment lengths tried.
VEC_ROWS = (17, 11, 9)
VEC_SEGMENTS = (4, 5, 6, 7, 9, 11, 13)
#: Device busy time (clock cycles of a spin kernel) ahead of each group of timed candidates, so the
#: host enqueues a group's launches while the device is still busy and the events time the kernels,
#: not the host's launch rate.
LEAD_CYCLES = 2_000_000
GROUP = 6

_stencil_cache: dict[tuple, dict] = {}
_stencil_seen: dict[tuple, int] = {}
#: Set False to keep the built-in launch rule (the solvers then tune nothing).
ENABLED = True
#: A solver tunes a problem shape on this solve of it in the process. The sweep takes 10-20 ms, which
#: a shape solved a few times (a scan over geometries) would not win back; a self-consistent solve
#: diagonalises its grid 13-30 times.
TUNE_ON_SOLVE = 5


def stencil_candidates(shape, dtype) -> list:
    """(tile rows, tile width, segment) settings tried for ``shape``: the two tile widths with the
    fewest idle columns, three tile heights, seven segment lengths; ``None`` is the built-in rule."""
    n0, n1, n2 = shape
    widths = VEC_WIDTHS[dtype]
    idle = {w: math.ceil(n2 / w) * w - n2 for w in widths}
    best_widths = sorted(widths, key=lambda w: (idle[w], -w))[:2]
    rows = sorted({math.ceil(n1 / math.ceil(n1 / r)) for r in VEC_ROWS})
    segments = [s for s in VEC_SEGMENTS if s <= n0]
    return [None] + [(r, w, s) for w in best_widths for r in rows for s in segments]


def _time_groups(candidates, put, run, repeats: int) -> dict:
    times = {}
    sleep = getattr(torch.cuda, "_sleep", None)
    for start in range(0, len(candidates), GROUP):
        group = candidates[start:start + GROUP]
        events = [torch.cuda.Event(enable_timing=True) for _ in range(len(group) + 1)]
        if sleep is not None:
            sleep(LEAD_CYCLES)
        put(group[0])
        run()  # first launch of the group (code and descriptors warm)
        events[0].record()
        for i, c in enumerate(group):
            put(c)
            for _ in range(repeats):
                run()
            events[i + 1].record()
        events[-1].synchronize()
        for i, c in enumerate(group):
            times[c] = events[i].elapsed_time(events[i + 1]) * 1e3 / repeats
    return times


def tune_stencil(geom, k: int) -> dict | None:
    """Time the order-8 vector kernel on ``geom`` for blocks of ``k`` vectors (the Chebyshev step, or
    the apply when ``k == 1``) under :func:`stencil_candidates`, put the fastest setting in force for
    this shape and return ``{"best", "us"}`` (``best`` None: the built-in rule). Runs on the current
    stream on scratch blocks; call it outside graph capture."""
    if not ENABLED or geom.order != 8 or geom.dtype not in VEC_WIDTHS:
        return None
    signature = (tuple(geom.shape), k, geom.dtype, geom.device)
    if signature in _stencil_cache:
        return _stencil_cache[signature]
    shape, dtype = tuple(geom.shape), geom.dtype
    n = geom.n_points
    y = torch.zeros((k, n), dtype=dtype, device=geom.device)
    prev = torch.zeros_like(y)
    out = torch.empty_like(y)
    table = torch.tensor([[0.0, 1.0, 0.0], [0.5, 1.0, 0.5]], dtype=torch.float64, device=geom.device)
    if k == 1:
        def run():
            ops.hdiv_apply(y, geom, out)
    else:
        def run():
            ops.cheb_step(y, prev, geom, table, 1, out)

    def put(c):
        if c is None:
            ops.hdiv_vec_set(shape, k, dtype, 0, 0, 0)
        else:
            ops.hdiv_vec_set(shape, k, dtype, *c)

    candidates = stencil_candidates(shape, dtype)
    try:
        first = _time_groups(candidates, put, run, repeats=4)
        finalists = sorted(first, key=first.get)[:4]
        rounds = [_time_groups(finalists, put, run, repeats=12) for _ in range(2)]
        final = {c: sorted(r[c] for r in rounds)[0] for c in finalists}
        best = min(final, key=final.get)
    except Exception:
        put(None)
        raise
    put(best)
    key = lambda c: "auto" if c is None else "x".join(map(str, c))  # noqa: E731
    result = {"best": best, "us": {key(c): round(t, 3) for c, t in final.items()}}
    _stencil_cache[signature] = result
    return result


def tune_for_solve(geom, k: int) -> bool:
    """Count a solve of ``(geom.shape, k, geom.dtype)`` and tune it (:func:`tune_stencil`) on the
    :data:`TUNE_ON_SOLVE`-th. True when this call changed the launch settings in force: graphs
    captured before it still hold the old launches and should be recaptured."""
    if not hasattr(geom, "packed"):
        return False  # the model systems' plain operator has one launch shape
    signature = (tuple(geom.shape), k, geom.dtype, geom.device)
    if not ENABLED or signature in _stencil_cache:
        return False
    seen = _stencil_seen.get(signature, 0) + 1
    _stencil_seen[signature] = seen
    if seen < TUNE_ON_SOLVE:
        return False
    tuning = tune_stencil(geom, k)
    return tuning is not None and tuning["best"] is not None
