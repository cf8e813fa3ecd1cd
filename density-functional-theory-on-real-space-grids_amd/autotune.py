"""A small autotuner for the marching stencil's launch: it times candidate x-segment lengths (the
knob that trades blocks against the 4P-2 plane warm-up of each segment), then candidate tile
extents (y rows and z width of a block's tile, which trade halo reads against block count) at the
best segment length, on the real buffers, and keeps the fastest. Results are cached per problem
signature for the life of the process.

Nothing here changes arithmetic: every segment length and every tiling computes each output with
the same operations in the same order, so tuning never moves a result bit.
"""

from __future__ import annotations

import torch

from . import ops

#: Planes per segment tried (0 = the built-in heuristic).
CANDIDATES = (0, 4, 6, 8, 11, 16, 22, 33)

#: (largest tile rows along y, tile width along z) tried at the best segment length ((0, 0) = the
#: built-in choice: 16 rows at most, 13 or 16 columns by the fewer idle ones).
TILES = ((0, 0), (16, 16), (16, 13), (13, 13), (12, 16), (8, 16))

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
