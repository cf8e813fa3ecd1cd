"""A small autotuner for the marching stencil's launch: it times candidate x-segment lengths (the
knob that trades blocks against the 4P-2 plane warm-up of each segment) on the real buffers and
keeps the fastest. Results are cached per problem signature for the life of the process.

Nothing here changes arithmetic: every segment length computes each output with the same
operations in the same order, so tuning never moves a result bit.
"""

from __future__ import annotations

import torch

from . import ops

#: Planes per segment tried (0 = the built-in heuristic).
CANDIDATES = (0, 4, 6, 8, 11, 16, 22, 33)

_cache: dict[tuple, dict] = {}


def tune(signature: tuple, run, candidates=CANDIDATES, repeats: int = 2) -> dict:
    """Time ``run()`` (which must build whatever plan it needs, so the segment setting in force at
    call time applies) under each candidate segment length; return ``{"best", "us"}`` and leave the
    library's segment setting at its automatic default. ``signature`` keys the cache."""
    if signature in _cache:
        return _cache[signature]
    timings = {}
    try:
        for c in candidates:
            ops.hdiv_set_segment(c)
            run()  # warm-up (and plan build)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(repeats):
                run()
            e1.record()
            torch.cuda.synchronize()
            timings[c] = e0.elapsed_time(e1) * 1e3 / repeats
    finally:
        ops.hdiv_set_segment(0)
    best = min(timings, key=timings.get)
    _cache[signature] = {"best": best, "us": timings}
    return _cache[signature]
