"""Autotuner for the LDS-tiled 3D push: tile shape, halo, workgroup variant
and (under the fixed sort policy) the sort interval.

Every candidate computes bit-identical physics (the tiled and untiled pushes
perform the same operations and the deposit is integer), so the choice is
purely a performance one.  Candidates are timed on the live context: the
simulation advances during tuning (the tuning steps count as warm-up).
"""
from __future__ import annotations

import time

# (tx, ty, tz, halo, block, prefetch) combinations built into libaurora_hip.so
SHAPES = [(8, 8, 8, 2, 512, 0), (8, 8, 8, 1, 512, 0), (8, 8, 4, 2, 512, 0),
          (16, 8, 4, 2, 512, 0), (4, 4, 4, 2, 512, 0), (8, 8, 4, 2, 512, 1),
          (8, 8, 4, 2, 256, 1), (8, 8, 4, 2, 256, 0), (8, 8, 8, 2, 512, 1),
          (8, 8, 8, 2, 256, 1), (8, 8, 4, 2, 512, 3), (8, 8, 4, 1, 512, 3),
          (8, 8, 4, 2, 512, 5), (8, 8, 8, 2, 512, 5), (8, 8, 8, 1, 512, 1),
          (8, 8, 8, 1, 512, 5), (8, 8, 8, 1, 512, 3), (8, 8, 8, 1, 512, 7),
          (8, 8, 8, 2, 512, 13), (8, 8, 8, 1, 512, 13), (8, 8, 8, 1, 512, 15),
          (8, 8, 4, 2, 512, 13), (8, 8, 4, 2, 512, 11), (8, 8, 4, 2, 512, 7),
          (8, 8, 4, 2, 512, 15)]


def tune(ctx, dt, shapes=SHAPES, sort_intervals=(20,), steps=20, log=None, adaptive=False,
         refine=3, refine_steps=None):
    """Time each candidate over `steps` steps (plus one warm step after the
    re-sort) and leave the fastest configured.  Under the adaptive sort
    policy (adaptive=True) the interval is not tuned: the window should span
    at least one sort cycle.  The `refine` fastest are then timed again over
    `refine_steps` (default 2.5 x steps): candidates within a few per cent
    of each other are told apart by the longer window, not by the noise of
    the short one.  Returns (best, table) with table rows (ms_per_step,
    shape, interval or None); a refined row carries the longer window's
    time."""
    from . import hip
    table = []
    intervals = (None,) if adaptive else sort_intervals
    if adaptive:
        ctx.sort_policy(hip.SORT_ADAPTIVE)

    def timed(shape, interval, n):
        ctx.tiling(*shape[:4])
        ctx.push_variant(*(shape[4:] or (512, 0)))
        ctx.sort()
        if interval is not None:
            ctx.sort_interval(interval)
        ctx.step(dt, 1)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.step(dt, n)
        ctx.sync()
        return 1e3 * (time.perf_counter() - t0) / n

    for shape in shapes:
        try:
            ctx.tiling(*shape[:4])
            ctx.push_variant(*(shape[4:] or (512, 0)))
        except Exception:  # shape not applicable to this mesh
            continue
        for interval in intervals:
            ms = timed(shape, interval, steps)
            table.append((ms, shape, interval))
            if log:
                what = "adaptive sort" if interval is None else f"sort every {interval}"
                log(f"autotune {shape} {what}: {ms:.3f} ms/step")
    if refine and len(table) > 1:
        n = refine_steps or int(2.5 * steps)
        order = sorted(range(len(table)), key=lambda i: table[i][0])[:refine]
        for i in order:
            _, shape, interval = table[i]
            ms = timed(shape, interval, n)
            table[i] = (ms, shape, interval)
            if log:
                log(f"autotune refined {shape} over {n} steps: {ms:.3f} ms/step")
        table_best = [table[i] for i in order]
        best = min(table_best, key=lambda r: r[0])
    else:
        best = min(table, key=lambda r: r[0])
    ctx.tiling(*best[1][:4])
    ctx.push_variant(*(best[1][4:] or (512, 0)))
    ctx.sort()
    if best[2] is not None:
        ctx.sort_interval(best[2])
    return best, table
