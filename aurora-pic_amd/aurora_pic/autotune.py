"""Autotuner for the LDS-tiled 3D push: tile shape, halo and sort interval.

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
          (8, 8, 8, 2, 256, 1)]


def tune(ctx, dt, shapes=SHAPES, sort_intervals=(20,), steps=20, log=None):
    """Time each (shape, interval) over `steps` steps (plus one warm step
    after the re-sort) and leave the fastest configured.  Returns
    (best, table) with table rows (ms_per_step, shape, interval)."""
    table = []
    for shape in shapes:
        try:
            ctx.tiling(*shape[:4])
            ctx.push_variant(*(shape[4:] or (512, 0)))
        except Exception:  # shape not applicable to this mesh
            continue
        for interval in sort_intervals:
            ctx.sort()
            ctx.sort_interval(interval)
            ctx.step(dt, 1)
            ctx.sync()
            t0 = time.perf_counter()
            ctx.step(dt, steps)
            ctx.sync()
            ms = 1e3 * (time.perf_counter() - t0) / steps
            table.append((ms, shape, interval))
            if log:
                log(f"autotune {shape} sort every {interval}: {ms:.3f} ms/step")
    best = min(table)
    ctx.tiling(*best[1][:4])
    ctx.push_variant(*(best[1][4:] or (512, 0)))
    ctx.sort()
    ctx.sort_interval(best[2])
    return best, table
