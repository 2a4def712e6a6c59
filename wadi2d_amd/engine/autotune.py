"""Launch-shape tuner for device-resident runs.

The tile is fixed at 32 x 32 own cells (the layout every caller binds to); what a launch can still choose is how
many threads work one tile in each kernel: 256 (four rows of the tile per thread) or 512 (two), for the face and
the cell kernel independently. Which is faster depends on the mix of wet and dry tiles and on how many tiles are
live, so the tuner measures it on the run itself: each candidate in turn advances the live run by a chunk of
captured steps (after one untimed chunk that captures its graph), and its wall time is divided by the steps and
the mean number of active tiles. No step is thrown away; the fields do not depend on the shape. The fastest
candidate stays set.
"""

from __future__ import annotations

import time

import numpy as np

CANDIDATES = ((512, 256), (512, 512), (256, 256), (256, 512))


def tune_launch(run, graph_steps: int = 64, chunk: int | None = None, rounds: int = 1,
                candidates=CANDIDATES) -> tuple[tuple[int, int], list[dict]]:  # fmt: skip
    """Time every candidate (faces_threads, cells_threads) on the live run; set and return the fastest, with the
    measurements (seconds per active tile-step)."""
    b = run.b
    warm = graph_steps if graph_steps > 0 else 16
    steps = chunk or 2 * warm
    log: list[dict] = []
    for _ in range(rounds):
        for c in candidates:
            if all(s.done for s in run.state):
                break
            b.set_launch(*c)
            run.steps(warm, graph_steps)  # captures this shape's graph
            before = b.nact
            t0 = time.perf_counter()
            run.steps(steps, graph_steps)
            wall = time.perf_counter() - t0
            tiles = max(0.5 * (before + b.nact), 1.0)
            log.append({"launch": c, "steps": steps, "tiles": tiles, "s_per_tile_step": wall / (steps * tiles)})
    if not log:
        return tuple(candidates[0]), log
    score = {c: float(np.median([r["s_per_tile_step"] for r in log if r["launch"] == c])) for c in
             {r["launch"] for r in log}}  # fmt: skip
    best = min(score, key=score.get)
    b.set_launch(*best)
    return best, log
