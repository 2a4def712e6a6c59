"""Launch tuner for device-resident runs.

The tile is fixed at 32 x 32 own cells: every step kernel stages a tile and its 2-cell ring in LDS, keeps per-tile
row caches and per-tile books (wave speed, wet flag, window source), and the ring of a neighbour is that neighbour's
own tile. What a run can still choose is how a step is launched: the pipelined step (a step's cell update made by the
next step's kernel together with its faces, each tile recomputing its neighbours' ring) or the two-kernel step (face
kernel, then cell kernel); and how many threads work one tile, 256 (four rows of the tile per thread) or 512 (two),
in the face kernel and in the cell kernel. Which is faster depends on the members' sizes, on the mix of wet and dry
tiles and on how many tiles are live, so the tuner measures it on the run itself: each candidate in turn advances
the live run by a chunk of captured steps (after one untimed chunk that captures its graph), and its wall time is
divided by the steps and the mean number of active tiles. No step is thrown away; the fields do not depend on the
choice (the pipelined step is the two-kernel step bit for bit). The fastest candidate stays set.
"""

from __future__ import annotations

import time

import numpy as np

# (pipelined step, face-kernel threads, cell-kernel threads); the pipe kernel always runs 512 threads, so with the
# pipelined step only the cell kernel's count (its updates before a check or a window's end) is a choice
CANDIDATES = ((True, 512, 256), (True, 512, 512), (False, 512, 256), (False, 512, 512), (False, 256, 256),
              (False, 256, 512))  # fmt: skip


def apply_launch(b, launch) -> None:
    """Set a launch: (face threads, cell threads) or (pipelined, face threads, cell threads)."""
    if len(launch) == 3:
        b.run_pipeline(bool(launch[0]))
        b.set_launch(int(launch[1]), int(launch[2]))
    else:
        b.set_launch(*launch)


def tune_launch(run, graph_steps: int = 64, chunk: int | None = None, rounds: int = 1,
                candidates=CANDIDATES) -> tuple[tuple, list[dict]]:  # fmt: skip
    """Time every candidate (pipelined, faces_threads, cells_threads) on the live run; set and return the fastest,
    with the measurements (seconds per active tile-step)."""
    b = run.b
    warm = graph_steps if graph_steps > 0 else 16
    steps = chunk or 2 * warm
    log: list[dict] = []
    for _ in range(rounds):
        for c in candidates:
            if all(s.done for s in run.state):
                break
            apply_launch(b, c)
            run.steps(warm, graph_steps)  # captures this launch's graph
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
    apply_launch(b, best)
    return best, log
