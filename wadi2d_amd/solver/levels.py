"""Time levels of the tiles for local time stepping (behaviour of wadi2d/solver/levels.py:1-196 on a uniform grid).

A tile at level L steps 2^L times the smallest step. Levels are decided on the host once per cycle from each tile's
fastest wave; neighbours differ by at most one level, so a coarse step spans exactly two of its finer neighbour's.
The device sees only lists: per level, the tiles that step, and the edge rows between a tile and a coarser neighbour.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

W_COL, N_COL, E_COL, S_COL = 3, 1, 4, 6
"""Columns of ``TileGrid.nbr`` holding the west, north, east and south neighbours; side k of an edge row is the k-th."""
SIDE_COLUMNS = (W_COL, N_COL, E_COL, S_COL)
LOG_MARGIN = 1e-9  # a tile exactly on a power of two is not rounded down by the logarithm
HOLD = 0.85  # a tile keeps its level while the smallest step stays above this share of what the fastest allows
CLIMB = 0.25  # and climbs only this far, in levels, past the mark


@dataclass(frozen=True)
class LevelRule:
    """What a tile's own step depends on (levels.py:26-37)."""

    cfl: float
    wave_floor: float
    buffer: float
    dt_max: float
    top: int
    hold: float = HOLD
    climb: float = CLIMB


def first_level_due(n: int, top: int) -> int:
    """The coarsest level whose step begins at smallest step n of a cycle: that of n's lowest set bit, capped at
    ``top``; every level begins at 0 (level_stepper.py:27-31)."""
    return top if n == 0 else min(top, (n & -n).bit_length() - 1)


def graded(level: np.ndarray, active: np.ndarray, nbr: np.ndarray) -> np.ndarray:
    """The largest levels at or under ``level`` with active neighbours at most one apart (levels.py:111-135).

    Each tile ends at the least, over the active tiles s joined to it through active tiles, of level[s] plus the
    number of tile steps from s: a breadth-first sweep outward in order of level, every tile settled once.
    """
    n = nbr.shape[0]
    on = np.zeros(n, bool)
    on[active] = True
    out = np.zeros(n, np.int64)
    if active.size == 0:
        return out.astype(level.dtype)
    big = np.iinfo(np.int64).max // 4
    best = np.full(n, big, np.int64)
    best[active] = level[active]
    settled = np.zeros(n, bool)
    v = int(best[active].min())
    while True:
        frontier = np.flatnonzero(on & ~settled & (best == v))
        if frontier.size == 0:
            rest = on & ~settled
            if not rest.any():
                break
            v = int(best[rest].min())
            continue
        settled[frontier] = True
        reach = nbr[frontier].ravel()
        reach = reach[on[reach] & ~settled[reach]]
        np.minimum.at(best, reach, v + 1)
    out[active] = best[active]
    return out.astype(level.dtype)


def assign_levels(wave: np.ndarray, active: np.ndarray, nbr: np.ndarray, dx: float, rule: LevelRule,
                  pinned: np.ndarray, previous: np.ndarray | None = None,
                  may_climb: bool = True) -> tuple[np.ndarray, float]:  # fmt: skip
    """(level per tile, smallest step) from the fastest wave seen in each tile (levels.py:43-108).

    A tile answers to the larger of its own wave and ``buffer`` times its fastest neighbour's (a front may enter it
    within the cycle); its step is cfl dx over that, at most dt_max. Its level is how many doublings of the fastest
    tile's step fit in its own, at most ``top``. A level already held stays while the step need not fall below
    ``hold`` of what the fastest tile allows, and (when ``may_climb``) until the tile is ``climb`` of a level past the
    next; pinned tiles stay at 0; then the grading. The smallest step is the least of each tile's step over its
    span.
    """
    n = nbr.shape[0]
    if active.size == 0:
        return np.zeros(n, np.int32), rule.dt_max
    speed = np.zeros(n)
    speed[active] = np.maximum(wave[active], rule.wave_floor)
    felt = np.maximum(speed[active], rule.buffer * speed[nbr[active]].max(axis=1))
    own_dt = np.minimum(rule.cfl * dx / felt, rule.dt_max)
    room = np.log2(own_dt / own_dt.min()) + LOG_MARGIN
    lv = np.clip(np.floor(room), 0, rule.top).astype(np.int32)
    if previous is not None:
        held = previous[active]
        stay = held <= room - np.log2(rule.hold)
        if may_climb:
            stay &= room < held + 1 + rule.climb
        lv = np.where(stay, held, lv)
    level = np.zeros(n, np.int32)
    level[active] = lv
    level[pinned] = 0
    level = graded(level, active, nbr).astype(np.int32)
    return level, float((own_dt / (1 << level[active])).min())


@dataclass
class CyclePlan:
    """One cycle's lists in host form (levels.py:138-196 ``Plan``), indexed by time level n.

    groups[n]: the tiles of level n; upto[n]: of every level <= n (what steps when level n's step begins);
    halo[n]: upto[n] and the coarse owners of the strips of levels <= n; coarse[n]: the tiles of level n with a finer
    neighbour (their old depth is kept); edges_x[n], edges_y[n]: the edge rows (fine tile, side, coarse tile, edge
    number) whose fine tile is at level n, by axis; strips[n]: those rows whose faces the coarse tile owns (side east
    or south).
    """

    level: np.ndarray
    top: int
    groups: list[np.ndarray]
    upto: list[np.ndarray]
    halo: list[np.ndarray]
    coarse: list[np.ndarray]
    edges_x: list[np.ndarray]
    edges_y: list[np.ndarray]
    strips: list[np.ndarray]


def plan_cycle(level: np.ndarray, active: np.ndarray, nbr: np.ndarray) -> CyclePlan:
    """The lists of a cycle at these levels (levels.py:147-196 for a grid of one cell size)."""
    active = np.asarray(active, np.int32)
    lv = level[active]
    top = int(lv.max()) if active.size else 0
    levels = range(top + 1)
    on = np.zeros(level.size, bool)
    on[active] = True
    # edge rows, side by side: west rows of every tile first, then north, east, south; numbered in that order
    blocks = []
    for side, col in enumerate(SIDE_COLUMNS):
        other = nbr[active, col]
        coarser = on[other] & (level[other] > lv)
        blocks.append(np.stack([active[coarser], np.full(int(coarser.sum()), side, np.int32), other[coarser]], 1))
    rows = np.concatenate(blocks).astype(np.int32) if blocks else np.zeros((0, 3), np.int32)
    rows = np.concatenate([rows, np.arange(rows.shape[0], dtype=np.int32)[:, None]], 1)
    fine_level = level[rows[:, 0]]
    side = rows[:, 1]

    def by_level(mask: np.ndarray) -> list[np.ndarray]:
        return [rows[mask & (fine_level == n)] for n in levels]

    strips = by_level(side >= 2)
    has_finer = np.zeros(level.size, bool)
    has_finer[rows[:, 2]] = True
    halo, owners = [], np.zeros(0, np.int32)
    upto = [active[lv <= n] for n in levels]
    for n in levels:
        owners = np.concatenate([owners, strips[n][:, 2]])
        halo.append(np.union1d(upto[n], owners).astype(np.int32))
    return CyclePlan(level=level, top=top, groups=[active[lv == n] for n in levels], upto=upto, halo=halo,
                     coarse=[active[(lv == n) & has_finer[active]] for n in levels],
                     edges_x=by_level(side % 2 == 0), edges_y=by_level(side % 2 == 1), strips=strips)  # fmt: skip
