"""Local time stepping, restated for the oracle. TEST INFRASTRUCTURE ONLY (the checker of the HIP level stepper).

What it restates:
  levels of the tiles .......... assign_levels, graded, plan_cycle, trailing_level
                                 (wadi2d/solver/levels.py:43-196, solver/level_stepper.py:27-31)
  the level stepper ............ OracleLevelStepper (wadi2d/solver/level_stepper.py:48-222), uniform grids only: no
                                 quadtree interfaces, no aligned strips' hooks or wake rules
  the operations ............... OracleLevels over oracle_lts_* of step_oracle.cpp (wadi2d/solver/cpu/levels.py)

The plan is written here as plain loops over tiles, apart from the package's vectorised one, so that a test can hold
one against the other.
"""

from __future__ import annotations

import math

import numpy as np

from oracle.oracle import STRIDE2, NumericalFailure, OracleModel, OracleStepper, lib

SIDES = (3, 1, 4, 6)  # west, north, east, south in TileGrid.nbr (levels.py:15)
EAST = 2
LOG_MARGIN = 1e-9
HOLD = 0.85
CLIMB = 0.25
CLIMB_EVERY = 16  # level_stepper.py:23


def trailing_level(n: int, top: int) -> int:
    """The coarsest level that starts a step at smallest step n of a cycle (level_stepper.py:27-31)."""
    if n == 0:
        return top
    low = 0
    while not (n >> low) & 1:
        low += 1
    return min(top, low)


def graded(level: np.ndarray, active: np.ndarray, nbr: np.ndarray) -> np.ndarray:
    """Levels lowered until active neighbours differ by at most one (levels.py:111-135), by sweeps to a fixed point."""
    big = np.iinfo(np.int32).max // 2
    cur = {int(t): int(level[t]) for t in active}
    while True:
        nxt = {}
        for t in cur:
            low = min(cur.get(int(o), big) for o in nbr[t])
            nxt[t] = min(cur[t], low + 1)
        if nxt == cur:
            break
        cur = nxt
    out = np.zeros_like(level)
    for t, v in cur.items():
        out[t] = v
    return out


def assign_levels(wave, active, nbr, dx, cfl, floor, buffer, dt_max, top, pinned, previous, may_climb):
    """(level per tile, smallest step) from each tile's fastest wave (levels.py:43-108)."""
    n = nbr.shape[0]
    speed = np.zeros(n)
    for t in active:
        speed[t] = max(float(wave[t]), floor)
    want = {}
    for t in active:
        around = max(speed[o] for o in nbr[t])
        felt = max(speed[t], buffer * around)
        want[int(t)] = min(cfl * dx / felt, dt_max)
    fastest = min(want.values()) if want else dt_max
    level = np.zeros(n, np.int32)
    for t in active:
        room = float(np.log2(want[int(t)] / fastest)) + LOG_MARGIN
        best = int(min(max(math.floor(room), 0), top))
        if previous is not None:
            old = int(previous[t])
            keep = old <= room - float(np.log2(HOLD))
            if may_climb:
                keep = keep and room < old + 1 + CLIMB
            if keep:
                best = old
        level[t] = best
    level[pinned] = 0
    level = graded(level, active, nbr)
    dt0 = min(want[int(t)] / (1 << int(level[t])) for t in active) if want else dt_max
    return level, dt0


class Plan:
    """One cycle's lists (levels.py:138-196) for a uniform grid: per time level n the tiles of that level (groups),
    of every level up to n (upto), the tiles whose halos a step of level n refreshes (halo), the tiles of level n with
    a finer neighbour (coarse), and the edge rows (fine tile, side, coarse tile, edge number) by the fine tile's level,
    split by axis (edges_x, edges_y) and those whose faces the coarse tile owns (strips)."""

    def __init__(self, level: np.ndarray, active: np.ndarray, nbr: np.ndarray) -> None:
        on = np.zeros(level.size, bool)
        on[active] = True
        self.top = int(max((level[t] for t in active), default=0))
        span = range(self.top + 1)
        self.groups = [np.array([t for t in active if level[t] == n], np.int32) for n in span]
        self.upto = [np.array([t for t in active if level[t] <= n], np.int32) for n in span]
        rows = []
        for side in range(4):  # the reference stacks the rows side by side (levels.py:160-164)
            for t in active:
                o = int(nbr[t, SIDES[side]])
                if on[o] and level[o] > level[t]:
                    rows.append((int(t), side, o, len(rows)))
        rows = np.array(rows, np.int32).reshape(-1, 4)
        fine = level[rows[:, 0]] if rows.size else np.zeros(0, np.int32)

        def pick(mask):
            return [rows[mask & (fine == n)] for n in span]

        self.edges_x = pick(rows[:, 1] % 2 == 0)
        self.edges_y = pick(rows[:, 1] % 2 == 1)
        self.strips = pick(rows[:, 1] >= EAST)
        finer = set(int(c) for c in rows[:, 2])
        self.coarse = [np.array([t for t in active if level[t] == n and int(t) in finer], np.int32) for n in span]
        self.halo = []
        owners: set[int] = set()
        for n in span:
            owners |= set(int(c) for c in self.strips[n][:, 2])
            self.halo.append(np.array(sorted(set(int(t) for t in self.upto[n]) | owners), np.int32))


def _ptr(a: np.ndarray):
    return a.ctypes.data


class OracleLevels:
    """The level stepper's operations (cpu/levels.py CpuLevels) on the oracle's C++ model."""

    def __init__(self, backend) -> None:
        self.b = backend
        self.dt32 = backend.prec == 4
        lib().oracle_lts_init(backend.h)

    def _dt(self, dt: float) -> float:
        return float(np.float32(dt)) if self.dt32 else float(dt)

    def _tiles(self, op: int, tiles: np.ndarray, dt: float = 0.0, flag: int = 0, arg: int = 0) -> None:
        t = np.ascontiguousarray(tiles, np.int32)
        lib().oracle_lts_tiles(self.b.h, op, _ptr(t), int(t.size), self._dt(dt), int(flag), int(arg))

    def _edges(self, op: int, rows: np.ndarray, dt: float = 0.0, flag: int = 0) -> None:
        e = np.ascontiguousarray(rows, np.int32)
        if e.size:
            lib().oracle_lts_edges(self.b.h, op, _ptr(e), int(e.shape[0]), self._dt(dt), int(flag))

    def set_level(self, level: np.ndarray) -> None:
        lv = np.ascontiguousarray(level, np.int32)
        lib().oracle_lts_set_level(self.b.h, _ptr(lv))

    def halo_state(self, tiles, n: int, want_plan: bool) -> None:
        self._tiles(0, tiles, flag=want_plan, arg=n)

    def faces(self, tiles, dt, want_wave) -> None:
        self._tiles(1, tiles, dt, want_wave)

    def faces_done(self, tiles) -> None:
        self._tiles(2, tiles)

    def limit(self, tiles, dt) -> None:
        self._tiles(3, tiles, dt)

    def limit_done(self, tiles) -> None:
        self._tiles(4, tiles)

    def save(self, tiles) -> None:
        self._tiles(5, tiles)

    def update(self, tiles, dt, want_plan) -> None:
        self._tiles(6, tiles, dt, want_plan)

    def commit(self, tiles) -> None:
        self._tiles(7, tiles)

    def clear_wave(self) -> None:
        self._tiles(8, np.zeros(0, np.int32))

    def clear_wet(self) -> None:
        self._tiles(9, np.zeros(0, np.int32))

    def boundary(self) -> None:
        self._tiles(10, np.zeros(0, np.int32))

    def strips(self, rows, dt, want_wave) -> None:
        self._edges(0, rows, dt, want_wave)

    def commit_strips(self, rows) -> None:
        self._edges(1, rows)

    def estimate(self, rows) -> None:
        self._edges(2, rows)

    def flux(self, rows, dt) -> None:
        self._edges(3, rows, dt)


class OracleLevelStepper(OracleStepper):
    """The level stepper (level_stepper.py:48-222) on a uniform grid. A cycle is a run of smallest steps after which
    every tile stands at one time; the wave speeds of its last span set the next cycle's levels and smallest step."""

    def __init__(self, backend, grid, num) -> None:
        super().__init__(backend, grid, num)
        self.ops = OracleLevels(backend)
        self.top = num.lts_levels if num.local_time_stepping else 0
        self.pinned = np.zeros(grid.ntiles, bool)
        self.level = np.zeros(grid.ntiles, np.int32)
        self.plan: Plan | None = None
        self.since_active = 0
        self.fresh_wave = False
        self.cycles = 0
        self.tile_steps = 0
        self.substep_hook = None

    def pin(self, tiles) -> None:
        self.pinned[tiles] = True

    def rebuild_active(self, retain: bool = False, check: bool = False) -> None:
        before = np.array(self.backend.active_tiles, copy=True)
        ntiles = self.grid.ntiles
        b = self.backend
        if not self.adaptive_active:
            if int(b.nact) != ntiles - 1:
                b.set_active(np.arange(1, ntiles, dtype=np.int32))
        else:  # the base rule (stepper.py:56-92) with the level stepper's wake: the wet share only
            wet = b.wet_flags().astype(bool) | self.keep_active
            wet[0] = False
            act = wet.copy()
            act[self.grid.nbr[wet].ravel()] = True
            current = np.zeros(ntiles, bool)
            current[b.active_tiles] = True
            if retain:
                act |= current
            elif check:
                self.idle[act] = 0
                self.idle[~act] += 1
                act |= current & (self.idle < 2)
            else:
                self.idle[:] = 0
            act[0] = False
            woken = act & ~current
            b.set_active(np.nonzero(act)[0].astype(np.int32))
            if (retain or check) and woken.any():
                b.refresh_plan(np.nonzero(woken)[0].astype(np.int32))
        if not np.array_equal(before, self.backend.active_tiles):
            self.plan = None

    def _set_plan(self) -> None:
        active = np.asarray(self.backend.active_tiles)
        self.level = graded(self.level, active, self.grid.nbr)
        self.plan = Plan(self.level, active, self.grid.nbr)
        self.ops.set_level(self.level)

    def _reassign(self) -> None:
        wave = self.backend.wave()
        if not np.all(np.isfinite(wave)):
            raise NumericalFailure(f"non-finite wave speed at t = {self.t:.3f} s")
        active = np.asarray(self.backend.active_tiles)
        keep = self.level.copy()
        pinned = self.pinned & np.isin(np.arange(self.pinned.size), active)
        self.cycles += 1
        n = self.num
        self.level, self.dt = assign_levels(wave, active, self.grid.nbr, self.grid.dx, n.cfl, n.lts_wave_floor,
                                            n.lts_buffer, n.dt_max, self.top, pinned, keep,
                                            self.cycles % CLIMB_EVERY == 0)  # fmt: skip
        if self.dt < n.dt_min:
            raise NumericalFailure(f"time step {self.dt:.2e} s below the minimum at t = {self.t:.3f} s")
        if not np.array_equal(keep, self.level):
            self.plan = None

    def _substep(self, n: int, dt0: float, wave_step: bool, wet_due: bool) -> None:  # level_stepper.py:137-194
        ops, p = self.ops, self.plan
        k = trailing_level(n, p.top)
        tiles = p.upto[k]
        if tiles.size == 0:
            return
        dts = [dt0 * (1 << a) for a in range(k + 2)]
        due = range(k + 1)
        wet_step = wave_step and wet_due
        ops.halo_state(p.halo[k], n, wave_step)
        if wave_step:
            ops.clear_wave()
        for a in due:
            if p.groups[a].size:
                ops.faces(p.groups[a], dts[a], wave_step)
        for a in due:
            ops.strips(p.strips[a], dts[a], wave_step)
        ops.faces_done(tiles)
        for a in due:
            if p.groups[a].size:
                ops.limit(p.groups[a], dts[a])
        ops.limit_done(tiles)
        for a in due:
            if p.coarse[a].size:
                ops.save(p.coarse[a])
            for rows in (p.edges_x[a], p.edges_y[a]):
                if rows.size:
                    if a < k:
                        ops.estimate(rows)
                    else:
                        ops.flux(rows, dts[a])
        if wet_step:
            ops.clear_wet()
        for a in due:
            if p.groups[a].size:
                ops.update(p.groups[a], dts[a], wave_step)
        ops.boundary()
        ops.commit(tiles)
        for a in due:
            ops.commit_strips(p.strips[a])
        if self.substep_hook is not None:
            self.substep_hook(dt0)
        self.nsteps += 1
        self.tile_steps += sum(int(p.groups[a].size) for a in due)

    def advance(self, t_end: float, on_step=None) -> float:  # level_stepper.py:196-222
        num = self.num
        while self.t < t_end - 1e-9:
            if self.plan is None:
                self._set_plan()
            span = 1 << self.plan.top
            cycle = span * ((max(num.dt_every, span) + span - 1) // span) if self.fresh_wave else 1
            dt0 = min(self.dt, num.dt_max, (t_end - self.t) / cycle)
            self.since_active += cycle
            wet_due = self.since_active >= num.active_every
            for n in range(cycle):
                self._substep(n % span, dt0, n == cycle - span, wet_due)
            self.t = self.t + cycle * dt0
            self.fresh_wave = True
            self._reassign()
            if wet_due:
                self.since_active = 0
                if self.adaptive_active:
                    self.rebuild_active(check=True)
            if on_step is not None:
                on_step(self)
        return self.t


class OracleLevelModel(OracleModel):
    """OracleModel with the level stepper; boundary tiles are pinned to the smallest step (model.py:145-150)."""

    def __init__(self, tables, num=None, prec: int = 8, threads: int = 1) -> None:
        super().__init__(tables, num, prec, threads)
        self.stepper = OracleLevelStepper(self.backend, self.grid, self.num)

    def set_boundary(self, bc) -> None:  # model.py:145-150: pinned before the list is rebuilt
        self.backend.set_boundary(bc)
        self.boundary_tiles = np.unique(np.asarray(bc.cell) // STRIDE2)
        self.stepper.keep_active[self.boundary_tiles] = True
        self.stepper.pin(self.boundary_tiles)
        self.stepper.rebuild_active(retain=True)
