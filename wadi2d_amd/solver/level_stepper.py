"""The time loop with local time stepping (behaviour of wadi2d/solver/level_stepper.py:48-222, uniform grids).

A cycle is a run of smallest steps after which every tile stands at the same time. At smallest step n the tiles of
every level up to ``first_level_due(n)`` begin a step of their own length; a fine tile beside a coarse one reads the
coarse depth interpolated in time, the edge faces step at the fine rate whoever owns them, and what the fine steps
carry across an edge beyond the coarse step's estimate is owed to the coarse cell's residual, so no water is made or
lost. The wave speeds of the cycle's last span decide the next cycle's levels and smallest step. The boundary tally
sees every smallest step through ``substep_hook``.
"""

from __future__ import annotations

from collections.abc import Callable

import numpy as np

from wadi2d_amd.config import Numerics
from wadi2d_amd.solver.levels import CyclePlan, LevelRule, assign_levels, first_level_due, graded, plan_cycle
from wadi2d_amd.solver.stepper import SLEEP_CHECKS, NumericalFailure, Stepper

CLIMB_EVERY = 16
"""Cycles between chances for a tile to climb a level: new lists cost more than a late climb (level_stepper.py:23)."""


class _DeviceLists:
    """A cycle plan's lists as device list handles (None for an empty list: no launch)."""

    def __init__(self, ops, plan: CyclePlan) -> None:
        ops.clear_lists()

        def put(a: np.ndarray):
            return ops.lists(a) if a.size else None

        self.groups = [put(a) for a in plan.groups]
        self.upto = [put(a) for a in plan.upto]
        self.halo = [put(a) for a in plan.halo]
        self.coarse = [put(a) for a in plan.coarse]
        self.edges = [(put(x), put(y)) for x, y in zip(plan.edges_x, plan.edges_y, strict=True)]
        self.strips = [put(a) for a in plan.strips]
        self.sizes = [int(a.size) for a in plan.groups]


class LevelStepper(Stepper):
    """Each tile steps a power of two of the smallest step. ``pin`` keeps tiles at the smallest step (the boundary
    tally counts every smallest step there)."""

    def __init__(self, backend, grid, num: Numerics) -> None:
        super().__init__(backend, grid, num)
        self.ops = backend.level_ops()
        top = num.lts_levels if num.local_time_stepping else 0
        self.rule = LevelRule(num.cfl, num.lts_wave_floor, num.lts_buffer, num.dt_max, top)
        self.pinned = np.zeros(grid.ntiles, bool)
        self.level = np.zeros(grid.ntiles, np.int32)
        self.plan: CyclePlan | None = None
        self.lists: _DeviceLists | None = None
        self.substep_hook: Callable[[float], None] | None = None
        self.since_active = 0
        self.fresh_wave = False
        self.cycles = 0
        self.tile_steps = 0

    def pin(self, tiles: np.ndarray) -> None:
        self.pinned[tiles] = True

    def check_period(self) -> int:
        """Smallest steps between list checks at most: a check ends the first cycle at or past ``active_every``
        smallest steps, its wet flags written at the start of that cycle's last span (level_stepper.py:74-80)."""
        n = self.num
        span = 1 << self.rule.top
        cycle = span * ((max(n.dt_every, span) + span - 1) // span)
        return int(n.active_every) + cycle + span

    def rebuild_active(self, retain: bool = False, check: bool = False) -> None:
        """The base rule; a tile that joins has its wet share rewritten (every halo is read live, so no refill), and a
        changed list asks for a new plan."""
        b, n = self.backend, self.grid.ntiles
        before = np.array(b.active_tiles, copy=True)
        if not self.adaptive_active:
            if int(b.nact) != n - 1:
                b.set_active(np.arange(1, n, dtype=np.int32))
        else:
            want = self._wanted()
            now = np.zeros(n, bool)
            now[b.active_tiles] = True
            if retain:
                stay = now
            elif check:
                self.idle[want] = 0
                self.idle[~want] += 1
                stay = now & (self.idle < SLEEP_CHECKS)
            else:
                self.idle[:] = 0
                stay = np.zeros(n, bool)
            act = want | stay
            act[0] = False
            b.set_active(np.flatnonzero(act).astype(np.int32))
            joined = act & ~now
            if (retain or check) and joined.any():
                b.refresh_plan(np.flatnonzero(joined).astype(np.int32))
        if not np.array_equal(before, b.active_tiles):
            self.plan = None

    def _replan(self) -> None:
        active = np.asarray(self.backend.active_tiles, np.int32)
        self.level = graded(self.level, active, self.grid.nbr).astype(np.int32)
        self.plan = plan_cycle(self.level, active, self.grid.nbr)
        self.lists = _DeviceLists(self.ops, self.plan)
        self.ops.set_level(self.level)

    def _reassign(self) -> None:
        """New levels and smallest step from the waves of the cycle that just ended (level_stepper.py:112-135)."""
        wave = self.ops.wave()
        if not np.all(np.isfinite(wave)):
            raise NumericalFailure(f"non-finite wave speed at t = {self.t:.3f} s")
        active = np.asarray(self.backend.active_tiles)
        held = self.level.copy()
        on = np.zeros(self.pinned.size, bool)
        on[active] = True
        self.cycles += 1
        self.level, self.dt = assign_levels(wave, active, self.grid.nbr, self.grid.dx, self.rule, self.pinned & on,
                                            held, self.cycles % CLIMB_EVERY == 0)  # fmt: skip
        if self.dt < self.num.dt_min:
            raise NumericalFailure(f"time step {self.dt:.2e} s below the minimum at t = {self.t:.3f} s")
        if not np.array_equal(held, self.level):
            self.plan = None

    def _substep(self, n: int, dt0: float, wave_step: bool, wet_due: bool) -> None:
        """Smallest step n of the cycle (level_stepper.py:137-194)."""
        ops, L = self.ops, self.lists
        k = first_level_due(n, self.plan.top)
        if L.upto[k] is None:
            return
        due = range(k + 1)
        dts = [dt0 * (1 << a) for a in range(k + 1)]
        ops.halo_state(L.halo[k], n, wave_step)
        if wave_step:
            ops.clear_wave()
        for a in due:
            if L.groups[a] is not None:
                ops.faces(L.groups[a], dts[a], wave_step)
        for a in due:
            if L.strips[a] is not None:
                ops.strips(L.strips[a], dts[a], wave_step)
        ops.faces_done(L.upto[k])
        for a in due:
            if L.groups[a] is not None:
                ops.limit(L.groups[a], dts[a])
        ops.limit_done(L.upto[k])
        for a in due:
            if L.coarse[a] is not None:
                ops.save(L.coarse[a])
            for rows in L.edges[a]:
                if rows is None:
                    continue
                if a < k:  # the coarse tile begins its step too: the edge's estimate
                    ops.estimate(rows)
                else:  # the coarse tile is in the middle of its step: what this fine step owes it
                    ops.flux(rows, dts[a])
        if wave_step and wet_due:
            ops.clear_wet()
        for a in due:
            if L.groups[a] is not None:
                ops.update(L.groups[a], dts[a], wave_step)
        ops.boundary()
        ops.commit(L.upto[k])
        for a in due:
            if L.strips[a] is not None:
                ops.commit_strips(L.strips[a])
        if self.substep_hook is not None:
            self.substep_hook(dt0)
        self.nsteps += 1
        self.tile_steps += sum(L.sizes[a] for a in due)

    def advance(self, t_end: float, on_step: Callable[[Stepper], None] | None = None) -> float:
        """Whole cycles to ``t_end`` exactly: the last cycle's smallest step is cut so that it lands there
        (level_stepper.py:196-222). The first cycle is one smallest step, to measure the waves."""
        num = self.num
        while self.t < t_end - 1e-9:
            if self.plan is None:
                self._replan()
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


def tile_step_share(st: LevelStepper) -> float:
    """Tile-steps taken over what a uniform step would have taken with the same tiles (level_stepper.py:225-228)."""
    return st.tile_steps / max(st.nsteps * max(int(st.backend.nact), 1), 1)
