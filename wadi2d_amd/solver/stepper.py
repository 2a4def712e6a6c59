"""Stepper: the one time loop (wadi2d/solver/stepper.py:14-133). It knows nothing of storms or files."""

from __future__ import annotations

from collections.abc import Callable

import numpy as np

from wadi2d_amd.config import Numerics

SLEEP_CHECKS = 2
"""Checks in a row that must find a tile and its surroundings dry before it leaves the active list."""


class NumericalFailure(RuntimeError):
    """The time step collapsed or a NaN appeared."""


class Stepper:
    def __init__(self, backend, grid, num: Numerics) -> None:
        self.backend = backend
        self.grid = grid
        self.num = num
        self.t = 0.0
        self.nsteps = 0
        self.adaptive_active = True
        self.keep_active = np.zeros(grid.ntiles, bool)
        self.refresh_due = True
        self.dt = num.cfl * grid.dx / num.wave_floor
        self.idle = np.zeros(grid.ntiles, np.int32)

    def rebuild_active(self, retain: bool = False, check: bool = False) -> None:
        """Active: a wet tile, a tile with a source, and all their neighbours (stepper.py:56-92)."""
        ntiles = self.grid.ntiles
        b = self.backend
        if not self.adaptive_active:
            if int(b.nact) != ntiles - 1:
                b.set_active(np.arange(1, ntiles, dtype=np.int32))
            return
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
            act |= current & (self.idle < SLEEP_CHECKS)
        else:
            self.idle[:] = 0
        act[0] = False
        woken = act & ~current
        b.set_active(np.nonzero(act)[0].astype(np.int32))
        if (retain or check) and woken.any():
            b.refresh_plan(np.nonzero(woken)[0].astype(np.int32))
            b.refresh_halos()

    def advance(self, t_end: float, on_step: Callable[[Stepper], None] | None = None) -> float:
        """Run to ``t_end`` exactly (stepper.py:101-129)."""
        b = self.backend
        num = self.num
        dx = self.grid.dx
        while self.t < t_end - 1e-9:
            refresh = self.refresh_due or self.nsteps % num.dt_every == 0
            self.refresh_due = False
            activate = self.nsteps % num.active_every == 0
            dt = min(self.dt, t_end - self.t, num.dt_max)
            b.step(dt, refresh, activate)
            self.t += dt
            self.nsteps += 1
            if refresh:
                w = b.max_wave()
                if not np.isfinite(w):
                    raise NumericalFailure(f"non-finite wave speed at t = {self.t:.3f} s")
                self.dt = min(num.cfl * dx / max(w, num.wave_floor), num.dt_max)
                if self.dt < num.dt_min:
                    raise NumericalFailure(f"time step {self.dt:.2e} s below the minimum at t = {self.t:.3f} s")
            if activate and self.adaptive_active:
                self.rebuild_active(check=True)
            if on_step is not None:
                on_step(self)
        return self.t
