"""The host time loop over a backend (behaviour of wadi2d/solver/stepper.py:14-133).

Each step takes dt from the latest wave-speed estimate, clipped to the time left and to ``dt_max``. The backend
reports a fresh maximum wave speed on every ``dt_every``-th step (and on the step after the host changes the
state or a source); every ``active_every``-th step the set of stepped tiles is decided again. Storms, files and
outputs live with the caller, which sees each step through ``on_step``.
"""

from __future__ import annotations

from collections.abc import Callable

import numpy as np

from wadi2d_amd.config import Numerics

SLEEP_CHECKS = 2
"""A tile stops being stepped only when this many activity checks in a row found neither it nor a neighbour wet or
held awake: by the second such check its face buffers and wave speed already hold what a dry tile's would."""


class NumericalFailure(RuntimeError):
    """The time step fell below ``dt_min``, or the wave speed was not a finite number."""


class Stepper:
    """Clock, step count and tile activity of one backend. ``keep_active`` marks tiles stepped whatever their water
    (sources, boundaries); ``idle`` counts, per tile, the checks in a row that did not want it; ``refresh_due``
    asks the next step for a new wave-speed estimate. With ``adaptive_active`` False every tile is stepped."""

    def __init__(self, backend, grid, num: Numerics) -> None:
        self.backend, self.grid, self.num = backend, grid, num
        self.t = 0.0
        self.nsteps = 0
        self.dt = num.cfl * grid.dx / num.wave_floor  # until the first estimate: the floor wave speed
        self.refresh_due = True
        self.adaptive_active = True
        self.keep_active = np.zeros(grid.ntiles, bool)
        self.idle = np.zeros(grid.ntiles, np.int32)

    def pin(self, tiles: np.ndarray) -> None:
        """Tiles that must see every smallest step: every tile does in the uniform step (stepper.py:49-50)."""

    def _wanted(self) -> np.ndarray:
        """Tiles that must be stepped now: each wet or held tile and its eight neighbours (water crosses less than a
        cell per step, so less than a tile between two checks)."""
        seed = self.backend.wet_flags().astype(bool) | self.keep_active
        seed[0] = False
        want = seed.copy()
        want[self.grid.nbr[seed]] = True
        return want

    def rebuild_active(self, retain: bool = False, check: bool = False) -> None:
        """Decide the stepped tiles again (stepper.py:56-92).

        ``retain`` (the host changed a source or a boundary): the tiles stepped now stay, since the wet flags may be
        older than the state. ``check`` (the time loop's own call): a tile no longer wanted stays until its
        ``SLEEP_CHECKS``-th check in a row. Neither (the host wrote the whole state): the set is made afresh. A tile
        that joins under ``retain`` or ``check`` was dry while asleep; its wet share is rewritten by the device and
        every ring is refilled, so it stands as if it had never slept.
        """
        b, n = self.backend, self.grid.ntiles
        if not self.adaptive_active:
            if int(b.nact) != n - 1:
                b.set_active(np.arange(1, n, dtype=np.int32))
            return
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
            b.refresh_halos()

    def _take_estimate(self, wave: float) -> None:
        """The next steps' dt from a maximum wave speed: cfl dx over the speed (at least the floor), at most dt_max."""
        num = self.num
        if not np.isfinite(wave):
            raise NumericalFailure(f"non-finite wave speed at t = {self.t:.3f} s")
        self.dt = min(num.cfl * self.grid.dx / max(wave, num.wave_floor), num.dt_max)
        if self.dt < num.dt_min:
            raise NumericalFailure(f"time step {self.dt:.2e} s below the minimum at t = {self.t:.3f} s")

    def advance(self, t_end: float, on_step: Callable[[Stepper], None] | None = None) -> float:
        """Step until the clock reads ``t_end`` (the last step is cut to land on it) (stepper.py:101-129)."""
        b, num = self.backend, self.num
        while self.t < t_end - 1e-9:
            k = self.nsteps
            estimate = self.refresh_due or k % num.dt_every == 0
            recheck = k % num.active_every == 0
            self.refresh_due = False
            dt = min(self.dt, t_end - self.t, num.dt_max)
            b.step(dt, estimate, recheck)
            self.t += dt
            self.nsteps = k + 1
            if estimate:
                self._take_estimate(b.max_wave())
            if recheck and self.adaptive_active:
                self.rebuild_active(check=True)
            if on_step is not None:
                on_step(self)
        return self.t
