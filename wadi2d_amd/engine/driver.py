"""The window loop of a run (wadi2d/run.py:1712-1739) with the water ledger (engine/monitor.py:35-79,197-248).

Every ``source_every_s`` the rain source is refreshed, its tiles kept active and a time-step refresh forced
(model.py:112-122); the model then advances to the window's end exactly. The boundary tally counts every step in
FP64 (maxima.py:149-192). Storage is the FP64 sum of h + res over the interior cells (monitor.py:214-218), and the
rain of each window is credited pro rata as the run advances through it (monitor.py:197-212).

The rain here is a weight per cell times one number per window, applied on the device (``source_scale``), so a
window costs one pass over the source array instead of a host computation and an upload of 16.8 M floats.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from wadi2d_amd.engine.forcing import BoundarySet, Rain
from wadi2d_amd.layout import STRIDE2


@dataclass
class Balance:
    """Water accounts of the interior cells, m3 (monitor.py:35-79)."""

    rain: float = 0.0
    inflow: float = 0.0
    outflow: float = 0.0
    storage0: float = 0.0
    storage: float = 0.0
    transmission: float = 0.0

    @property
    def error(self) -> float:
        return self.rain + self.inflow - self.transmission - self.outflow - (self.storage - self.storage0)

    @property
    def relative(self) -> float:
        scale = max(self.rain + self.inflow + max(-self.outflow, 0.0) + self.storage0, 1e-9)
        return self.error / scale


class StepBudgetSpent(Exception):
    """The driver's step budget ran out inside a window."""


class Driver:
    def __init__(self, model, boundary: BoundarySet, rain: Rain, source_every_s: float = 30.0,
                 maxima_every: int = 0, peak_flow: bool = False) -> None:  # fmt: skip
        self.model = model
        self.b = model.backend
        self.rain = rain
        self.every = float(source_every_s)
        self.area = model.grid.dx**2
        if boundary.size:
            model.set_boundary(boundary.cells)
        if rain.excess_form:
            n = rain.state.size
            self.b.source_losses(rain.state, rain.depths if rain.depths is not None else np.ones(n),
                                 rain.cn if rain.cn is not None else np.full(n, np.nan))  # fmt: skip
            if rain.pixels is not None:
                self.b.source_pixels(rain.state, rain.pixels.column, rain.pixels.used.size)
            self.source_tiles = None
        else:
            weight = rain.weight if rain.weight is not None else np.ones(rain.state.size)
            self.b.source_weights(rain.state, weight)
            self.source_tiles = np.unique(rain.state[weight != 0.0] // STRIDE2)
        faces = np.concatenate((boundary.xfaces, boundary.yfaces)).astype(np.int64)
        axis = np.concatenate((np.zeros(boundary.xfaces.size), np.ones(boundary.yfaces.size))).astype(np.int32)
        sign = np.concatenate((boundary.xsign, boundary.ysign)).astype(np.float32)
        self.b.tally_set(faces, axis, sign)
        self.b.storage_set(rain.state)
        self.maxima_every = int(maxima_every)
        if self.maxima_every:
            self.b.maxima_set(0.1, 0.05)
        self.peak_flow = peak_flow
        self.peak_q = 0.0
        self.balance = Balance(storage0=self.storage())
        self.window = (0.0, 1.0, 0.0)
        self.settled = 0.0
        self.last_t = model.stepper.t
        self.dry = False
        self.budget: int | None = None
        # the level stepper calls back on every smallest step, where the boundary faces carry their flow; the tally
        # counts there (monitor.py:111-113), not once per cycle
        self.counted_inside = hasattr(model.stepper, "substep_hook")
        if self.counted_inside:
            model.stepper.substep_hook = self.b.tally_add

    def storage(self) -> float:
        return self.b.storage_sum() * self.area

    def settle(self, t: float) -> None:
        t0, t1, rain = self.window
        share = (min(t, t1) - self.settled) / (t1 - t0)
        self.balance.rain += rain * share
        self.settled = min(t, t1)

    def on_step(self, st) -> None:
        """After every step: the boundary tally, the maxima every ``maxima_every`` steps, the step budget."""
        if not self.counted_inside:
            self.b.tally_add(st.t - self.last_t)
        self.last_t = st.t
        if self.peak_flow:
            self.peak_q = max(self.peak_q, self.b.tally_flow())
        if self.maxima_every and st.nsteps % self.maxima_every == 0:
            self.b.maxima_update(st.t)
        if self.budget is not None and st.nsteps >= self.budget:
            raise StepBudgetSpent

    def open_window(self, t: float, t_next: float) -> None:
        """The source of window [t, t_next]: its rate on the device, its tiles kept, a forced refresh, its rain."""
        if self.rain.excess_form:
            self._open_excess(t, t_next)
            return
        s = self.rain.scale(t, t_next)
        if s != 0.0 or not self.dry:
            self.b.source_scale(s)
            m, st = self.model, self.model.stepper
            keep = st.keep_active
            keep[:] = False
            if s != 0.0:
                keep[self.source_tiles] = True
            keep[m.boundary_tiles] = True
            st.rebuild_active(retain=True)
            st.refresh_due = True
        rain = self.b.source_sum() * self.area * (t_next - t) if s != 0.0 else 0.0
        self.dry = rain <= 0.0
        self.window = (t, t_next, rain)
        self.settled = t

    def _open_excess(self, t: float, t_next: float) -> None:
        """The excess form (run.py:1719-1729): every cell's rate formed on the device; the window opens (kept tiles:
        those with a non-zero rate, model.py:112-122) when its rain is positive or the last window's was."""
        if self.rain.pixels is not None:
            d = self.rain.pixel_terms(t, t_next)
            total, tiles = self.b.source_excess_pixels(d[0], d[1], self.rain.terms(t, t_next)[2])
        else:
            total, tiles = self.b.source_excess(*self.rain.terms(t, t_next))
        rain = total * self.area * (t_next - t)
        if rain > 0.0 or not self.dry:
            m, st = self.model, self.model.stepper
            keep = st.keep_active
            keep[:] = tiles
            keep[m.boundary_tiles] = True
            st.rebuild_active(retain=True)
            st.refresh_due = True
        self.dry = rain <= 0.0
        self.window = (t, t_next, rain)
        self.settled = t

    def run(self, t_end: float, max_steps: int | None = None) -> bool:
        """Advance window by window to ``t_end``; with ``max_steps`` stop once the stepper has taken that many steps
        in total. Returns True when t_end was reached."""
        st = self.model.stepper
        self.budget = max_steps
        try:
            while st.t < t_end - 1e-9:
                if st.t >= self.window[1] - 1e-9 or self.window == (0.0, 1.0, 0.0):
                    t_next = min((np.floor(st.t / self.every + 1e-9) + 1.0) * self.every, t_end)
                    self.open_window(st.t, t_next)
                self.model.advance(self.window[1], self.on_step)
                self.settle(st.t)
        except StepBudgetSpent:
            self.settle(st.t)
            return False
        finally:
            self.budget = None
        return True

    def accounts(self) -> Balance:
        self.settle(self.model.stepper.t)
        self.balance.outflow = self.b.tally_volume()
        self.balance.storage = self.storage()
        return self.balance
