"""A run with its time loop on the device: the stepper (solver/stepper.py:101-129) and the window loop
(run.py:1712-1739) of one model, or of an ensemble of models side by side in one set of arrays, advanced by
captured graphs of steps with no host in the loop.

Each member keeps the reference's sequence: a window opens every ``source_every_s`` (new rates, kept tiles, a
forced wave refresh), a step is ``min(dt, window end - t, dt_max)``, the wave speed is reduced every ``dt_every``
steps into the next ``dt``, the active list is checked every ``active_every`` steps. The window's source rates are
its precomputed scale times each cell's weight, float(weight * scale), as ``Rain.rate`` forms them on the host; in
the excess form (curve numbers, a depth grid, gridded rain) the device forms every cell's rate from the window's
storm depths (with gridded rain, its pixel's) when the window is due, and the window's rain decides whether it
opens.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from wadi2d_amd.engine.driver import Balance
from wadi2d_amd.engine.forcing import BoundarySet, Rain
from wadi2d_amd.solver.hip_backend import Member, RunNumerics


@dataclass
class MemberPlan:
    """One member's share of a device run."""

    tile0: int
    tile1: int
    rain: Rain
    boundary: BoundarySet
    t0: float
    t_end: float
    dt0: float
    nsteps0: int = 0
    refresh_due: bool = True
    storage_index: np.ndarray | None = None
    dry: bool = False  # the last window's rain was nothing (a resumed run's state)


def window_bounds(t0: float, t_end: float, every: float) -> list[tuple[float, float]]:
    """The windows of [t0, t_end] as the run loop opens them."""
    out = []
    t = t0
    while t < t_end - 1e-9:
        t_next = min(t + every, t_end)
        out.append((t, t_next))
        t = t_next
    return out


def window_scales(rain: Rain, t0: float, t_end: float, every: float) -> np.ndarray:
    """The scale of every window of [t0, t_end] (the weight form's; the excess form's windows carry 0)."""
    return np.asarray([0.0 if rain.excess_form else rain.scale(a, b) for a, b in window_bounds(t0, t_end, every)],
                      np.float64)  # fmt: skip


def window_terms(rain: Rain, t0: float, t_end: float, every: float) -> np.ndarray:
    """(d0, d1, 1000 (t1 - t0)) of every window of [t0, t_end]: the excess form's windows as the device takes them."""
    return np.asarray([rain.terms(a, b) for a, b in window_bounds(t0, t_end, every)], np.float64).reshape(-1, 3)


class DeviceRun:
    def __init__(self, backend, plans: list[MemberPlan], num, source_every_s: float = 30.0, maxima: int = 0,
                 idle: np.ndarray | None = None, pipeline: bool | None = None) -> None:  # fmt: skip
        """``maxima``: sample the running maxima after every step whose count is a multiple of it (0: never).
        ``pipeline``: take the pipelined step (a step's cell update made by the next step's kernel, with its faces)
        where the members allow it, or not; None keeps the library's default (on in FP32, off in FP64). The fields
        are the same either way."""
        self.b = backend
        self.plans = plans
        self.area = backend.grid.dx ** 2
        scales, members = [], []
        off = 0
        for p in plans:
            sc = window_scales(p.rain, p.t0, p.t_end, source_every_s)
            members.append(Member(p.tile0, p.tile1, p.t0, p.t_end, p.dt0, source_every_s, sc.size, off, p.nsteps0,
                                  int(p.refresh_due), int(p.dry)))  # fmt: skip
            scales.append(sc)
            off += sc.size
        self.scales = np.concatenate(scales) if scales else np.zeros(0)
        excess = {p.rain.excess_form for p in plans}
        if len(excess) > 1:
            raise ValueError("a device run's members take one source form: all weighted, or all in the excess form")
        if excess == {True}:
            backend.run_terms(np.concatenate([window_terms(p.rain, p.t0, p.t_end, source_every_s) for p in plans]))
            if plans[0].rain.pixels is not None:
                if len(plans) != 1:
                    raise ValueError("gridded rain runs one member per device run")
                p = plans[0]
                rows = [p.rain.pixel_terms(t0, t1) for t0, t1 in window_bounds(p.t0, p.t_end, source_every_s)]
                backend.run_pixel_terms(np.asarray(rows).reshape(len(rows), 2, -1))
        rn = RunNumerics(num.cfl, num.dt_max, num.dt_min, num.wave_floor, num.dt_every, num.active_every, int(maxima))
        if pipeline is not None:
            backend.run_pipeline(pipeline)
        backend.run_setup(members, self.scales, rn, idle)
        self.state = backend.run_read()
        self.storage0 = [self._storage(p) for p in plans]

    def _storage(self, p: MemberPlan) -> float:
        if p.storage_index is None:
            return 0.0
        self.b.storage_set(p.storage_index)
        return self.b.storage_sum() * self.area

    def steps(self, n: int, graph_steps: int = 64) -> list:
        self.b.run_steps(n, graph_steps)
        self.state = self.b.run_read()
        return self.state

    def run(self, chunk: int = 256, graph_steps: int = 64, on_chunk=None) -> list:
        """Step until every member is done or has failed (a failed member held at a stop time is not retired)."""
        while not all(s.done or s.failed for s in self.state):
            self.steps(chunk, graph_steps)
            if on_chunk is not None:
                on_chunk(self.state)
        return self.state

    def balance(self, m: int, outflow: float) -> Balance:
        p, s = self.plans[m], self.state[m]
        return Balance(rain=s.rain, outflow=outflow, storage0=self.storage0[m], storage=self._storage(p))
