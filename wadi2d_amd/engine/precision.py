"""One run of a scenario with a start depth and no rain, for the FP32-against-FP64 study: max depth (sampled every
step), peak outfall discharge (max over steps of the signed sum over the free boundary's open faces) and the FP64
books (engine/monitor.py:35-79), through the host-driven window loop of run.py:1712-1739."""

from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np

from wadi2d_amd.engine.driver import Driver
from wadi2d_amd.engine.forcing import Rain, boundary_set
from wadi2d_amd.solver.model import Model
from wadi2d_amd.solver.state import from_state


@dataclass
class StudyRun:
    max_depth: np.ndarray
    peak_q: float
    mass_error: float
    steps: int
    wall_s: float


def run_study(scen, h0: np.ndarray, backend: str, log=None) -> StudyRun:
    tables = scen.build()
    bset = boundary_set(scen.boundaries, tables)
    rain = Rain(tables, bset, scen.cumulative_mm, scen.weight)
    model = Model(tables, scen.numerics, backend)
    model.set_depth(h0)
    drv = Driver(model, bset, rain, maxima_every=1, peak_flow=True)
    t0 = time.perf_counter()
    t = 0.0
    while t < scen.duration_s - 1e-9:
        t = min(t + 600.0, scen.duration_s)
        drv.run(t)
        if log is not None:
            log(f"[study] {backend}: t={model.stepper.t:.0f} s, {model.stepper.nsteps} steps, dt={model.stepper.dt:.3f}")
    model.backend.synchronize()
    wall = time.perf_counter() - t0
    dmax = from_state(tables.grid, model.backend.maxima_get(0)).astype(np.float64)
    return StudyRun(dmax, drv.peak_q, drv.accounts().relative, model.stepper.nsteps, wall)
