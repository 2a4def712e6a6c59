"""Whole runs of a scenario through the oracle and through the HIP path, and the quantities compared between them:
the per-cell maximum depth (sampled every step), the peak outfall discharge (max over steps of the sum of sign * q
over the free boundary's open faces) and the relative mass error of the FP64 books (monitor.py:59-78)."""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from oracle.oracle import Ledger, OracleModel
from oracle.oracle import Numerics as ONumerics
from wadi2d_amd.engine.driver import Driver
from wadi2d_amd.engine.forcing import Rain, boundary_set
from wadi2d_amd.solver.model import Model
from wadi2d_amd.solver.state import from_state


@dataclass
class RunResult:
    max_depth: np.ndarray  # (ny, nx)
    peak_q: float
    mass_error: float
    steps: int
    h: np.ndarray


def run_oracle(scen, prec: int = 8, threads: int = 8, t_end: float | None = None,
               h0: np.ndarray | None = None) -> RunResult:  # fmt: skip
    """The run loop of run._run (run.py:1712-1739) over the oracle, with its ledger; dry or from mean depth h0."""
    tables = scen.build()
    bset = boundary_set(scen.boundaries, tables)
    rain = Rain(tables, bset, scen.cumulative_mm, scen.weight)
    m = OracleModel(tables, ONumerics.like(scen.numerics), prec=prec, threads=threads)
    m.set_depth(np.zeros(tables.grid.shape) if h0 is None else h0)
    if bset.size:
        m.set_boundary(bset.cells)
    interior = tables.grid.mask & ~bset.mask
    ledger = Ledger(m, interior, bset.xfaces, bset.xsign, bset.yfaces, bset.ysign)
    end = scen.duration_s if t_end is None else t_end
    t, dry = 0.0, False
    while t < end - 1e-9:
        t_next = min(t + 30.0, end)
        rate, vol = rain.rate(t, t_next)
        if vol > 0.0 or not dry:
            m.set_source_cells(rain.state, rate)
        dry = vol <= 0.0
        ledger.open_window(t, t_next, vol, 0.0)
        m.advance(t_next, ledger.on_step)
        ledger.settle(t_next)
        t = t_next
    bal = ledger.accounts(t)
    return RunResult(from_state(tables.grid, ledger.dmax), ledger.peak_q, bal.relative, m.stepper.nsteps,
                     m.field("h"))  # fmt: skip


def run_hip(scen, backend: str = "hip-f64", t_end: float | None = None) -> RunResult:
    tables = scen.build()
    bset = boundary_set(scen.boundaries, tables)
    rain = Rain(tables, bset, scen.cumulative_mm, scen.weight)
    m = Model(tables, scen.numerics, backend)
    m.set_depth(np.zeros(tables.grid.shape))
    drv = Driver(m, bset, rain, maxima_every=1, peak_flow=True)
    drv.run(scen.duration_s if t_end is None else t_end)
    bal = drv.accounts()
    dmax = from_state(tables.grid, m.backend.maxima_get(0))
    return RunResult(dmax.astype(np.float64), drv.peak_q, bal.relative, m.stepper.nsteps, m.field("h"))
