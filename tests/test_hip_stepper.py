"""The stepper's edge cases on the HIP path, as the reference's own tests state them for its backends: a NaN stops
the run (tests/solver/test_review_regressions.py:96-105), the first step on dry ground is no longer than
cfl dx / wave_floor (:66-69), an empty grid runs (:89-93), a step below the minimum stops the run
(stepper.py:117-124); and the device-resident loop reports the same failures through its member state."""

from __future__ import annotations

import numpy as np
import pytest

from wadi2d_amd.config import Numerics, Tables
from wadi2d_amd.grid import TileGrid
from wadi2d_amd.tables import ArrayBed, build_tables


def _flat(n=30, dx=10.0):
    grid = TileGrid.from_mask(np.ones((n, n), bool), 0.0, 0.0, dx)
    return build_tables(grid, ArrayBed(np.zeros(grid.shape), dx), Tables(), 0.03), grid


@pytest.mark.gpu
@pytest.mark.parametrize("backend", ["hip", "hip-exact", "hip-f64"])
def test_nan_stops_the_run(backend):
    from wadi2d_amd.solver.model import Model
    from wadi2d_amd.solver.stepper import NumericalFailure

    tables, grid = _flat()
    m = Model(tables, Numerics(), backend)
    m.set_depth(np.where(grid.mask, 1.0, 0.0))
    u = m.backend.get_field("u")
    u[m.cell_state_index(10, 10)] = float("nan")
    m.backend.set_field("u", u)
    m.stepper.refresh_due = True
    with pytest.raises(NumericalFailure):
        m.advance(5.0)


@pytest.mark.gpu
def test_first_step_on_dry_ground_and_an_empty_grid():
    from wadi2d_amd.solver.model import Model

    tables, _ = _flat(30, 20.0)
    m = Model(tables, Numerics(), "hip")
    m.set_depth(np.zeros(tables.grid.shape))
    assert m.stepper.dt <= m.num.cfl * 20.0 / m.num.wave_floor + 1e-9
    m.advance(10.0)
    assert m.stepper.t == pytest.approx(10.0)
    assert m.backend.nact == 0 or not m.field("h").any()


@pytest.mark.gpu
def test_a_step_below_the_minimum_stops_the_run():
    from wadi2d_amd.solver.model import Model
    from wadi2d_amd.solver.stepper import NumericalFailure

    tables, grid = _flat()
    h = np.where(grid.mask, 1.0, 0.0)
    h[:, : grid.shape[1] // 2] = 20.0 * grid.mask[:, : grid.shape[1] // 2]
    m = Model(tables, Numerics(dt_min=1.0), "hip")  # 20 m of water: dt ~ 0.9 * 10 / 14 s, below the minimum
    m.set_depth(h)
    with pytest.raises(NumericalFailure):
        m.advance(60.0)


@pytest.mark.gpu
def test_the_device_loop_reports_a_failed_member():
    from wadi2d_amd.engine.device_run import DeviceRun, MemberPlan
    from wadi2d_amd.engine.forcing import Rain, boundary_set, uniform_storm
    from wadi2d_amd.solver.model import Model

    tables, grid = _flat()
    m = Model(tables, Numerics(), "hip")
    m.set_depth(np.where(grid.mask, 1.0, 0.0))
    u = m.backend.get_field("u")
    u[m.cell_state_index(10, 10)] = float("nan")
    m.backend.set_field("u", u)
    bset = boundary_set([], tables)
    rain = Rain(tables, bset, uniform_storm(0.0, 600.0))
    m.backend.source_weights(rain.state, np.zeros(rain.state.size))
    m.backend.tally_set(np.zeros(0, np.int64), np.zeros(0, np.int32), np.zeros(0, np.float32))
    plan = MemberPlan(1, grid.ntiles, rain, bset, 0.0, 600.0, m.stepper.dt, 0, True)
    run = DeviceRun(m.backend, [plan], m.num, idle=m.stepper.idle)
    s = run.run(chunk=64, graph_steps=16)[0]
    print(f"failed {s.failed} after {s.nsteps} steps at t={s.t}", flush=True)
    # a non-finite face velocity reports a wave speed of 1e30 (faces.cu:133): the next step falls below the minimum
    assert s.done and s.failed in (1, 2)
    assert s.nsteps <= 2
