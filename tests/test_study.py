"""The depth-area study as one batched launch (wadi2d/arf.py:328-388, whose runs go one after another): every storm
area a member of one device run over copies of the same grid. Each member must be exactly the run of its area alone,
and the host loop's run of that area (the reference's own way of running it), in exact float32."""

from __future__ import annotations

import numpy as np
import pytest

from wadi2d_amd import synth
from wadi2d_amd.engine.arf import ArfTable, ReducedStorm, depth_area_study
from wadi2d_amd.engine.driver import Driver
from wadi2d_amd.engine.forcing import Rain, boundary_set
from wadi2d_amd.engine.storms import ScsStorm
from wadi2d_amd.solver.model import Model
from wadi2d_amd.solver.state import from_state

pytestmark = pytest.mark.gpu

ARF_FM = ArfTable(
    np.array([1.0, 5.0, 10.0, 20.0, 30.0, 40.0, 50.0, 100.0, 200.0, 500.0, 1000.0, 2000.0]),
    np.array([0.999, 0.994, 0.987, 0.976, 0.965, 0.956, 0.948, 0.913, 0.864, 0.776, 0.694, 0.606]),
    reduces_below_table=False,
)
AREAS = [10.0, 100.0, 500.0]


def _ground():
    return synth.c0_style_case(128, 20.0, 4, 7, ScsStorm(80.0, 1.0), 3600.0, "study ground 128x128")


def test_a_study_in_one_launch_is_every_area_run_alone_and_the_host_loop():
    scen = _ground()
    res = depth_area_study(scen, ARF_FM, AREAS, "hip-exact", design=[("outlet", 50.0), ("far", 300.0)])
    print(f"study: factors {res.factors}, peaks {[round(q, 3) for q in res.peak_q]} m3/s at "
          f"{res.t_peak} s, steps {res.steps}, {res.wall_s:.2f} s", flush=True)
    assert res.factors == [ARF_FM.factor(a) for a in AREAS]
    assert res.peak_q[0] > res.peak_q[1] > res.peak_q[2] > 0.0  # less rain over a larger storm, a lower peak
    assert max(abs(m) for m in res.mass_error) <= 1e-6
    for k, area in enumerate(AREAS):
        alone = depth_area_study(scen, ARF_FM, [area], "hip-exact")
        assert alone.steps[0] == res.steps[k]
        assert np.array_equal(alone.max_depth[0], res.max_depth[k])
        assert alone.peak_q[0] == res.peak_q[k] and alone.t_peak[0] == res.t_peak[k]
        # the reference's way: one run of the area through the host loop
        tables = scen.build()
        bset = boundary_set(scen.boundaries, tables)
        rain = Rain(tables, bset, ReducedStorm(scen.cumulative_mm, ARF_FM.factor(area)))
        m = Model(tables, scen.numerics, "hip-exact")
        m.set_depth(np.zeros(tables.grid.shape))
        drv = Driver(m, bset, rain, maxima_every=1, peak_flow=True)
        drv.run(scen.duration_s)
        assert m.stepper.nsteps == res.steps[k]
        assert np.array_equal(from_state(tables.grid, m.backend.maxima_get(0)), res.max_depth[k])
        assert res.peak_q[k] == pytest.approx(drv.peak_q, rel=1e-12)
    outlet, far = res.rows
    assert (outlet["storm_area_low_km2"], outlet["storm_area_high_km2"]) == (10.0, 100.0)
    assert (far["storm_area_low_km2"], far["storm_area_high_km2"]) == (100.0, 500.0)
    assert res.peak_q[1] < outlet["design_peak_m3s"] < res.peak_q[0]
    assert res.peak_q[2] < far["design_peak_m3s"] < res.peak_q[1]
