"""The excess form of the rain source on the device (w2d_source_losses / w2d_source_excess / w2d_run_terms) against
the host restatement of forcing.py:644-678 (``Rain.rate``) and the oracle.

* One window's rates, formed on the device in float64 and cast once: equal to ``Rain.rate``'s float32 rates bit for
  bit (curve numbers from 55 to 100 and none, a depth grid, windows before, across and after the initial
  abstraction and after the storm), their FP64 sum to 1e-12, the kept tiles exactly those with a non-zero rate.
* The host-driven window loop (engine/driver.py) on the exact FP32 variant against the oracle in the reference's run
  loop (run.py:1712-1739) over the same rates: the same steps and the same fields bit for bit.
* The device-resident loop against the host-driven loop: the same steps, active tiles and fields bit for bit.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle.oracle import OracleModel
from oracle.oracle import Numerics as ONumerics
from tests.cases import sloped_valley
from wadi2d_amd.engine.device_run import DeviceRun, MemberPlan
from wadi2d_amd.engine.driver import Driver
from wadi2d_amd.config import Numerics
from wadi2d_amd.engine.forcing import Boundary, Rain, boundary_set, uniform_storm
from wadi2d_amd.layout import STRIDE2
from wadi2d_amd.solver.model import Model
from wadi2d_amd.solver.state import state_index

from test_losses import frames_over

pytestmark = pytest.mark.gpu

FIELDS = ("h", "d", "u", "v", "mx", "my", "qx", "qy", "res", "src")


def _valley(form: str):
    """The crested valley with curve numbers (55-98, 100 on a strip, none on the upper valley) under one of three
    storms: 'uniform' (50 mm over 30 min), 'depth_grid' (30-80 mm per cell), 'gridded' (seeded frames of 170 m
    pixels, a gap and a missing frame value, area factor 0.9)."""
    tables = sloped_valley()
    ny, nx = tables.grid.shape
    east = Boundary("free", np.arange(1, ny - 1), np.full(ny - 2, nx - 2))
    bset = boundary_set([east], tables)
    rng = np.random.default_rng(11)
    cn = rng.uniform(55.0, 98.0, (ny, nx)).round(1)
    cn[:, : nx // 4] = np.nan  # no losses on the upper valley
    cn[: ny // 3, nx // 2 : nx // 2 + 20] = 100.0
    if form == "gridded":
        frames = frames_over(tables.grid)
        rain = Rain(tables, bset, None, curve_number=cn, frames=frames, area_factor=0.9)
    elif form.startswith("gauges"):
        rain = Rain(tables, bset, None, curve_number=cn, gauges=gauges_over(tables.grid, form.split("_")[1]))
    else:
        depths = rng.uniform(30.0, 80.0, (ny, nx)) if form == "depth_grid" else None
        storm = uniform_storm(1.0 if depths is not None else 50.0, 1800.0)
        rain = Rain(tables, bset, storm, curve_number=cn, depth_grid=depths)
    return tables, bset, rain


def gauges_over(grid, method: str, seed: int = 4):
    """Five seeded gauges over and beside the grid, cumulative records every 10 min for 70 min (0-12 mm per period,
    one gauge dry for a period), spread by nearest gauge ('thiessen') or inverse distance ('idw')."""
    from wadi2d_amd.engine.gauges import INVERSE_DISTANCE, THIESSEN, GaugeStorm

    ny, nx = grid.shape
    rng = np.random.default_rng(seed)
    xy = np.column_stack([grid.x0 + rng.uniform(-0.1, 1.1, 5) * nx * grid.dx,
                          grid.y0 - rng.uniform(-0.1, 1.1, 5) * ny * grid.dx])  # fmt: skip
    inc = rng.uniform(0.0, 12.0, (7, 5))
    inc[3, 1] = 0.0
    cumulative = np.vstack([np.zeros(5), np.cumsum(inc, axis=0)])
    method = THIESSEN if method == "thiessen" else INVERSE_DISTANCE
    return GaugeStorm(tuple(f"g{k}" for k in range(5)), xy, np.arange(8) * 10.0, cumulative, method)


FORMS = ["uniform", "depth_grid", "gridded", "gauges_thiessen", "gauges_idw"]


def _window(b, rain, t0, t1):
    if rain.pixels is not None:
        d = rain.pixel_terms(t0, t1)
        return b.source_excess_pixels(d[0], d[1], rain.terms(t0, t1)[2])
    return b.source_excess(*rain.terms(t0, t1))


@pytest.mark.parametrize("form", FORMS)
def test_one_window_of_device_rates_is_the_host_restatement_bit_for_bit(form):
    tables, bset, rain = _valley(form)
    m = Model(tables, None, "hip-exact")
    m.set_depth(np.zeros(tables.grid.shape))
    Driver(m, bset, rain)  # w2d_source_losses
    b = m.backend
    print(f"excess rates: {form}", flush=True)
    jb, ib = np.nonzero(bset.mask)
    boundary = np.asarray(state_index(tables.grid, jb, ib), np.int64)
    assert boundary.size > 0
    seen_zero = seen_rain = False
    for t0 in (0.0, 120.0, 300.0, 600.0, 1200.0, 1770.0, 1800.0, 2100.0, 2400.0, 4500.0):
        t1 = t0 + 30.0
        want, vol = rain.rate(t0, t1)
        total, tiles = _window(b, rain, t0, t1)
        src = b.get_field("src")
        assert np.array_equal(src[rain.state], want), t0
        assert not np.any(src[boundary]), t0  # the boundary cells take no rain
        assert total * rain.area * (t1 - t0) == pytest.approx(vol, rel=1e-12, abs=1e-300)
        kept = np.zeros(tables.grid.ntiles, bool)
        kept[rain.state[want != 0.0] // STRIDE2] = True
        assert np.array_equal(tiles, kept), t0
        seen_zero |= vol == 0.0
        seen_rain |= vol > 0.0
    assert seen_zero and seen_rain


def _oracle_loop(tables, bset, rain, t_end):
    """The reference's run loop (run.py:1712-1739) over the oracle in float32, the rates from the host restatement."""
    m = OracleModel(tables, ONumerics.like(Numerics()), prec=4, threads=8)
    m.set_depth(np.zeros(tables.grid.shape))
    m.set_boundary(bset.cells)
    t, dry = 0.0, False
    while t < t_end - 1e-9:
        t_next = min(t + 30.0, t_end)
        rate, vol = rain.rate(t, t_next)
        if vol > 0.0 or not dry:
            m.set_source_cells(rain.state, rate)
        dry = vol <= 0.0
        m.advance(t_next)
        t = t_next
    return m


@pytest.mark.parametrize("form", FORMS)
def test_the_host_loop_with_losses_is_the_oracle_bit_for_bit(form):
    t_end = 2400.0 if form in ("uniform", "depth_grid") else 4800.0
    tables, bset, rain = _valley(form)
    print(f"host loop vs oracle, excess form: {form}", flush=True)
    ref = _oracle_loop(tables, bset, rain, t_end)
    m = Model(tables, None, "hip-exact")
    m.set_depth(np.zeros(tables.grid.shape))
    drv = Driver(m, bset, rain)
    drv.run(t_end)
    print(f"steps {m.stepper.nsteps}/{ref.stepper.nsteps} rain {drv.accounts().rain:.3f} m3", flush=True)
    assert m.stepper.nsteps == ref.stepper.nsteps
    for f in ("h", "d", "u", "v", "qx", "qy"):
        a, b = m.field(f), ref.field(f)
        assert np.array_equal(a, b), f"{f}: {np.abs(a.astype(np.float64) - b).max()}"
    assert drv.accounts().rain > 0.0 and (m.field("h") > 0.01).sum() > 500


def _plan(m, bset, rain, t_end):
    st = m.stepper
    return MemberPlan(1, m.grid.ntiles, rain, bset, st.t, t_end, st.dt, st.nsteps, st.refresh_due, rain.state)


@pytest.mark.parametrize("graph, form", [(0, "uniform"), (16, "uniform"), (16, "gridded"), (16, "gauges_idw")])
def test_the_device_loop_with_losses_is_the_host_loop_bit_for_bit(graph, form):
    t_end = 2400.0 if form in ("uniform", "depth_grid") else 4800.0
    tables, bset, rain = _valley(form)
    host = Model(tables, None, "hip-exact")
    host.set_depth(np.zeros(tables.grid.shape))
    drv = Driver(host, bset, rain)
    drv.run(t_end)
    _, bset2, rain2 = _valley(form)
    dev = Model(tables, None, "hip-exact")
    dev.set_depth(np.zeros(tables.grid.shape))
    Driver(dev, bset2, rain2)
    run = DeviceRun(dev.backend, [_plan(dev, bset2, rain2, t_end)], dev.num, idle=dev.stepper.idle)
    s = run.steps(host.stepper.nsteps, graph)[0]
    print(f"graph {graph}: steps {s.nsteps}/{host.stepper.nsteps} t {s.t}/{host.stepper.t}", flush=True)
    assert not s.failed and s.nsteps == host.stepper.nsteps
    assert s.t == host.stepper.t and s.dt == host.stepper.dt
    assert np.array_equal(np.sort(dev.backend.active_tiles), np.sort(host.backend.active_tiles))
    for f in FIELDS:
        a, b = dev.field(f), host.field(f)
        assert np.array_equal(a, b), f"{f}: {np.abs(a - b).max()}"
    assert s.rain == pytest.approx(drv.accounts().rain, rel=1e-12)
    assert dev.backend.tally_volume() == host.backend.tally_volume()
