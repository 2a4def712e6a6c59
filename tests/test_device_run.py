"""The time loop on the device (wadi2d_amd/engine/device_run.py) against the host-driven loop and the oracle.

The device loop must take the same steps as the host loop (solver/stepper.py:101-129 driven by the window loop of
run.py:1712-1739): same step count, same active tiles, and in the exact float32 variant the same fields bit for bit.
Run through it, the plain 256 x 256 case must meet the FP64 oracle's limits as the host loop does."""

from __future__ import annotations

import numpy as np
import pytest

from tests.cases import sloped_valley
from tests.runs import run_oracle
from wadi2d_amd import synth
from wadi2d_amd.engine.device_run import DeviceRun, MemberPlan
from wadi2d_amd.engine.driver import Driver
from wadi2d_amd.engine.forcing import Boundary, Rain, boundary_set, uniform_storm
from wadi2d_amd.solver.model import Model
from wadi2d_amd.solver.state import from_state

FIELDS = ("h", "d", "u", "v", "mx", "my", "qx", "qy", "res", "om", "src")


def _valley_setup(backend):
    tables = sloped_valley()
    ny, nx = tables.grid.shape
    east = Boundary("free", np.arange(1, ny - 1), np.full(ny - 2, nx - 2))
    bset = boundary_set([east], tables)
    weight = np.zeros(tables.grid.shape)
    weight[:, : nx // 3] = 1.0
    rain = Rain(tables, bset, uniform_storm(40.0, 1800.0), weight)
    m = Model(tables, None, backend)
    m.set_depth(np.zeros(tables.grid.shape))
    return tables, bset, rain, m


def _device_plan(m, bset, rain, t_end):
    st = m.stepper
    return MemberPlan(1, m.grid.ntiles, rain, bset, st.t, t_end, st.dt, st.nsteps, st.refresh_due,
                      rain.state)  # fmt: skip


@pytest.mark.gpu
@pytest.mark.parametrize("graph", [0, 16])
def test_device_loop_equals_the_host_loop_bit_for_bit(graph):
    t_end = 2400.0
    tables, bset, rain, host = _valley_setup("hip-exact")
    drv = Driver(host, bset, rain)
    drv.run(t_end)
    _, bset2, rain2, dev = _valley_setup("hip-exact")
    Driver(dev, bset2, rain2)  # the same boundary, source weights, tally and storage setup
    run = DeviceRun(dev.backend, [_device_plan(dev, bset2, rain2, t_end)], dev.num, idle=dev.stepper.idle)
    # exactly the host's steps: the member has reached t_end and is retired (its tiles leave the list) by the next
    s = run.steps(host.stepper.nsteps, graph)[0]
    print(f"graph {graph}: steps {s.nsteps}/{host.stepper.nsteps} t {s.t}/{host.stepper.t} dt {s.dt}/{host.stepper.dt}",
          flush=True)
    assert not s.done and not s.failed
    assert s.nsteps == host.stepper.nsteps
    assert s.t == host.stepper.t and s.dt == host.stepper.dt
    assert np.array_equal(np.sort(dev.backend.active_tiles), np.sort(host.backend.active_tiles))
    for f in FIELDS:
        a, b = dev.field(f), host.field(f)
        assert np.array_equal(a, b), f"{f}: {np.abs(a - b).max()}"
    assert dev.backend.tally_volume() == host.backend.tally_volume()
    bal_host = drv.accounts()
    assert s.rain == pytest.approx(bal_host.rain, rel=1e-12)
    assert host.stepper.nsteps > 300
    h = dev.field("h")
    s = run.steps(1, 0)[0]
    assert s.done and s.nsteps == host.stepper.nsteps and dev.backend.nact == 0
    assert np.array_equal(dev.field("h"), h)  # a retired member keeps its state


@pytest.mark.gpu
def test_plain_case_through_the_device_loop_meets_the_fp64_oracle():
    scen = synth.plain_cpu_case()
    ref = run_oracle(scen, prec=8, threads=16)
    tables = scen.build()
    bset = boundary_set(scen.boundaries, tables)
    rain = Rain(tables, bset, scen.cumulative_mm, scen.weight)
    m = Model(tables, scen.numerics, "hip-f64")
    m.set_depth(np.zeros(tables.grid.shape))
    Driver(m, bset, rain, maxima_every=1)
    run = DeviceRun(m.backend, [_device_plan(m, bset, rain, scen.duration_s)], m.num, maxima=1,
                    idle=m.stepper.idle)  # fmt: skip
    s = run.run(chunk=512, graph_steps=64)[0]
    bal = run.balance(0, m.backend.tally_volume())
    dmax = from_state(tables.grid, m.backend.maxima_get(0)).astype(np.float64)
    err = np.abs(dmax - ref.max_depth).max()
    print(f"steps {s.nsteps}/{ref.steps} mass {bal.relative:.2e} max |d max depth| {err:.3e}", flush=True)
    assert s.nsteps == ref.steps
    assert err <= 1e-3
    assert abs(bal.relative) <= 1e-7
