"""The pipelined step of device-resident runs against the two-kernel step, through the C-ABI.

With the pipelined step a step's cell update is made by the next step's kernel, together with that step's faces, and
each tile recomputes the update of its west column and north row instead of reading it back. Nothing of the scheme
changes (cells.cu:4-111, faces.cu:244-271 in the reference's order, solver/stepper.py:101-129 around them): every
field, running maximum, tally total, clock and book of a run must be the two-kernel run's, bit for bit, whatever the
run asks of it (sleeping tiles, boundary cells, maxima every step or every fourth, graphs or single launches, a member
held while another steps)."""

from __future__ import annotations

import numpy as np
import pytest

from tests.test_device_run import _device_plan, _valley_setup
from wadi2d_amd.engine.device_run import DeviceRun
from wadi2d_amd.engine.driver import Driver
from wadi2d_amd.engine.ensemble import Ensemble, MemberSpec, setup_batched
from wadi2d_amd.solver.state import from_state

FIELDS = ("h", "d", "u", "v", "mx", "my", "qx", "qy", "res", "om", "src")
STATE = ("t", "dt", "rain", "nsteps", "done", "failed", "dt_min", "q_peak", "t_peak")


def _valley_run(backend, pipeline, maxima, t_end, chunks):
    _, bset, rain, m = _valley_setup(backend)
    Driver(m, bset, rain, maxima_every=max(maxima, 1))
    run = DeviceRun(m.backend, [_device_plan(m, bset, rain, t_end)], m.num, maxima=maxima, idle=m.stepper.idle,
                    pipeline=pipeline)  # fmt: skip
    for n, graph in chunks:
        s = run.steps(n, graph)
    return m, s[0]


def _same_state(a, b):
    for k in STATE:
        assert getattr(a, k) == getattr(b, k), k


def _same_model(ma, mb, grid):
    for f in FIELDS:
        x, y = ma.field(f), mb.field(f)
        assert np.array_equal(x, y), f"{f}: {np.abs(x - y).max()}"
    for which in range(5):
        x = from_state(grid, ma.backend.maxima_get(which))
        y = from_state(grid, mb.backend.maxima_get(which))
        assert np.array_equal(x, y, equal_nan=True), which
    assert np.array_equal(ma.backend.tally_totals(), mb.backend.tally_totals())
    assert np.array_equal(np.sort(ma.backend.active_tiles), np.sort(mb.backend.active_tiles))


@pytest.mark.gpu
@pytest.mark.parametrize("backend,maxima", [("hip-exact", 1), ("hip-exact", 4), ("hip-f64", 4), ("hip", 4)])
def test_pipelined_run_is_the_two_kernel_run_bit_for_bit(backend, maxima):
    """A rained valley draining through a free outflow: tiles wake and sleep, windows open every 30 s, the run ends
    with its member retired; chunks in graphs and one launch at a time."""
    chunks = [(100, 16), (37, 0), (2000, 64), (5, 0), (400, 16)]
    a, sa = _valley_run(backend, True, maxima, 2400.0, chunks)
    b, sb = _valley_run(backend, False, maxima, 2400.0, chunks)
    print(f"{backend} maxima {maxima}: {sa.nsteps} steps, pipelined {sa.pipelined}/{sb.pipelined}", flush=True)
    assert sa.pipelined == 1 and sb.pipelined == 0
    assert sa.done and sa.nsteps > 300
    _same_state(sa, sb)
    _same_model(a, b, a.grid)


@pytest.mark.gpu
def test_pipelined_run_read_mid_run_is_the_two_kernel_state():
    """run_steps returns with no update waiting: a read between chunks (fields, maxima, tally, clock) sees the state
    the two-kernel run has after the same steps."""
    for n in (1, 2, 3, 17, 64, 200):
        a, sa = _valley_run("hip-exact", True, 1, 2400.0, [(n, 0 if n < 16 else 16)])
        b, sb = _valley_run("hip-exact", False, 1, 2400.0, [(n, 0 if n < 16 else 16)])
        _same_state(sa, sb)
        _same_model(a, b, a.grid)


@pytest.mark.gpu
@pytest.mark.parametrize("backend", ["hip-exact", "hip-f64"])
def test_pipelined_ensemble_with_a_held_member_is_the_two_kernel_one(backend):
    """Two members with SCS storms in one mosaic; one is held at t = 600 s for an odd number of the other's steps,
    then both run to their ends."""
    specs = [MemberSpec(0, 21, 96, 90.0, 0.5, 0.9), MemberSpec(1, 22, 64, 80.0, 0.5, 1.0)]
    ens = Ensemble.build(specs, dx=20.0, k=4, duration_s=1500.0)
    out = []
    for pipeline in (True, False):
        bt = setup_batched(ens, backend, maxima_every=4, pipeline=pipeline)
        bt.model.backend.run_pause([600.0, np.inf])
        s = bt.run.steps(40, 16)
        while s[0].t < 600.0:
            s = bt.run.steps(7, 0)
        for chunk, graph in ((7, 0), (32, 16), (6, 0)):
            s = bt.run.steps(chunk, graph)
        bt.model.backend.run_pause([np.inf, np.inf])
        s = bt.run.run(chunk=64, graph_steps=16)
        out.append((bt, s))
    (a, sa), (b, sb) = out
    assert [x.pipelined for x in sa] == [1, 1] and [x.pipelined for x in sb] == [0, 0]
    for m in range(2):
        _same_state(sa[m], sb[m])
    _same_model(a.model, b.model, ens.tables.grid)


@pytest.mark.gpu
@pytest.mark.parametrize("edge", [False, True])
def test_following_cells_keep_the_pipeline_only_where_each_tile_updates_their_reference(edge):
    """A column of following boundary cells (kind 2) inside the domain: in the middle of a tile, each following its
    west neighbour (every tile that updates the cell updates its reference: the pipelined step applies), or at a
    tile's east column following its east neighbour (the next tile's west column, which the cell's own tile does not
    update: the member takes the two-kernel step). Either way the device loop is the host loop bit for bit
    (cells.cu:95-111 after every update)."""
    import dataclasses

    from wadi2d_amd.solver.backend import BoundaryCells
    from wadi2d_amd.solver.state import state_index

    t_end = 1800.0
    runs = []
    steps = 0
    for loop in ("host", "device"):
        tables, bset, rain, m = _valley_setup("hip-exact")
        ny, _ = tables.grid.shape
        j = np.arange(2, ny - 2)
        i = np.full(j.size, 63 if edge else 47)  # column 31 or 15 of the second tile column
        cell = state_index(tables.grid, j, i).astype(np.int32)
        ref = state_index(tables.grid, j, i + (1 if edge else -1)).astype(np.int32)
        extra = BoundaryCells(cell, ref, np.full(j.size, 2, np.int32), np.full(j.size, -0.05, np.float32),
                              np.zeros(j.size, np.float32))  # fmt: skip
        c = bset.cells
        both = BoundaryCells(*(np.concatenate((getattr(c, f), getattr(extra, f))) for f in
                               ("cell", "ref", "kind", "d0", "h0")))  # fmt: skip
        mask = bset.mask.copy()
        mask[j, i] = True
        bset2 = dataclasses.replace(bset, cells=both, mask=mask)
        drv = Driver(m, bset2, rain, maxima_every=1)
        if loop == "host":
            drv.run(t_end)
            steps = m.stepper.nsteps
            runs.append((m, None))
        else:  # exactly the host's steps (one more would retire the member)
            run = DeviceRun(m.backend, [_device_plan(m, bset2, rain, t_end)], m.num, maxima=1, idle=m.stepper.idle)
            s = run.steps(steps, 16)[0]
            runs.append((m, s))
    (host, _), (dev, s) = runs
    print(f"edge {edge}: pipelined {s.pipelined}, {s.nsteps}/{host.stepper.nsteps} steps", flush=True)
    assert s.pipelined == (0 if edge else 1)
    assert s.nsteps == host.stepper.nsteps and s.t == host.stepper.t
    for f in FIELDS:
        assert np.array_equal(dev.field(f), host.field(f)), f
    for which in range(5):
        x = from_state(host.grid, host.backend.maxima_get(which))
        y = from_state(dev.grid, dev.backend.maxima_get(which))
        assert np.array_equal(x, y, equal_nan=True), which


def _wet_valley(backend, storm_s):
    """The valley wet all over (0.3 m mean depth) under a storm on every cell for storm_s seconds."""
    from wadi2d_amd.engine.forcing import Rain, uniform_storm

    tables, bset, _, m = _valley_setup(backend)
    rain = Rain(tables, bset, uniform_storm(40.0 * storm_s / 1800.0, storm_s), np.ones(tables.grid.shape))
    m.set_depth(np.full(tables.grid.shape, 0.3))
    drv = Driver(m, bset, rain, maxima_every=1)
    return drv, bset, rain, m


def _wet_valley_run(backend, pipeline, t_end, chunks, storm_s):
    _, bset, rain, m = _wet_valley(backend, storm_s)
    run = DeviceRun(m.backend, [_device_plan(m, bset, rain, t_end)], m.num, maxima=1, idle=m.stepper.idle,
                    pipeline=pipeline)  # fmt: skip
    for n, graph in chunks:
        s = run.steps(n, graph)
    return m, s[0]


@pytest.mark.gpu
@pytest.mark.parametrize("backend,storm_s", [("hip-exact", 1800.0), ("hip", 1800.0), ("hip-exact", 600.0)])
def test_window_ends_and_checks_keep_the_pipeline_when_the_list_cannot_change(backend, storm_s):
    """With every tile listed a window's end cannot change the list, and while it rains on every tile a check cannot
    either (each tile is kept whatever its wet flag says): the step before either leaves its update to the next step's
    pipe kernel like any other, and only each chunk's last step (whose update run_steps makes before it returns) and
    the run's last take another path. A storm that stops at 600 s turns the checks' rule off mid-run (and with it,
    as tiles dry and leave, the windows'). Either way the two-kernel run bit for bit, and the host loop's (whose
    stepper checks the list on the host), to its last step."""
    t_end = 1500.0
    chunks = [(90, 16), (33, 0), (5000, 64)]
    a, sa = _wet_valley_run(backend, True, t_end, chunks, storm_s)
    b, sb = _wet_valley_run(backend, False, t_end, chunks, storm_s)
    n, every = sa.nsteps, a.num.active_every
    print(f"{backend}, storm {storm_s:.0f} s: {n} steps, {sa.pipelined_updates} pipelined updates, checks every "
          f"{every}", flush=True)  # fmt: skip
    assert sa.done and sb.done
    if storm_s >= t_end:
        # a check or a window's end left to the cell kernel would allow at most n - n // every pipelined updates
        assert sa.pipelined_updates >= n - len(chunks) - 2 > n - n // every
    else:  # the checks after the storm (some 15 of them) take the cell kernel
        assert sa.pipelined_updates < n - 5
    _same_state(sa, sb)
    _same_model(a, b, a.grid)
    drv, _, _, host = _wet_valley(backend, storm_s)
    drv.run(t_end)
    c, sc = _wet_valley_run(backend, True, t_end, [(host.stepper.nsteps, 16)], storm_s)
    assert sc.nsteps == host.stepper.nsteps and sc.t == host.stepper.t and sc.dt == host.stepper.dt
    assert c.backend.tally_volume() == host.backend.tally_volume()
    for f in FIELDS:
        assert np.array_equal(c.field(f), host.field(f)), f
    for which in range(5):
        x = from_state(host.grid, host.backend.maxima_get(which))
        y = from_state(c.grid, c.backend.maxima_get(which))
        assert np.array_equal(x, y, equal_nan=True), which
    assert np.array_equal(np.sort(c.backend.active_tiles), np.sort(host.backend.active_tiles))
