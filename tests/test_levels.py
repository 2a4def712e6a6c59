"""Local time stepping (wadi2d/solver/level_stepper.py, solver/levels.py, solver/cpu/levels.py, levels.cu).

On the CPU: the package's vectorised levels and cycle plans against the oracle's loop restatement on random tile
sets, the grading's defining property checked by brute force, and the oracle's level stepper pinned by what the
method fixes: in float64 a closed basin keeps its volume to rounding whatever the levels (the edge corrections are
what make that so; the test plants their loss in the oracle and sees the volume drift), tiles do sit at several
levels, and the result stays near the uniform step's. On the GPU: the HIP level stepper against the oracle's, bit
for bit in exact float32 and float64, over cycles in which the levels change, and the fast variant within the
reference's parity limits.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import levels as olv
from oracle.oracle import Numerics as ONumerics
from tests.cases import deep_and_shallow
from wadi2d_amd.config import Numerics
from wadi2d_amd.grid import TileGrid
from wadi2d_amd.solver import levels as plv

LTS = dict(local_time_stepping=True, lts_levels=3)


def _random_tiles(rng, ny=9, nx=11):
    mask = np.zeros((ny * 32, nx * 32), bool)
    held = rng.random((ny, nx)) < 0.8
    for j, i in zip(*np.nonzero(held)):
        mask[j * 32 : (j + 1) * 32, i * 32 : (i + 1) * 32] = True
    g = TileGrid(0.0, 0.0, 10.0, mask)
    tiles = np.arange(1, g.ntiles)
    active = np.sort(rng.choice(tiles, size=max(1, int(0.7 * tiles.size)), replace=False)).astype(np.int32)
    return g, active


def test_grading_equals_the_oracle_and_its_definition():
    rng = np.random.default_rng(1)
    for _ in range(25):
        g, active = _random_tiles(rng)
        level = np.zeros(g.ntiles, np.int32)
        level[active] = rng.integers(0, 5, active.size)
        got = plv.graded(level, active, g.nbr)
        assert np.array_equal(got, olv.graded(level, active, g.nbr))
        on = np.zeros(g.ntiles, bool)
        on[active] = True
        # the definition: at or under the asked level, active neighbours at most one apart, and as high as that allows
        assert np.all(got[active] <= level[active])
        for t in active:
            for o in g.nbr[t]:
                if on[o]:
                    assert abs(int(got[t]) - int(got[o])) <= 1
        for t in active:  # raising any one tile by one breaks the rule or exceeds its asked level
            if got[t] < level[t]:
                assert any(on[o] and got[t] + 1 - got[o] > 1 for o in g.nbr[t])
        assert np.all(got[~on] == 0)


def test_levels_and_cycle_plans_equal_the_oracle():
    rng = np.random.default_rng(2)
    for trial in range(25):
        g, active = _random_tiles(rng)
        wave = rng.uniform(0.0, 9.0, g.ntiles) * (rng.random(g.ntiles) < 0.9)
        pinned = np.zeros(g.ntiles, bool)
        pinned[rng.choice(active, 2)] = True
        previous = rng.integers(0, 4, g.ntiles).astype(np.int32) if trial % 2 else None
        climb = trial % 3 == 0
        rule = plv.LevelRule(0.45, 1.0, 1.0, 5.0, 3)
        lv, dt = plv.assign_levels(wave, active, g.nbr, g.dx, rule, pinned, previous, climb)
        lv_o, dt_o = olv.assign_levels(wave, active, g.nbr, g.dx, 0.45, 1.0, 1.0, 5.0, 3, pinned, previous, climb)
        assert np.array_equal(lv, lv_o) and dt == dt_o
        p = plv.plan_cycle(lv, active, g.nbr)
        q = olv.Plan(lv, active, g.nbr)
        assert p.top == q.top
        for name in ("groups", "upto", "halo", "coarse", "edges_x", "edges_y", "strips"):
            for a, b in zip(getattr(p, name), getattr(q, name), strict=True):
                assert np.array_equal(a, b.reshape(a.shape)), name


def test_every_level_edge_is_listed_once_and_only_between_neighbouring_levels():
    rng = np.random.default_rng(3)
    g, active = _random_tiles(rng)
    level = np.zeros(g.ntiles, np.int32)
    level[active] = rng.integers(0, 4, active.size)
    level = plv.graded(level, active, g.nbr)
    p = plv.plan_cycle(level, active, g.nbr)
    rows = np.concatenate([np.concatenate(p.edges_x + p.edges_y)]).reshape(-1, 4)
    assert np.array_equal(np.sort(rows[:, 3]), np.arange(rows.shape[0]))
    on = np.zeros(g.ntiles, bool)
    on[active] = True
    want = set()
    for t in active:
        for side, col in enumerate(plv.SIDE_COLUMNS):
            o = g.nbr[t, col]
            if on[o] and level[o] != level[t]:
                fine, crs = (t, o) if level[t] < level[o] else (o, t)
                if fine == t:
                    want.add((int(t), side, int(o)))
                assert abs(int(level[o]) - int(level[t])) == 1
    assert {(int(a), int(b), int(c)) for a, b, c, _ in rows} == want
    for n in range(p.top + 1):  # strips: the coarse tile east or south owns the faces
        assert np.all(p.strips[n][:, 1] >= 2)


def _held(m) -> float:
    """Water held: the mean depths and the residuals (a coarse cell's residual carries what the fine steps beside it
    owe it until its next step, and an overdrawn cell's debt), times the cell area."""
    return float((m.field("h") + m.field("res")).sum()) * m.grid.dx**2


def _oracle(num, prec=8):
    tables, h0 = deep_and_shallow()
    m = olv.OracleLevelModel(tables, ONumerics.like(num), prec=prec, threads=4)
    m.set_depth(h0)
    return m


def test_oracle_level_stepper_keeps_a_closed_basin_volume_across_levels():
    num = Numerics(**LTS)
    m = _oracle(num)
    v0 = _held(m)
    tops = set()
    for t_end in (20.0, 60.0, 120.0):
        m.advance(t_end, on_step=lambda st: tops.add(int(st.level.max())))
        assert abs(_held(m) - v0) <= 1e-13 * v0, (t_end, _held(m) - v0)
    assert max(tops) >= 2, tops  # tiles did step at several rates
    assert m.stepper.tile_steps < m.stepper.nsteps * m.backend.nact  # and fewer tile-steps than a uniform run


def test_the_volume_drifts_when_the_edge_corrections_are_left_out():
    """A planted mistake in the oracle: without the fine steps' corrections across level edges the basin loses or
    gains water, so the exact-volume pin above would see a dropped correction."""
    num = Numerics(**LTS)
    m = _oracle(num)
    v0 = _held(m)
    m.stepper.ops.flux = lambda rows, dt: None
    m.advance(120.0)
    assert abs(_held(m) - v0) > 1e-9 * v0


def test_oracle_level_stepper_stays_near_the_uniform_step():
    from oracle.oracle import OracleModel

    tables, h0 = deep_and_shallow()
    lts = _oracle(Numerics(**LTS))
    uni = OracleModel(tables, ONumerics.like(Numerics()), prec=8, threads=4)
    uni.set_depth(h0)
    lts.advance(120.0)
    uni.advance(120.0)
    dh = np.abs(lts.field("h") - uni.field("h"))
    assert float(np.percentile(dh[tables.grid.mask], 99)) < 0.05
    assert float(dh.max()) < 0.5


# ---- the HIP level stepper against the oracle's


def _hip_pair(backend, prec, num, outflow=False):
    from wadi2d_amd.solver.backend import BoundaryCells
    from wadi2d_amd.solver.model import Model

    tables, h0 = deep_and_shallow()
    hip = Model(tables, num, backend)
    ora = olv.OracleLevelModel(tables, ONumerics.like(num), prec=prec, threads=4)
    for m in (hip, ora):
        m.set_depth(h0)
        if outflow:  # free cells on the east edge: their tiles are pinned to the smallest step
            ny, nx = tables.grid.shape
            cells = np.array([m.cell_state_index(j, nx - 2) for j in range(1, ny - 1)], np.int32)
            n = cells.size
            m.set_boundary(BoundaryCells(cells, cells.copy(), np.ones(n, np.int32), np.zeros(n, np.float32),
                                         np.zeros(n, np.float32)))  # fmt: skip
    return hip, ora


FIELDS = ("h", "d", "u", "v", "mx", "my", "qx", "qy", "res", "om")


@pytest.mark.gpu
@pytest.mark.parametrize("backend,prec,outflow", [("hip-exact", 4, False), ("hip-f64", 8, False),
                                                  ("hip-exact", 4, True)])  # fmt: skip
def test_hip_level_stepper_is_the_oracle_bit_for_bit(backend, prec, outflow):
    from wadi2d_amd.solver.level_stepper import LevelStepper

    num = Numerics(**LTS)
    hip, ora = _hip_pair(backend, prec, num, outflow)
    assert isinstance(hip.stepper, LevelStepper)
    tops = set()
    for t_end in (1.0, 30.0, 90.0):
        hip.advance(t_end, on_step=lambda st: tops.add(int(st.level.max())))
        ora.advance(t_end)
        print(f"{backend} outflow={outflow} t={t_end}: {hip.stepper.nsteps} smallest steps, levels up to "
              f"{max(tops)}", flush=True)
        assert hip.stepper.nsteps == ora.stepper.nsteps and hip.stepper.t == ora.stepper.t
        assert np.array_equal(hip.stepper.level, ora.stepper.level)
        assert np.array_equal(hip.backend.active_tiles, ora.backend.active_tiles)
        for f in FIELDS:
            a, b = hip.field(f), ora.field(f)
            if prec == 8:
                # the device's cube root and the host's may differ in the last bit of double; the residual is the
                # rounding left over from a depth update, so it is measured against the depth
                scale = max(float(np.abs(ora.field("h") if f == "res" else b).max()), 1e-30)
                assert float(np.abs(a - b).max()) <= 1e-9 * scale, f
            else:
                assert np.array_equal(a, b), f
    assert max(tops) >= (1 if outflow else 2)  # pinned east tiles cap the grading beside them
    if outflow:
        assert hip.boundary_tiles.size and not hip.stepper.level[hip.boundary_tiles].any()


@pytest.mark.gpu
def test_fast_hip_level_stepper_meets_the_fp64_oracle():
    num = Numerics(**LTS)
    hip, ora = _hip_pair("hip", 8, num)
    v0 = _held(ora)
    hip.advance(90.0)
    ora.advance(90.0)
    dh = np.abs(hip.field("h") - ora.field("h"))
    print(f"fast LTS: |dh| max {dh.max():.2e}, volume error {(_held(hip) - v0) / v0:.1e}", flush=True)
    assert float(dh.max()) < 1e-3
    assert abs(_held(hip) - v0) <= 1e-6 * v0


@pytest.mark.gpu
def test_level_stepper_in_the_run_loop_with_rain_and_an_outflow_is_the_oracle():
    """The plain CPU case (rain, a free outflow edge whose tiles are pinned to the smallest step) through the run
    loop: the HIP level stepper driven by the Driver, the oracle's by the same windows; exact float32, bit for bit.
    The Driver's tally counts every smallest step (substep_hook), and the books close."""
    from wadi2d_amd import synth
    from wadi2d_amd.engine.driver import Driver
    from wadi2d_amd.engine.forcing import Rain, boundary_set
    from wadi2d_amd.solver.model import Model

    scen = synth.plain_cpu_case()
    scen.numerics = scen.numerics.model_copy(update=LTS)
    tables = scen.build()
    bset = boundary_set(scen.boundaries, tables)
    rain = Rain(tables, bset, scen.cumulative_mm, scen.weight)
    end = 2400.0
    hip = Model(tables, scen.numerics, "hip-exact")
    hip.set_depth(np.zeros(tables.grid.shape))
    drv = Driver(hip, bset, rain)
    drv.run(end)
    bal = drv.accounts()
    ora = olv.OracleLevelModel(tables, ONumerics.like(scen.numerics), prec=4, threads=8)
    ora.set_depth(np.zeros(tables.grid.shape))
    ora.set_boundary(bset.cells)
    t, dry = 0.0, False
    while t < end - 1e-9:
        t_next = min(t + 30.0, end)
        rate, vol = rain.rate(t, t_next)
        if vol > 0.0 or not dry:
            ora.set_source_cells(rain.state, rate)
        dry = vol <= 0.0
        ora.advance(t_next)
        t = t_next
    print(f"run loop LTS: {hip.stepper.nsteps} smallest steps, top level {int(hip.stepper.level.max())}, "
          f"tile-step share {hip.stepper.tile_steps / max(hip.stepper.nsteps * hip.backend.nact, 1):.2f}, "
          f"mass error {bal.relative:.1e}", flush=True)
    assert hip.stepper.nsteps == ora.stepper.nsteps
    assert bool(hip.stepper.pinned[hip.boundary_tiles].all()) and hip.boundary_tiles.size
    for f in FIELDS:
        assert np.array_equal(hip.field(f), ora.field(f)), f
    assert abs(bal.relative) <= 1e-6
