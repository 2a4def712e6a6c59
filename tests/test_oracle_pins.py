"""Pins of the oracle's rules that the physics cases alone leave loose, each a check the reference itself holds or one
that the mathematics fixes, so that a plausible mistake in the oracle fails here:

* the mirror rule for a dry neighbour face (``gated``): the reference's 45-degree channel (tests/solver/
  test_dry_faces.py:17-59): with the rule on, the depth in a channel 3 cells wide laid at 45 degrees is within 3 % of
  Manning's normal depth; with it off, 5-12 % too deep;
* the normal (kind 2) boundary (forcing.py:299-332, cells.cu:95-111): steady uniform flow in the reference's 1 %
  channel (tests/solver/test_face_level.py:55-94) holds the normal depth within 1.5 % with the outlet fixed at the
  normal level, as there, and with a normal boundary instead, both at mid-reach and just above the outlet; and still
  water behind a raised rim of normal-boundary cells at slope 0 stays still (to the float32 rounding of d0), which
  holds only if the rim follows the lake's level (d0 = zmin_ref - zmin - S dx enters with its sign);
* sleeping tiles (stepper.py:56-92, ``rebuild_active``): the reference's own invariant (tests/solver/
  test_sleeping.py:19-111): a run whose dry tiles sleep is, to the last bit, the run that steps every tile.

The oracle runs here on the CPU; the same sleeping invariant through the HIP path is the GPU test at the end.
"""

from __future__ import annotations

import numpy as np
import pytest
from scipy.optimize import brentq

from oracle.oracle import Numerics as ONumerics
from oracle.oracle import OracleModel
from wadi2d_amd.config import Tables
from wadi2d_amd.grid import TileGrid
from wadi2d_amd.solver.backend import BoundaryCells
from wadi2d_amd.solver.state import level_of
from wadi2d_amd.tables import ArrayBed, build_tables, hbar_at_level

THREADS = 8


def _cells(model, mask):
    jj, ii = np.nonzero(mask)
    return np.array([model.cell_state_index(int(j), int(i)) for j, i in zip(jj, ii, strict=True)], np.int32)


# ---- the mirror rule: a channel three cells wide at 45 degrees


def _channel_45(mirror: bool) -> float:
    """Depth error against Manning's normal depth, %, of a 60 m channel with vertical walls laid at 45 degrees on a 20 m
    grid (1 m pixels), fed 60 m3/s at its head and held at the normal level at its foot."""
    dx, k, slope, n, b, q, side, offset = 20.0, 20, 0.002, 0.03, 60.0, 60.0, 128, 5.0
    tg = TileGrid.from_mask(np.ones((side, side), bool), 0.0, 0.0, dx)
    ny, nx = tg.shape
    p = dx / k
    c = np.sqrt(0.5)  # cos and sin of 45 degrees
    x = (np.arange(nx * k + 2) - 0.5) * p - 0.5 * nx * dx
    y = (np.arange(ny * k + 2) - 0.5) * p - 0.5 * ny * dx
    px, py = np.meshgrid(x, y)
    along, across = c * (px + py), c * (py - px) - offset
    z = -slope * along + 5.0 * (np.abs(across) >= b / 2)
    normal = brentq(lambda d: b * d * (b * d / (b + 2 * d)) ** (2 / 3) * np.sqrt(slope) / n - q, 0.01, 20.0)
    tables = build_tables(tg, ArrayBed(z, p, 1, 1), Tables(), n)
    model = OracleModel(tables, ONumerics(mirror_dry_faces=mirror), prec=8, threads=THREADS)
    cx = (np.arange(nx) + 0.5) * dx - 0.5 * nx * dx
    cy = (np.arange(ny) + 0.5) * dx - 0.5 * ny * dx
    gx, gy = np.meshgrid(cx, cy)
    s, t = c * (gx + gy), c * (gy - gx) - offset
    level = -slope * s + normal
    hc, dtab = tg.from_tiles(tables.hc).astype(np.float64), tg.from_tiles(tables.dtab).astype(np.float64)
    h0 = np.where(np.abs(t) < b / 2 + 2 * dx, hbar_at_level(level, tables.zmin, hc, dtab), 0.0)
    model.set_depth(h0)
    reach = 0.45 * side * dx
    head = (s > -reach) & (s < -reach + 120.0) & (np.abs(t) < b / 2 + dx)
    model.set_source(np.where(head, q / (head.sum() * dx * dx), 0.0))
    foot = (s > reach - 120.0) & (s < reach) & (np.abs(t) < b + 3 * dx)
    cells = _cells(model, foot)
    zero = np.zeros(cells.size, np.int32)
    d0 = np.maximum(level - tables.zmin, 0.0)[foot].astype(np.float32)
    model.set_boundary(BoundaryCells(cells, zero, zero.copy(), d0, h0[foot].astype(np.float32)))
    model.advance(4000.0)
    middle = (np.abs(s) < 0.15 * side * dx) & (np.abs(t) < 0.71 * dx)
    surface = tables.zmin + model.field("d") + slope * s
    return float(np.mean(surface[middle]) / normal - 1.0) * 100.0


@pytest.mark.parametrize(("mirror", "low", "high"), [(True, -3.0, 3.0), (False, 5.0, 12.0)])
def test_the_mirror_rule_in_the_45_degree_channel(mirror, low, high):
    err = _channel_45(mirror)
    print(f"mirror {mirror}: depth {err:+.2f} % against Manning's normal depth", flush=True)
    assert low < err < high


# ---- the normal boundary: steady uniform flow in the 1 % channel, and still water behind a rim

SIDE, DX, K = 96, 20.0, 10
S, N, WALL = 0.01, 0.03, 5.0
B, BANK, Q = 60.0, 2.0, 300.0


def _normal_depth() -> float:
    """Manning's normal depth of the trapezoid as the sub-grid conveyance sums it (no wall friction)."""
    return brentq(lambda y: (B * y ** (5 / 3) + 2 * BANK * (3 / 8) * y ** (8 / 3)) * S**0.5 / N - Q, 0.01, 20.0)


def _uniform_channel(outlet: str) -> tuple[float, float]:
    """The depth over the normal depth at mid-reach and just above the outlet, after 3000 s of 300 m3/s into the 1 %
    channel; ``outlet`` "fixed" holds the last cells at the normal level, "normal" makes them follow their upstream
    neighbour lowered by S dx (kind 2, its d0 as forcing.py:320-330 forms it)."""
    tg = TileGrid.from_mask(np.ones((SIDE, SIDE), bool), 0.0, 0.0, DX)
    ny, nx = tg.shape
    p = DX / K
    x = tg.x0 + (np.arange(nx * K) + 0.5) * p
    y = tg.y0 - (np.arange(ny * K) + 0.5) * p
    px, py = np.meshgrid(x, y)
    xm, ym = 0.5 * SIDE * DX, -0.5 * SIDE * DX
    z = (100.0 - (px - xm) * S + np.clip((np.abs(py - ym) - B / 2) / BANK, 0.0, WALL)).astype(np.float32)
    tables = build_tables(tg, ArrayBed(z.astype(np.float64), p), Tables(), N)
    jj, ii = np.mgrid[0:ny, 0:nx]
    gx, gy = tg.xy_of(jj, ii)
    s, t = gx - xm, gy - ym
    yn = _normal_depth()
    hc, dtab = tg.from_tiles(tables.hc).astype(np.float64), tg.from_tiles(tables.dtab).astype(np.float64)
    level = 100.0 - s * S + yn
    hb = np.where((np.abs(t) < B / 2 + BANK * yn + 2 * DX) & tg.mask, hbar_at_level(level, tables.zmin, hc, dtab), 0.0)
    model = OracleModel(tables, ONumerics(), prec=8, threads=THREADS)
    model.set_depth(hb)
    lim = 0.45 * SIDE * DX
    up = (s > -lim) & (s < -lim + 6 * DX) & (np.abs(t) < B / 2 + DX) & tg.mask
    model.set_source(np.where(up, Q / (up.sum() * DX * DX), 0.0))
    if outlet == "fixed":
        down = (s > lim - 6 * DX) & (s < lim) & (np.abs(t) < B + 3 * DX) & tg.mask
        cells = _cells(model, down)
        zero = np.zeros(cells.size, np.int32)
        d0 = level_of(hb[down], hc[down], dtab[down]).astype(np.float32)
        model.set_boundary(BoundaryCells(cells, zero, zero.copy(), d0, hb[down].astype(np.float32)))
    else:
        # one column of normal-boundary cells across the channel; each follows its upstream (west) neighbour, the
        # untaken neighbour with the highest mean bed
        col = int(np.argmin(np.abs(s[0] - (lim - 0.5 * DX))))
        rows = np.nonzero((np.abs(t[:, col]) < B + 3 * DX) & tg.mask[:, col])[0]
        cells = np.array([model.cell_state_index(int(j), col) for j in rows], np.int32)
        refs = np.array([model.cell_state_index(int(j), col - 1) for j in rows], np.int32)
        d0 = np.minimum(tables.zmin[rows, col - 1] - tables.zmin[rows, col] - S * DX, 0.0).astype(np.float32)
        model.set_boundary(BoundaryCells(cells, refs, np.full(cells.size, 2, np.int32), d0, np.zeros_like(d0)))
    model.advance(3000.0)
    d = model.field("d").astype(np.float64)
    depth = tables.zmin + d + s * S - 100.0
    mid = (np.abs(s) < 0.15 * SIDE * DX) & (np.abs(t) < 0.71 * DX) & tg.mask
    near = (s > lim - 8 * DX) & (s < lim - 3 * DX) & (np.abs(t) < 0.71 * DX) & tg.mask
    return float(np.mean(depth[mid])) / yn, float(np.mean(depth[near])) / yn


def test_steady_uniform_flow_holds_the_normal_depth_with_the_outlet_fixed():
    mid, _ = _uniform_channel("fixed")
    print(f"fixed outlet: depth {mid:.4f} of the normal depth at mid-reach", flush=True)
    assert abs(mid - 1.0) < 0.015


def test_a_normal_boundary_holds_the_normal_depth_up_to_the_outlet():
    mid, near = _uniform_channel("normal")
    print(f"normal outlet: depth {mid:.4f} of the normal depth at mid-reach, {near:.4f} just above it", flush=True)
    assert abs(mid - 1.0) < 0.015
    assert abs(near - 1.0) < 0.03


def _rim_lake(sign: float = 1.0) -> tuple[float, float]:
    """A flat lake 0.5 m deep on a 10 m grid whose east column is a rim 0.3 m higher, every rim cell a normal boundary
    at slope 0 following its west neighbour: d0 = zmin_ref - zmin - 0 dx = -0.3 m. Returns the largest speed and the
    relative change of the lake's volume after 600 s (the kernel's d0 taken times ``sign``)."""
    side, dx, k = 60, 10.0, 4
    tg = TileGrid.from_mask(np.ones((side, side), bool), 0.0, 0.0, dx)
    ny, nx = tg.shape
    rim = int(np.nonzero(tg.mask.any(axis=0))[0][-1])  # the last column of the domain
    z = np.full((ny * k, nx * k), 50.0)
    z[:, rim * k :] += 0.3
    tables = build_tables(tg, ArrayBed(z, dx / k), Tables(rows=8), 0.03)
    model = OracleModel(tables, ONumerics(), prec=8, threads=THREADS)
    hc, dtab = tg.from_tiles(tables.hc).astype(np.float64), tg.from_tiles(tables.dtab).astype(np.float64)
    level = np.full(tg.shape, 50.5)
    h0 = np.where(tg.mask, hbar_at_level(level, tables.zmin, hc, dtab), 0.0)
    model.set_depth(h0)
    rows = np.nonzero(tg.mask[:, rim])[0]
    cells = np.array([model.cell_state_index(int(j), rim) for j in rows], np.int32)
    refs = np.array([model.cell_state_index(int(j), rim - 1) for j in rows], np.int32)
    d0 = np.minimum(tables.zmin[rows, rim - 1] - tables.zmin[rows, rim], 0.0)
    assert rows.size == side and np.allclose(d0, -0.3)
    model.set_boundary(BoundaryCells(cells, refs, np.full(cells.size, 2, np.int32), (sign * d0).astype(np.float32),
                                     np.zeros(cells.size, np.float32)))  # fmt: skip
    lake = tg.mask.copy()
    lake[:, rim:] = False
    v0 = float(model.field("h")[lake].sum())
    model.advance(600.0)
    speed = float(np.abs(model.field("u")).max() + np.abs(model.field("v")).max())
    return speed, float(model.field("h")[lake].sum()) / v0 - 1.0


def test_still_water_behind_a_rim_of_normal_boundary_cells_stays_still():
    """Still, to the float32 rounding of d0 (the rim then stands about 1e-8 m off the lake's level)."""
    speed, change = _rim_lake()
    print(f"rim lake: largest speed {speed:.2e} m/s, volume change {change:.2e}", flush=True)
    assert speed < 1e-6 and abs(change) < 1e-7


def test_the_rim_check_catches_a_wrong_sign_of_d0():
    """The same lake with d0 entering with the wrong sign: the rim stands 0.6 m over the lake and water pours in."""
    speed, change = _rim_lake(sign=-1.0)
    print(f"rim lake, d0 of the wrong sign: largest speed {speed:.2e} m/s, volume change {change:.2e}", flush=True)
    assert speed > 0.05 and change > 1e-3


# ---- sleeping tiles: the run equals the run that steps every tile


VDX, VK = 20.0, 2
VSLOPE, VSIDE, VMANNING = 0.01, 0.02, 0.035


def _valley(tiles_y: int, tiles_x: int, sleeping: bool, prec: int) -> OracleModel:
    """A long valley falling east, hills rising either side of its axis (the reference's sleeping-tile ground)."""
    ny, nx = tiles_y * 32 - 2, tiles_x * 32 - 2
    tg = TileGrid.from_mask(np.ones((ny, nx), bool), 0.0, 0.0, VDX)
    axis = tg.shape[0] // 2
    p = VDX / VK
    xs = (np.arange(tg.shape[1] * VK) + 0.5) * p
    ys = (np.arange(tg.shape[0] * VK) + 0.5) * p
    px, py = np.meshgrid(xs, ys)
    z = (200.0 - VSLOPE * px + VSIDE * np.abs(py - (axis + 0.5) * VDX)).astype(np.float32).astype(np.float64)
    tables = build_tables(tg, ArrayBed(z, p), Tables(), VMANNING)
    m = OracleModel(tables, ONumerics(), prec=prec, threads=THREADS)
    m.stepper.adaptive_active = sleeping
    m.axis = axis
    m.set_depth(np.zeros(tg.shape))
    return m


def _same(a: OracleModel, b: OracleModel, where: str) -> None:
    sa, sb = a.stepper, b.stepper
    assert (sa.nsteps, sa.t, sa.dt) == (sb.nsteps, sb.t, sb.dt), where
    for f in ("h", "d", "u", "v", "mx", "my", "qx", "qy", "res"):
        x, y = a.field(f), b.field(f)
        assert np.array_equal(x, y), f"{where}: {f} differs in {int(np.sum(x != y))} cells"


def _run_same(a, b, times) -> list[int]:
    awake = []
    for t in times:
        a.advance(t)
        b.advance(t)
        _same(a, b, f"t = {t:g} s")
        awake.append(int(a.backend.nact))
    assert b.backend.nact == b.grid.ntiles - 1  # the reference run never sleeps
    return awake


def _head(m, q: float, column: int = 8) -> np.ndarray:
    rate = np.zeros(m.grid.shape)
    rate[[m.axis - 2, m.axis + 2], column] = q / (2 * VDX * VDX)
    return rate


@pytest.mark.parametrize("prec", [4, 8])
def test_a_front_crossing_tile_borders_in_every_direction_sleeps_as_the_all_tiles_run(prec):
    a, b = (_valley(5, 5, sleeping, prec) for sleeping in (True, False))
    for m in (a, b):
        h = np.zeros(m.grid.shape)
        h[8:24, 70:90] = 2.0  # on the north hillside, running south across two rows of tiles
        h[136:152, 70:90] = 2.0  # on the south hillside, running north
        h[72:88, 70:90] = 3.0  # on the axis: it backs up west and runs east
        m.set_depth(h * m.grid.mask)
    assert a.backend.nact == 15  # the three wet tiles and their neighbours, of 25
    awake = _run_same(a, b, [20.0, 60.0, 150.0, 400.0])
    assert awake[-1] > 15


def test_rain_starting_later_on_a_sleeping_tile_wakes_it_and_the_runs_stay_equal():
    a, b = (_valley(2, 8, sleeping, 4) for sleeping in (True, False))
    for m in (a, b):
        m.set_source(_head(m, 30.0))
    _run_same(a, b, [200.0])
    assert a.backend.nact < 16
    far = a.cell_state_index(10, 210) // 1296
    assert far not in set(a.backend.active_tiles.tolist())
    for m in (a, b):
        rate = _head(m, 30.0)
        rate[5:25, 200:230] = 60.0 / 3.6e6  # 60 mm/h on a dry, sleeping hillside six tiles down the valley
        m.set_source(rate * m.grid.mask)
    assert far in set(a.backend.active_tiles.tolist())
    _run_same(a, b, [260.0, 500.0])
    assert a.field("h")[5:25, 200:230].min() > 0.0


def test_a_tile_that_dries_sleeps_again_and_wakes_again():
    a, b = (_valley(2, 8, sleeping, 4) for sleeping in (True, False))

    def source(mm_h: float, everywhere: bool = False) -> None:
        for m in (a, b):
            rate = np.zeros(m.grid.shape)
            rate[8:20, 200:215] = mm_h / 3.6e6
            if everywhere:
                rate[:] = mm_h / 3.6e6
            m.set_source(rate * m.grid.mask)

    source(40.0)
    assert a.backend.nact == 6  # the rained tile and its neighbours
    _run_same(a, b, [120.0])
    source(-4000.0, everywhere=True)  # a loss that takes every drop wherever it ran
    _run_same(a, b, [600.0])
    assert a.volume() == 0.0
    source(0.0)
    assert a.backend.nact == 16  # a change of source drops no tile: only a check does
    _run_same(a, b, [700.0, 1200.0])
    assert a.backend.nact == 0  # dry at two checks in a row: every tile left the list
    source(40.0)
    assert a.backend.nact == 6
    _run_same(a, b, [1300.0, 1800.0])
    assert a.volume() > 0.0


@pytest.mark.gpu
@pytest.mark.parametrize("backend", ["hip-exact", "hip"])
def test_the_hip_path_sleeps_and_gives_the_all_tiles_run(backend):
    """The same invariant through the C-ABI: an inflow at the head of the valley and an open end, tiles waking as the
    flood runs down it (tests/solver/test_sleeping.py:111-116)."""
    from wadi2d_amd.solver.model import Model

    def valley(sleeping):
        ref = _valley(2, 8, True, 4)  # for the ground and tables
        m = Model(ref.tables, None, backend)
        m.stepper.adaptive_active = sleeping
        m.axis = ref.axis
        m.set_depth(np.zeros(m.grid.shape))
        cols = np.nonzero(m.grid.mask.any(axis=0))[0]
        j = np.nonzero(m.grid.mask[:, cols[-1]])[0]
        cells = np.array([m.cell_state_index(int(r), int(cols[-1])) for r in j], np.int32)
        zero = np.zeros(cells.size, np.int32)
        f = np.zeros(cells.size, np.float32)
        m.set_boundary(BoundaryCells(cells, zero, np.ones(cells.size, np.int32), f, f.copy()))
        m.set_source(_head(m, 40.0))
        return m

    a, b = valley(True), valley(False)
    awake = []
    for t in (30.0, 300.0, 1200.0, 3600.0):
        a.advance(t)
        b.advance(t)
        assert (a.stepper.nsteps, a.stepper.t, a.stepper.dt) == (b.stepper.nsteps, b.stepper.t, b.stepper.dt)
        for f in ("h", "d", "u", "v", "mx", "my", "qx", "qy", "res"):
            assert np.array_equal(a.field(f), b.field(f)), (t, f)
        awake.append(int(a.backend.nact))
    print(f"{backend}: awake tiles {awake} of {a.grid.ntiles - 1}", flush=True)
    assert min(awake) <= 8 and awake[0] < awake[-1]
    assert b.backend.nact == b.grid.ntiles - 1


# ---- the running maxima, against their definitions written on the (ny, nx) grid


def test_the_oracle_maxima_are_their_definitions_on_the_grid():
    """engine/maxima.py:29-51 written a second way: on whole (ny, nx) fields, the speed at a cell centre from the mean
    of its west and east face velocities (the east face is the next column's west face, zero past the grid) and of
    its north and south ones; a planted mistake (no half in the centre speed) differs."""
    from oracle.oracle import Maxima, from_state
    from tests.cases import rough_dam_break

    tables, h0 = rough_dam_break()
    m = OracleModel(tables, ONumerics.like(ONumerics()), prec=8, threads=THREADS)
    m.set_depth(h0)
    mx = Maxima(m, arrive=0.1, deep=0.05)
    g = tables.grid
    want = {"depth": np.zeros(g.shape), "mean_depth": np.zeros(g.shape), "speed": np.zeros(g.shape),
            "hazard": np.zeros(g.shape), "arrival": np.full(g.shape, -1.0), "planted": np.zeros(g.shape)}  # fmt: skip

    def sample(st):
        mx.update(st.t)
        d, h, u, v = (m.field(f) for f in ("d", "h", "u", "v"))
        east = np.concatenate([u[:, 1:], np.zeros((u.shape[0], 1))], axis=1)
        south = np.concatenate([v[1:, :], np.zeros((1, v.shape[1]))], axis=0)
        sp = np.hypot(0.5 * (u + east), 0.5 * (v + south))
        want["depth"] = np.maximum(want["depth"], d)
        want["mean_depth"] = np.maximum(want["mean_depth"], h)
        deep = d > 0.05
        want["speed"] = np.where(deep, np.maximum(want["speed"], sp), want["speed"])
        want["hazard"] = np.where(deep, np.maximum(want["hazard"], sp * d), want["hazard"])
        want["arrival"] = np.where((want["arrival"] < 0) & (d >= 0.1), st.t, want["arrival"])
        west_only = np.hypot(u, v)  # the planted mistake: the west and north faces' speed, no centre mean
        want["planted"] = np.where(deep, np.maximum(want["planted"], west_only), want["planted"])

    m.advance(20.0, sample)
    mask = g.mask
    for name, arr in zip(Maxima.NAMES, mx.arrays, strict=True):
        got = from_state(g, arr)
        assert np.allclose(got[mask], want[name][mask], rtol=1e-12, atol=1e-12), name
    assert (want["speed"][mask] > 0.5).any() and (want["arrival"][mask] > 0).any()
    assert not np.allclose(from_state(g, mx.arrays[2])[mask], want["planted"][mask], rtol=1e-3)
