"""The oracle against what the mathematics fixes, on the CPU: closed forms and invariants that a dropped term, a
wrong sign or a transposed operand in the restated step would break.

Each case restates one of the reference's own solver tests (tests/solver/*.py) and holds the oracle to the same
limit: Stoker's dam break (test_analytic.py:17-26), still water on rough ground at altitude (:52-68), rain on deep
water stored in full (test_review_regressions.py:26-38), critical flow over a thin crest (test_subgrid.py:65-92), the
kinematic wave on a plane (:49-62), a NaN stopping the run (test_review_regressions.py:96-105), the first step on dry
ground (:66-69) and an empty grid (:89-93).
"""

from __future__ import annotations

import math

import numpy as np
import pytest
from scipy.ndimage import gaussian_filter
from scipy.optimize import brentq

from oracle.oracle import NumericalFailure, Numerics, OracleModel
from wadi2d_amd.config import CRITICAL_WEIR_COEFF, G, Tables
from wadi2d_amd.grid import TileGrid
from wadi2d_amd.solver.backend import BoundaryCells
from wadi2d_amd.tables import ArrayBed, build_tables, hbar_at_level

PLAIN = Numerics(sheet_flow=False, e_wet=1e-5, h_cfl=1e-4, dt_every=1, dt_max=60.0)


def _flat_model(n, dx, prec=8, num=None, k=1, manning=0.03):
    grid = TileGrid.from_mask(np.ones((n, n), bool), 0.0, 0.0, dx)
    z = np.zeros((grid.shape[0] * k, grid.shape[1] * k))
    return OracleModel(build_tables(grid, ArrayBed(z, dx / k), Tables(), manning), num or Numerics(), prec=prec,
                       threads=4), grid  # fmt: skip


def _stoker(h_left, h_right):
    """The plateau between rarefaction and shock: the velocity behind the rarefaction equals the one behind the
    shock (Stoker 1957)."""
    c_left = math.sqrt(G * h_left)

    def mismatch(h):
        return 2.0 * (c_left - math.sqrt(G * h)) - (h - h_right) * math.sqrt(G * (h + h_right) / (2.0 * h * h_right))

    hm = brentq(mismatch, h_right, h_left, xtol=1e-13)
    um = 2.0 * (c_left - math.sqrt(G * hm))
    return hm, um, math.sqrt(G * hm), hm * um / (hm - h_right)


def test_stoker_dam_break():
    nx, ny, dx, hl, hr = 400, 8, 0.5, 2.0, 0.5
    grid = TileGrid.from_mask(np.ones((ny, nx), bool), 0.0, 0.0, dx)
    tables = build_tables(grid, ArrayBed(np.zeros(grid.shape), dx), Tables(), 1e-6)
    m = OracleModel(tables, PLAIN, prec=8, threads=4)
    x0 = 0.5 * nx * dx
    xc = (np.arange(grid.shape[1]) - 1 + 0.5) * dx  # padded column 1 is the domain's first
    h0 = np.where(xc <= x0, hl, hr)[None, :] * grid.mask
    m.set_depth(h0)
    v0 = m.volume()
    t_end = 15.0
    m.advance(t_end)
    hm, um, cm, s = _stoker(hl, hr)
    c_left = math.sqrt(G * hl)
    xi = (xc - x0) / t_end
    fan = (2.0 * c_left - xi) ** 2 / (9.0 * G)
    exact = np.select([xi <= -c_left, xi <= um - cm, xi <= s], [hl, fan, hm], hr)
    h = m.field("h")[1 : 1 + ny, 1 : 1 + nx]
    rmse = float(np.sqrt(((h - exact[None, 1 : 1 + nx]) ** 2).mean()))
    assert rmse < 0.02
    assert abs(m.volume() / v0 - 1.0) < 1e-6


@pytest.mark.parametrize("sheet", [False, True])
@pytest.mark.parametrize("prec", [8, 4])
def test_still_water_stays_still_on_rough_ground_at_altitude(sheet, prec):
    rng = np.random.default_rng(11)
    n, dx, k = 100, 10.0, 4
    grid = TileGrid.from_mask(np.ones((n, n), bool), 0.0, 0.0, dx)
    z = 1500.0 + gaussian_filter(rng.normal(0.0, 1.0, (grid.shape[0] * k, grid.shape[1] * k)), 6) * 8.0
    tables = build_tables(grid, ArrayBed(z, dx / k), Tables(), 0.03)
    eta = float(np.median(z))
    h = hbar_at_level(eta, tables.zmin, grid.from_tiles(tables.hc).astype(np.float64),
                      grid.from_tiles(tables.dtab).astype(np.float64)) * grid.mask  # fmt: skip
    m = OracleModel(tables, Numerics(sheet_flow=sheet, dt_every=1), prec=prec, threads=4)
    m.set_depth(h)
    m.advance(2.0)
    v0 = m.volume()
    m.advance(10.0)
    hh = m.field("h")
    deep = (hh > 0.01) & (np.roll(hh, 1, axis=1) > 0.01) & (np.roll(hh, 1, axis=0) > 0.01)
    assert deep.sum() > 1000
    assert np.abs(m.field("u"))[deep].max() < 1e-2
    assert np.abs(m.field("v"))[deep].max() < 1e-2
    assert abs(m.volume() / v0 - 1.0) < 1e-6


@pytest.mark.parametrize("depth, mm_h", [(20.0, 10.0), (20.0, 30.0), (2.0, 3.0)])
def test_rain_on_deep_water_is_stored_in_full(depth, mm_h):
    """In float32 without the compensated residual 0 % is stored at 20 m and 10 mm/h; with it, all of it."""
    m, grid = _flat_model(30, 10.0, prec=4)
    h0 = np.where(grid.mask, depth, 0.0)
    m.set_depth(h0)
    rate = np.where(grid.mask, mm_h / 3.6e6, 0.0)
    m.set_source(rate)
    m.advance(3600.0)
    stored = float((m.field("h") + m.field("res") - h0).sum())
    assert stored == pytest.approx(float(rate.sum()) * 3600.0, rel=0.01)
    assert float((m.field("h") - h0).sum()) == pytest.approx(float(rate.sum()) * 3600.0, rel=0.01)


def test_thin_crest_passes_critical_flow():
    nx, ny, k, dx, crest, head = 120, 4, 4, 5.0, 3.0, 1.0
    grid = TileGrid.from_mask(np.ones((ny, nx), bool), 0.0, 0.0, dx)
    p = dx / k
    xs = grid.x0 + (np.arange(grid.shape[1] * k) + 0.5) * p
    i_crest = 61
    z = np.zeros((grid.shape[0] * k, grid.shape[1] * k))
    z[:, (xs >= i_crest * dx) & (xs < i_crest * dx + p)] = crest
    tables = build_tables(grid, ArrayBed(z, p), Tables(), 0.012)
    m = OracleModel(tables, Numerics(cfl=0.4, dt_every=1), prec=8, threads=4)
    h0 = np.zeros(grid.shape)
    h0[:, :i_crest] = crest + head
    m.set_depth(h0 * grid.mask)
    cells = [m.cell_state_index(j + 1, c) for j in range(ny) for c in (1, 2, 3)]
    nb = len(cells)
    m.set_boundary(BoundaryCells(np.array(cells, np.int32), np.zeros(nb, np.int32), np.zeros(nb, np.int32),
                                 np.full(nb, crest + head, np.float32), np.full(nb, crest + head, np.float32)))
    m.advance(300.0)
    q = float(m.field("qx")[2, i_crest]) / dx
    hup = float(m.field("h")[2, i_crest - 3]) - crest
    assert q == pytest.approx(CRITICAL_WEIR_COEFF * hup**1.5, rel=0.01)


def test_sheet_flow_on_a_plane_matches_the_kinematic_wave():
    """Storage and outflow at equilibrium on a 5 % plane; a flat-surface reading would hold a seventh of it."""
    slope, n, length, dx, k = 0.05, 0.05, 2000.0, 20.0, 20
    nx, ny = int(length / dx), 3
    grid = TileGrid.from_mask(np.ones((ny, nx), bool), 0.0, 0.0, dx)
    p = dx / k
    xs = grid.x0 + (np.arange(grid.shape[1] * k) + 0.5) * p
    z = np.tile((length - xs) * slope, (grid.shape[0] * k, 1))
    tables = build_tables(grid, ArrayBed(z, p), Tables(), n)
    m = OracleModel(tables, Numerics(dt_max=2.0), prec=8, threads=4)
    i = 50.0 / 3.6e6
    rain = np.zeros(grid.shape)
    rain[1 : 1 + ny, 1:nx] = i
    m.set_source(rain)
    cells = [m.cell_state_index(j + 1, nx) for j in range(ny)]
    m.set_boundary(BoundaryCells(np.array(cells, np.int32), np.zeros(ny, np.int32), np.ones(ny, np.int32),
                                 np.zeros(ny, np.float32), np.zeros(ny, np.float32)))
    tc = (n * length / (slope**0.5 * i ** (2.0 / 3.0))) ** 0.6
    m.advance(2.5 * tc)
    q_end = i * length
    h_end = (q_end * n / slope**0.5) ** 0.6
    hmean = float(m.field("h")[1 + ny // 2, 1:nx].mean())
    assert hmean == pytest.approx(0.625 * h_end, rel=0.03)
    assert float(m.field("qx")[1 + ny // 2, nx]) / dx == pytest.approx(q_end, rel=0.03)


def test_nan_stops_the_run():
    m, grid = _flat_model(30, 10.0)
    m.set_depth(np.where(grid.mask, 1.0, 0.0))
    u = m.backend.get_field("u")
    u[m.cell_state_index(10, 10)] = float("nan")
    m.backend.set_field("u", u)
    m.stepper.refresh_due = True
    with pytest.raises(NumericalFailure):
        m.advance(5.0)


def test_first_step_on_dry_ground_is_not_the_longest_and_an_empty_grid_runs():
    m, _ = _flat_model(30, 20.0)
    assert m.stepper.dt <= m.num.cfl * 20.0 / m.num.wave_floor + 1e-9
    m.advance(10.0)
    assert m.stepper.t == pytest.approx(10.0)
