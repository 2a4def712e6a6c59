"""Rain losses and the excess form of the rain source on the CPU (wadi2d_amd/engine/losses.py, engine/forcing.py).

Pins: the SCS runoff of USDA TR-55 (1986) worked by hand, the closed forms at the ends (CN 100 passes the rain
whole; nothing runs off up to the initial abstraction), the reference's own loss tests
(tests/hydrology/test_losses_transforms.py:24-57) and forcing test (tests/engine/test_forcing.py:67-80), and the
excess form reducing to the weight form where it must."""

from __future__ import annotations

import itertools

import numpy as np
import pytest

from wadi2d_amd.config import Tables
from wadi2d_amd.engine.forcing import Rain, boundary_set, uniform_storm
from wadi2d_amd.engine.losses import NoLoss, ScsCurveNumber, cell_excess
from wadi2d_amd.grid import TileGrid
from wadi2d_amd.tables import ArrayBed, build_tables

RAIN = np.array([0.0, 2.0, 5.0, 12.0, 20.0, 8.0, 3.0, 1.0, 0.0, 0.5])


def test_tr55_runoff_worked_by_hand():
    # TR-55: P = 5 in on CN 80: S = 1000/80 - 10 = 2.5 in, Ia = 0.5 in, Q = 4.5^2 / 7 = 2.893 in
    loss = ScsCurveNumber(80.0)
    assert loss.retention_mm == pytest.approx(2.5 * 25.4, rel=1e-15)
    assert loss.abstraction_mm == pytest.approx(0.5 * 25.4, rel=1e-15)
    assert float(loss.cumulative_excess(5.0 * 25.4)) == pytest.approx(4.5**2 / 7.0 * 25.4, rel=1e-14)
    # P = 2 in on CN 61 (S = 6.393 in, Ia = 1.279 in): Q = 0.721^2 / 7.114 = 0.0731 in
    s61 = 1000.0 / 61.0 - 10.0
    q = (2.0 - 0.2 * s61) ** 2 / (2.0 - 0.2 * s61 + s61)
    assert float(ScsCurveNumber(61.0).cumulative_excess(2.0 * 25.4)) == pytest.approx(q * 25.4, rel=1e-13)
    assert q == pytest.approx(0.0731, abs=1e-4)


def test_the_ends_cn_100_and_the_initial_abstraction():
    p = np.array([0.0, 0.5, 7.0, 120.0])
    # CN 100: S = 0, no abstraction; 0 / 0 at zero depth reads 0 through nan_to_num
    assert np.array_equal(cell_excess(p, np.full(4, 100.0)), p)
    # below Ia (0.2 S = 16.93 mm on CN 75) nothing runs off; no curve number passes the depth
    assert np.all(cell_excess(np.array([0.0, 10.0, 16.9]), np.full(3, 75.0)) == 0.0)
    assert np.array_equal(cell_excess(p, np.full(4, np.nan)), p)


@pytest.mark.parametrize("cn", [55.0, 70.0, 85.0, 98.0])
def test_excess_never_exceeds_rain_and_is_never_negative(cn):
    excess = ScsCurveNumber(cn).excess(RAIN)
    assert np.all(excess >= 0.0)
    assert np.all(excess <= RAIN + 1e-12)


def test_excess_rises_with_curve_number():
    totals = [ScsCurveNumber(cn).excess(RAIN).sum() for cn in (60.0, 70.0, 80.0, 90.0, 98.0)]
    assert all(low < high for low, high in itertools.pairwise(totals))


def test_nothing_runs_off_before_the_initial_abstraction():
    loss = ScsCurveNumber(90.0, initial_abstraction_mm=10.0)
    assert loss.excess(np.array([3.0, 3.0, 3.0])).sum() == 0.0


def test_the_cumulative_form_does_not_depend_on_how_the_rain_is_split():
    loss = ScsCurveNumber(85.0, initial_abstraction_mm=7.0)
    assert loss.excess(np.full(40, 2.5)).sum() == pytest.approx(loss.excess(np.array([100.0])).sum())


def test_impervious_rain_runs_off_whole_and_no_loss_passes_it():
    assert ScsCurveNumber(60.0, percent_impervious=100.0).excess(RAIN) == pytest.approx(RAIN)
    assert np.array_equal(NoLoss().excess(RAIN), RAIN)


def _small(n=32, dx=4.0):
    grid = TileGrid.from_mask(np.ones((n, n), bool), 0.0, 0.0, dx)
    ny, nx = grid.shape
    k = 2
    y, x = np.mgrid[0 : ny * k, 0 : nx * k] * (dx / k)
    tables = build_tables(grid, ArrayBed(50.0 - 0.01 * x - 0.002 * y, dx / k), Tables(rows=8), 0.04, workers=2)
    return tables, boundary_set([], tables)


def test_window_rain_adds_up_to_the_whole_and_the_abstraction_takes_the_first_millimetres():
    tables, bset = _small()
    cn = np.full(tables.grid.shape, 75.0)
    rain = Rain(tables, bset, uniform_storm(60.0, 3600.0), curve_number=cn)
    whole = rain.rate(0.0, 3600.0)[1]
    parts = sum(rain.rate(t, t + 90.0)[1] for t in np.arange(0.0, 3600.0, 90.0))
    assert parts == pytest.approx(whole, rel=1e-5)
    assert rain.rate(0.0, 60.0)[1] == 0.0
    # the whole storm's runoff: 60 mm on CN 75 gives (60 - 16.93)^2 / (60 - 16.93 + 84.67) mm per cell
    s = (1000.0 / 75.0 - 10.0) * 25.4
    q = (60.0 - 0.2 * s) ** 2 / (60.0 - 0.2 * s + s)
    assert whole == pytest.approx(q / 1000.0 * rain.area * rain.state.size, rel=1e-6)


def test_the_excess_form_reduces_to_the_weight_form_and_scales_by_the_depth_grid():
    tables, bset = _small()
    ny, nx = tables.grid.shape
    storm = uniform_storm(45.0, 1800.0)
    plain = Rain(tables, bset, storm)
    ones = Rain(tables, bset, storm, depth_grid=np.ones((ny, nx)))
    nan_cn = Rain(tables, bset, storm, curve_number=np.full((ny, nx), np.nan))
    grid_depth = np.random.default_rng(2).uniform(20.0, 90.0, (ny, nx))
    unit = uniform_storm(1.0, 1800.0)
    scaled = Rain(tables, bset, unit, depth_grid=grid_depth)
    for t0, t1 in ((0.0, 30.0), (600.0, 630.0), (1790.0, 1820.0), (1900.0, 1930.0)):
        a, va = plain.rate(t0, t1)
        for other in (ones, nan_cn):
            b, vb = other.rate(t0, t1)
            assert np.array_equal(a, b) and va == vb
        c, _ = scaled.rate(t0, t1)
        d0, d1 = unit(t0), unit(t1)
        w = grid_depth[scaled.jj, scaled.ii]
        assert np.array_equal(c, ((d1 * w - d0 * w) / (1000.0 * (t1 - t0))).astype(np.float32))


def test_curve_numbers_out_of_range_are_refused():
    tables, bset = _small()
    cn = np.full(tables.grid.shape, 75.0)
    cn[5, 5] = 0.0
    with pytest.raises(ValueError):
        Rain(tables, bset, uniform_storm(10.0, 600.0), curve_number=cn)
    with pytest.raises(ValueError):
        Rain(tables, bset, uniform_storm(10.0, 600.0), np.ones(tables.grid.shape), curve_number=cn * 0 + 70.0)


# ---- gridded rain (storm/rain_frames.py, storm/gridded.py, hydrology/cell_rain.py)

def test_a_pixel_series_is_linear_in_a_period_and_level_across_a_gap():
    from wadi2d_amd.engine.gridded import PixelSeries, RainFrames

    depths = np.array([[[10.0, 2.0]], [[5.0, np.nan]]])  # 2 frames, 1 row, 2 pixels; pixel 1 has no data in frame 2
    frames = RainFrames(depths, np.array([0.0, 90.0]), np.array([60.0, 120.0]), 0.0, 100.0, 50.0, 100.0)
    s = PixelSeries(frames, np.array([0, 1]))
    for minutes, want in ((0.0, (0.0, 0.0)), (30.0, (5.0, 1.0)), (60.0, (10.0, 2.0)), (75.0, (10.0, 2.0)),
                          (105.0, (12.5, 2.0)), (200.0, (15.0, 2.0))):  # fmt: skip
        assert np.allclose(s.at(minutes), want, rtol=0, atol=1e-12), minutes
    assert np.array_equal(frames.pixel_of(np.array([10.0, 60.0, 120.0, 10.0]), np.array([50.0, 50.0, 50.0, -5.0])),
                          [0, 1, -1, -1])  # fmt: skip
    assert np.array_equal(frames.totals(), [[15.0, 2.0]])
    with pytest.raises(ValueError):
        RainFrames(depths, np.array([0.0, 30.0]), np.array([60.0, 120.0]), 0.0, 100.0, 50.0, 100.0)  # overlap


def frames_over(grid, pixel: float = 170.0, seed: int = 9, nan_frame: bool = True):
    """Seeded rain frames covering the grid: 6 periods of 10 min with a gap (30-40 min), 0-15 mm each."""
    from wadi2d_amd.engine.gridded import RainFrames

    ny, nx = grid.shape
    west, north = grid.x0 - 5.0, grid.y0 + 5.0
    cols = int(np.ceil((nx * grid.dx + 10.0) / pixel))
    rows = int(np.ceil((ny * grid.dx + 10.0) / pixel))
    rng = np.random.default_rng(seed)
    depths = rng.uniform(0.0, 15.0, (6, rows, cols))
    if nan_frame:
        depths[2, 0, 0] = np.nan  # a pixel missing in one frame counts as no rain there
    starts = np.array([0.0, 10.0, 20.0, 40.0, 50.0, 60.0])
    return RainFrames(depths, starts, starts + 10.0, west, north, pixel, pixel)


def test_gridded_rain_gives_each_cell_its_pixel_through_its_losses():
    tables, bset = _small()
    grid = tables.grid
    frames = frames_over(grid, pixel=37.0)
    cn = np.full(grid.shape, 70.0)
    cn[:, :10] = np.nan
    rain = Rain(tables, bset, None, curve_number=cn, frames=frames, area_factor=0.9)
    assert rain.excess_form and rain.pixels.used.size > 4
    x, y = grid.xy_of(rain.jj, rain.ii)
    pix = frames.pixel_of(x, y)
    for t0 in (0.0, 600.0, 1500.0, 2700.0, 4200.0):
        t1 = t0 + 30.0
        got, vol = rain.rate(t0, t1)
        # a cell's depth is its pixel's, times the area factor; no-CN cells pass it whole
        for m in (0, 5, rain.state.size // 2, rain.state.size - 1):
            col = np.nan_to_num(frames.depths.reshape(6, -1)[:, pix[m]])
            knots = np.concatenate(([0.0], np.cumsum(col[:3]), np.sum(col[:3]) + np.cumsum(col[3:])))

            def cum(t):  # the knots every 10 min, level from 30 to 40 min (the gap), times the area factor
                return 0.9 * float(np.interp(t / 60.0, [0.0, 10.0, 20.0, 30.0, 40.0, 50.0, 60.0, 70.0],
                                             np.insert(knots, 4, knots[3])))  # fmt: skip

            cn_m = np.array([cn[rain.jj[m], rain.ii[m]]])
            e0 = cell_excess(np.array([cum(t0)]), cn_m)[0]
            e1 = cell_excess(np.array([cum(t1)]), cn_m)[0]
            assert got[m] == pytest.approx((e1 - e0) / (1000.0 * 30.0), rel=1e-6, abs=1e-12), (t0, m)
        assert vol >= 0.0
    # a rained cell the grid does not cover is refused
    small = frames_over(grid, pixel=37.0)
    from wadi2d_amd.engine.gridded import RainFrames

    cut = RainFrames(small.depths[:, 1:, :], small.starts_min, small.ends_min, small.west, small.north - 37.0,
                     37.0, 37.0)  # fmt: skip
    with pytest.raises(ValueError, match="does not cover"):
        Rain(tables, bset, None, frames=cut)
