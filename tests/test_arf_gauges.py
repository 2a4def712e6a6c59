"""Area reduction by table, the depth-area study's reading of its runs, and gauge rain (wadi2d/storm/arf.py,
wadi2d/arf.py:79-247, wadi2d/storm/gauges.py, storm/gauge_weights.py, hydrology/gauge_rain.py).

The expected numbers are the reference's own test vectors (tests/storm/test_arf_steps.py: Block 2's ``ARF FM`` table
and the areas it places; tests/storm/test_gauges.py: weights worked by hand) and closed forms: a power law of peak
against area is read back exactly by the log-log reading, and the log-area weights are 0 or 1 at the runs' own areas.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from wadi2d_amd.engine.arf import (ArfTable, GivenFactor, bracket_areas, composite_map, design_row, factor_at,
                                   interpolate_peak, log_area_weights)  # fmt: skip
from wadi2d_amd.engine.gauges import GaugeCells, GaugeStorm, fill_gaps, inverse_distance, nearest

ARF_FM = ArfTable(
    np.array([1.0, 5.0, 10.0, 20.0, 30.0, 40.0, 50.0, 100.0, 200.0, 500.0, 1000.0, 2000.0]),
    np.array([0.999, 0.994, 0.987, 0.976, 0.965, 0.956, 0.948, 0.913, 0.864, 0.776, 0.694, 0.606]),
    reduces_below_table=False,
)


def _steps(table, areas):
    return np.abs(np.diff([table.factor(a) for a in areas]))


@pytest.mark.parametrize(("limit", "runs"), [(0.05, 9), (0.02, 21), (0.01, 40)])
def test_step_areas_keep_every_change_within_the_limit_in_equal_steps(limit, runs):
    areas = ARF_FM.step_areas(10.0, 2000.0, limit)
    assert len(areas) == runs == math.ceil((0.987 - 0.606) / limit) + 1
    assert areas[0] == 10.0 and areas[-1] == 2000.0 and np.all(np.diff(areas) > 0.0)
    found = _steps(ARF_FM, areas)
    assert found.max() <= limit * (1.0 + 1e-12)
    assert found == pytest.approx(np.full(runs - 1, (0.987 - 0.606) / (runs - 1)), rel=1e-9)
    assert (0.987 - 0.606) / (len(areas) - 2) > limit  # one run fewer would break the limit


def test_the_table_reads_as_hms_logs_it_and_places_the_reference_areas():
    assert ARF_FM.factor(20.971) == pytest.approx(200.251 / 205.4, abs=5e-7)
    areas = ARF_FM.step_areas(10.0, 2000.0, 0.05)
    assert [round(a) for a in areas] == [10, 62, 143, 268, 430, 665, 956, 1459, 2000]
    assert ARF_FM.factor(areas[1]) == pytest.approx(0.987 - (0.987 - 0.606) / 8.0, abs=1e-12)
    areas = ARF_FM.step_areas(10.0, 2000.0, 0.02)
    assert areas[-1] / areas[-2] > 1.08 and areas[1] / areas[0] > 1.08  # no run a hair below either end


def test_edge_cases_of_the_table():
    assert ARF_FM.step_areas(20.971, 20.971, 0.05) == [20.971]
    flat = ArfTable(np.array([1.0, 100.0]), np.array([0.9, 0.9]))
    assert flat.step_areas(10.0, 50.0, 0.05) == [10.0, 50.0]
    odd = ArfTable(np.array([1.0, 10.0, 100.0, 1000.0]), np.array([1.0, 0.8, 0.9, 0.5]))
    areas = odd.step_areas(1.0, 1000.0, 0.05)
    assert _steps(odd, areas).max() <= 0.05 * (1.0 + 1e-12)
    assert len(areas) == math.ceil((0.2 + 0.1 + 0.4) / 0.05) + 1
    areas = ARF_FM.step_areas(0.5, 50.0, 0.01)  # below the table the depth is whole: the jump counts
    assert _steps(ARF_FM, areas).max() <= 0.01 * (1.0 + 1e-12) and ARF_FM.factor(areas[0]) == 1.0
    with pytest.raises(ValueError, match="above zero"):
        ARF_FM.step_areas(10.0, 2000.0, 0.0)
    with pytest.raises(ValueError):
        ArfTable(np.array([1.0, 1.0]), np.array([0.9, 0.8]))
    assert factor_at(None, 10.0) == 1.0 and factor_at(ARF_FM, None) == 1.0
    assert factor_at(GivenFactor(0.7), 1e6) == 0.7 and factor_at(ARF_FM, 2000.0) == 0.606
    assert ArfTable(np.array([1.0, 5.0]), np.array([0.99, 0.9])).factor(0.1) == 0.99  # reduces below the table


def test_the_study_reads_a_power_law_back_exactly_and_weights_runs_by_log_area():
    areas = np.array([10.0, 40.0, 160.0, 640.0])
    peaks = 3.0 * areas**0.6
    for a in (10.0, 25.0, 100.0, 639.0):
        assert interpolate_peak(areas, peaks, a) == pytest.approx(3.0 * a**0.6, rel=1e-12)
    assert interpolate_peak(areas, peaks, 5.0) == pytest.approx(peaks[0], rel=1e-12)  # the end runs hold
    assert interpolate_peak(areas, np.zeros(4), 50.0) == 0.0
    assert interpolate_peak(areas, np.array([0.0, 2.0, 4.0, 8.0]), 10.0) == pytest.approx(2.0)  # dry run passed over
    low, high, w = log_area_weights(areas, np.array([10.0, 40.0, 80.0, 5000.0]))
    assert low.tolist() == [0, 0, 1, 2] and high.tolist() == [1, 1, 2, 3]
    assert w == pytest.approx([0.0, 1.0, 0.5, 1.0])
    maps = [np.full((2, 2), p) for p in peaks]
    cell = np.array([[10.0, 20.0], [640.0, 1.0]])
    got = composite_map(maps, areas, cell)
    assert got[0, 0] == peaks[0] and got[1, 0] == peaks[3] and got[1, 1] == peaks[0]
    assert got[0, 1] == pytest.approx(0.5 * (peaks[0] + peaks[1]))
    row = design_row("outlet", 100.0, areas, peaks, np.array([1.0, 2.0, 3.0, 4.0]))
    assert row["storm_area_low_km2"] == 40.0 and row["storm_area_high_km2"] == 160.0
    assert row["design_peak_m3s"] == pytest.approx(3.0 * 100.0**0.6) and row["time_of_peak_s"] == 3.0
    assert bracket_areas([10.0, 100.0], 2.0, 20) == pytest.approx(list(np.geomspace(10.0, 100.0, 5)))
    assert bracket_areas([7.0, 7.0], 2.0, 20) == [7.0]


TWO = np.array([[0.0, 0.0], [10.0, 0.0]])


def test_gauge_weights_as_the_reference_works_them_by_hand():
    assert nearest(np.array([1.0, 9.0, 5.0]), np.array([0.0, 0.0, 0.0]), TWO).tolist() == [0, 1, 0]
    w, beyond = inverse_distance(np.array([0.0]), np.array([0.0]), np.array([[1.0, 0.0], [0.0, 2.0]]), 2.0, None,
                                 quadrants=False)  # fmt: skip
    assert w[0] == pytest.approx([0.8, 0.2]) and beyond == 0
    three = np.array([[1.0, 0.0], [0.0, 2.0], [-2.0, 0.0]])
    w, _ = inverse_distance(np.array([0.0]), np.array([0.0]), three, 2.0, None, quadrants=True)
    assert w[0] == pytest.approx([0.8, 0.0, 0.2])
    w, _ = inverse_distance(np.array([0.0]), np.array([0.0]), three, 2.0, None, quadrants=False)
    assert w[0] == pytest.approx(np.array([1.0, 0.25, 0.25]) / 1.5)
    w, beyond = inverse_distance(np.array([0.0, 50.0]), np.array([0.0, 0.0]), np.array([[1.0, 0.0], [0.0, 3.0]]),
                                 2.0, 2.0, quadrants=False)  # fmt: skip
    assert w[0] == pytest.approx([1.0, 0.0]) and w[1] == pytest.approx([1.0, 0.0]) and beyond == 1
    w, _ = inverse_distance(np.array([10.0]), np.array([0.0]), TWO, 2.0, None, quadrants=False)
    assert w[0] == pytest.approx([0.0, 1.0])


def test_gauge_depths_and_cells():
    storm = GaugeStorm(("a", "b"), TWO, np.array([0.0, 10.0, 30.0]), np.array([[0.0, 0.0], [5.0, 1.0], [5.0, 9.0]]))
    assert storm.depths_at(-5.0) == pytest.approx([0.0, 0.0])
    assert storm.depths_at(5.0) == pytest.approx([2.5, 0.5])
    assert storm.depths_at(20.0) == pytest.approx([5.0, 5.0])
    assert storm.depths_at(99.0) == pytest.approx([5.0, 9.0])
    cells = GaugeCells(storm, np.array([1.0, 9.0, 2.0]), np.array([0.0, 0.0, 1.0]))
    assert cells.used.tolist() == [0, 1] and cells.column.tolist() == [0, 1, 0]
    assert cells.cell_depths(1200.0) == pytest.approx([5.0, 5.0, 5.0])
    idw = GaugeStorm(("a", "b"), TWO, storm.edges, storm.cumulative, method="inverse_distance", quadrants=False)
    cells = GaugeCells(idw, np.array([2.0]), np.array([0.0]))
    w = np.array([1 / 4, 1 / 64]) / (1 / 4 + 1 / 64)
    assert cells.cell_depths(600.0) == pytest.approx([w @ [5.0, 1.0]])
    with pytest.raises(ValueError):
        GaugeStorm(("a",), np.zeros((1, 2)), np.array([0.0, 1.0]), np.array([[2.0], [1.0]]))  # a falling record


def test_gaps_rain_nothing_or_what_the_neighbours_read_at_the_gauge():
    gauges = np.array([[0.0, 0.0], [10.0, 0.0], [5.0, 0.0]])
    depths = np.array([[2.0, 4.0, np.nan], [1.0, 1.0, 1.0]])
    assert fill_gaps(depths, gauges, "zero", None)[0].tolist() == [2.0, 4.0, 0.0]

    def weigh(x, y, have):
        return inverse_distance(x, y, have, 2.0, None, quadrants=False)[0]

    out = fill_gaps(depths, gauges, "neighbours", weigh)
    assert out[0] == pytest.approx([2.0, 4.0, 3.0]) and out[1].tolist() == [1.0, 1.0, 1.0]
