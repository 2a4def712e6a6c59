"""The table builder (wadi2d_amd/csrc/tables.cpp) against the reference's own golden tables and closed forms.

* W85 golden tables: three analytic beds built by the reference's builder (``W85 what limits the time
  step/scripts/s08_builder_margin.py:29-42``), extracted unchanged into tests/golden. ``long`` and ``short_flat`` are
  the current builder's output to the digit (W85 README:174-181): every one of the 20 arrays must be equal.
  ``short_slope`` was saved before the window cap: equal except face-table rows whose local-radius window reaches
  past the bed's own length, so it is compared at rows shallower than the bed length over beta.
* Row invariants that fix each table without the reference: a cell row's level is exactly the level whose mean
  depth over the cell's pixels is Hc (r/m)^2; a face row's area is the sum of the pixel depths under the row's level;
  in strips mode its conveyance^(3/5) is (pixel/n sum dep^(5/3))^0.6 by brute force.
* The local hydraulic radius on the reference's gorge section (tests/tables/test_local_radius.py:20-111) against a
  direct O(n^2) evaluation of Spada et al. (2015) eqs. 14-17, within 1 %; on a bed with no relief across the faces it
  is the strips sum to 1e-5.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from wadi2d_amd.config import Tables
from wadi2d_amd.grid import TileGrid
from wadi2d_amd.tables import ArrayBed, build_tables

GOLDEN = Path(__file__).resolve().parent / "golden"
FIELDS = ("zmin", "zbar", "hc", "dtab", "flx", "frx", "glx", "grx", "efx", "cwx", "ftabx",
          "fly", "fry", "gly", "gry", "efy", "cwy", "ftaby", "zcx", "zcy")  # fmt: skip


def w85_bed(name: str):
    """The three beds of s08_builder_margin.py:29-42, restated: a 30 x 30 grid padded to one tile."""
    dx, k = {"long": (10.0, 10), "short_flat": (1.25, 5), "short_slope": (1.25, 5)}[name]
    grid = TileGrid.from_mask(np.ones((30, 30), bool), 0.0, 0.0, dx)
    ny, nx = grid.shape
    pixel = dx / k
    y, x = np.mgrid[0 : ny * k, 0 : nx * k] * pixel
    if name == "short_flat":
        z = np.full(x.shape, 50.0)
    else:
        width = nx * dx
        z = 50.0 - 0.02 * x + 0.3 * np.abs(y - 0.5 * width) + 0.5 * np.sin(x / (0.05 * width))
    return grid, ArrayBed(z, pixel)


@pytest.mark.parametrize("name", ["long", "short_flat"])
def test_golden_tables_equal_to_the_digit(name):
    grid, bed = w85_bed(name)
    t = build_tables(grid, bed, Tables(), 0.03, workers=2)
    with np.load(GOLDEN / f"w85_tables_{name}.npz", allow_pickle=False) as g:
        for key in FIELDS:
            assert np.array_equal(getattr(t, key), g[key], equal_nan=True), key


def test_golden_short_slope_equal_where_the_window_reaches_no_further_than_the_bed():
    grid, bed = w85_bed("short_slope")
    cfg = Tables()
    t = build_tables(grid, bed, cfg, 0.03, workers=2)
    with np.load(GOLDEN / "w85_tables_short_slope.npz", allow_pickle=False) as g:
        for key in FIELDS:
            if not key.startswith("ftab"):
                assert np.array_equal(getattr(t, key), g[key]), key
        reach = bed.span * bed.pixel / cfg.local_radius_beta  # 4.44 m: a window this deep ends at the bed's edge
        for axis in "xy":
            a, b = getattr(t, "ftab" + axis), g["ftab" + axis]
            top = getattr(t, "ef" + axis).astype(np.float64)
            rows = a.shape[-2]
            depth = (np.arange(rows) / (rows - 1)) ** 2 * top[..., None]
            shallow = depth < reach
            assert np.array_equal(a[shallow], b[shallow])
            assert shallow.sum() > 0.4 * shallow.size
            # the areas do not depend on the window at all
            assert np.array_equal(a[..., 0], b[..., 0])


def _random_case(seed=4, n=40, k=6, dx=12.0, mode="strips", rows=10):
    rng = np.random.default_rng(seed)
    grid = TileGrid.from_mask(np.ones((n, n), bool), 0.0, 0.0, dx)
    ny, nx = grid.shape
    z = rng.normal(0.0, 1.0, (ny * k, nx * k)).cumsum(axis=1) * 0.2 + 100.0
    bed = ArrayBed(z, dx / k)
    return grid, bed, build_tables(grid, bed, Tables(rows=rows, face_conveyance=mode), 0.04), k


def test_cell_rows_hold_the_mean_depth_they_stand_for():
    grid, bed, t, k = _random_case()
    last = t.rows - 1
    for (j, i) in [(5, 7), (20, 3), (33, 33), (1, 1)]:
        tt, oj, oi = grid.locate(j, i)
        pix = bed.z[j * k : (j + 1) * k, i * k : (i + 1) * k].ravel()
        rel = pix - pix.min()
        hc = float(t.hc[tt, oj, oi])
        assert hc == pytest.approx(rel.max() - rel.mean(), rel=1e-6)
        for r in range(last):
            level = float(t.dtab[tt, oj, oi, r])
            hbar = np.maximum(level - rel, 0.0).mean()
            assert hbar == pytest.approx(hc * (r / last) ** 2, rel=1e-5, abs=1e-6)


def test_face_rows_by_brute_force_in_strips_mode():
    grid, bed, t, k = _random_case()
    last = t.rows - 1
    pixel = bed.pixel
    for (j, i) in [(5, 7), (20, 3), (33, 33)]:
        tt, oj, oi = grid.locate(j, i)
        # the west face: the higher of the two pixels either side, along the face
        prof = np.maximum(bed.z[j * k : (j + 1) * k, i * k - 1], bed.z[j * k : (j + 1) * k, i * k])
        rel = prof - prof.min()
        top = float(t.efx[tt, oj, oi])
        assert top == pytest.approx(max(min(rel.max(), 16.0), 8.0), rel=1e-6)
        for r in range(t.rows):
            e = top * (r / last) ** 2
            dep = np.maximum(e - rel, 0.0)
            area = pixel * dep.sum()
            conv = (pixel / 0.04 * (dep ** (5.0 / 3.0)).sum()) ** 0.6
            assert float(t.ftabx[tt, oj, oi, r, 0]) == pytest.approx(area, rel=1e-5, abs=1e-6)
            assert float(t.ftabx[tt, oj, oi, r, 1]) == pytest.approx(conv, rel=1e-5, abs=1e-6)


# ---- the reference's gorge (tests/tables/test_local_radius.py:20-37), restated
PIXEL, S, N, BETA = 0.5, 0.01, 0.05, 9.0
NY_PIX, NX_PIX = 640, 160


def _section(t):
    a = np.abs(t)
    h = np.where(a < 6.0, 0.0, (a - 6.0) / 1.5 * 7.0)
    h = np.where(a >= 7.5, 7.0 + (a - 7.5) * 0.05, h)
    return np.where(a >= 17.5, 7.5 + (a - 17.5) * 0.5, h)


def _gorge(flat=False):
    y = (np.arange(NY_PIX) + 0.5) * PIXEL - 0.5 * NY_PIX * PIXEL
    x = (np.arange(NX_PIX) + 0.5) * PIXEL
    across = np.zeros(NY_PIX) if flat else _section(y)
    return (100.0 - S * x[None, :] + across[:, None]).astype(np.float32).astype(np.float64)


def _lookup_t(e, top, tab, rows):
    """Conveyance^(3/5) at depth e from a face table, float64 (the step's face lookup, quadratic rows)."""
    last = rows - 1
    m = float(last)
    if e >= top:
        a1, a0 = tab[last, 0], tab[last - 1, 0]
        c = (m - 1.0) / m
        w = (a1 - a0) / (top * (1.0 - c * c))
        area = a1 + w * (e - top)
        return tab[last, 1] * area / max(a1, 1e-12)
    s = np.sqrt(e / top) * m
    r = min(int(s), last - 1)
    e0 = (r / m) ** 2 * top
    de = ((r + 1) ** 2 - r**2) / m**2 * top
    return tab[r, 1] + (e - e0) / de * (tab[r + 1, 1] - tab[r, 1])


def _column_conveyance(dx, mode, over, z):
    k = round(dx / PIXEL)
    ny, nx = NY_PIX // k, NX_PIX // k
    grid = TileGrid.from_mask(np.ones((ny, nx), bool), 0.0, 0.0, dx)
    t = build_tables(grid, ArrayBed(z, PIXEL, row_offset=-k, col_offset=-k), Tables(face_conveyance=mode), N)
    i = nx // 2
    level = 100.0 - S * i * dx + over
    ftab, ef, fl = grid.from_tiles(t.ftabx), grid.from_tiles(t.efx), grid.from_tiles(t.flx)
    total = 0.0
    for j in range(ny):
        jj, ii = j + 1, i + 1
        e = level - (fl[jj, ii] + t.zmin[jj, ii - 1])
        if e > 0.0:
            total += _lookup_t(float(e), float(ef[jj, ii]), ftab[jj, ii].astype(np.float64), t.rows) ** (5.0 / 3.0)
    return total


def _direct_spada(over):
    """Eqs. 14-17 of Spada et al. evaluated directly at the pixel spacing: O(n^2), no prefix sums."""
    y = (np.arange(NY_PIX) + 0.5) * PIXEL - 0.5 * NY_PIX * PIXEL
    z = _section(y)
    d = np.clip(over - z, 0.0, None)
    wet = d > 0.0
    per = np.zeros(z.size)
    for q in range(z.size - 1):
        seg = float(np.hypot(PIXEL, z[q + 1] - z[q]))
        if wet[q] and wet[q + 1]:
            per[q] += 0.5 * seg
            per[q + 1] += 0.5 * seg
        elif wet[q] or wet[q + 1]:
            low = q if wet[q] else q + 1
            per[low] += min(d[low] / abs(z[q + 1] - z[q]), 1.0) * seg
    total = 0.0
    idx = np.nonzero(wet)[0]
    for i in idx:
        lo, hi = i, i
        while lo > 0 and wet[lo - 1]:
            lo -= 1
        while hi < z.size - 1 and wet[hi + 1]:
            hi += 1
        w = BETA * d[i] / PIXEL
        j = np.arange(lo, hi + 1)
        nw = np.clip(1.0 - np.abs(j - i) / w, 0.0, None)
        radius = PIXEL * float((d[j] * nw).sum()) / float((per[j] * nw).sum())
        total += d[i] * radius ** (2.0 / 3.0) * PIXEL / N
    return total


def test_no_relief_across_the_faces_gives_the_strips_sum():
    z = _gorge(flat=True)
    a = _column_conveyance(20.0, "strips", 2.0, z)
    b = _column_conveyance(20.0, "local_radius", 2.0, z)
    assert b == pytest.approx(a, rel=1e-5)


@pytest.mark.parametrize("dx", [20.0, 10.0])
def test_the_gorge_conveyance_is_spada_et_al_s(dx):
    z = _gorge()
    for over in (2.0, 4.0, 5.3):
        assert _column_conveyance(dx, "local_radius", over, z) == pytest.approx(_direct_spada(over), rel=0.01)
