"""The device table builder (w2d_build_tables_device, wadi2d_amd/csrc/tables_gpu.hip) against the host builder
and the reference's golden tables.

Both builders run the rows of csrc/table_rows.hpp in float64 and cast to float32 at the end. Arrays made of
additions, products, quotients, minima and maxima only (zmin, zbar, the cell tables, the tops, the areas, the
inverts, the centre beds) must be equal to the bit. The conveyance column goes through cbrt and pow, which the
device's math library computes to within an ulp of the host's in float64: there every entry must lie within one
float32 ulp of the host's and almost all must be equal.
"""

from __future__ import annotations


import numpy as np
import pytest

from wadi2d_amd.config import Tables
from wadi2d_amd.grid import TileGrid
from wadi2d_amd.tables import ArrayBed, CrestFaces, build_tables

from test_tables import FIELDS, GOLDEN, w85_bed

pytestmark = pytest.mark.gpu

EXACT = tuple(f for f in FIELDS if not f.startswith("ftab"))


def _close(name, a, b):
    """Every entry within one float32 ulp, at least 99 % equal; the area column (even entries) equal."""
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, name
    assert np.array_equal(a[..., 0], b[..., 0]), name + " areas"
    gap = np.abs(a.astype(np.float64) - b.astype(np.float64))
    ulp = np.spacing(np.abs(b).astype(np.float32)).astype(np.float64)
    assert np.all(gap <= ulp), (name, float((gap / ulp).max()))
    assert np.mean(a == b) >= 0.99, (name, float(np.mean(a == b)))


def _same(host, dev):
    for key in EXACT:
        assert np.array_equal(getattr(dev, key), getattr(host, key), equal_nan=True), key
    for key in ("ftabx", "ftaby"):
        _close(key, getattr(dev, key), getattr(host, key))


@pytest.mark.parametrize("name", ["long", "short_flat"])
def test_the_device_tables_are_the_golden_ones(name):
    grid, bed = w85_bed(name)
    print(f"device tables: W85 {name}", flush=True)
    t = build_tables(grid, bed, Tables(), 0.03, device=True)
    with np.load(GOLDEN / f"w85_tables_{name}.npz", allow_pickle=False) as g:
        for key in EXACT:
            assert np.array_equal(getattr(t, key), g[key], equal_nan=True), key
        for key in ("ftabx", "ftaby"):
            _close(key, getattr(t, key), g[key])


def _ground(seed, ny, nx, k):
    rng = np.random.default_rng(seed)
    z = rng.normal(0.0, 1.0, (ny * k, nx * k)).cumsum(axis=1) * 0.2
    z += rng.normal(0.0, 1.0, (ny * k, nx * k)).cumsum(axis=0) * 0.1 + 100.0
    return z


@pytest.mark.parametrize("mode", ["local_radius", "strips"])
@pytest.mark.parametrize("k", [3, 4])
def test_device_and_host_tables_agree_on_rough_ground(mode, k):
    dx = 2.0 * k
    mask = np.ones((70, 45), bool)
    mask[60:, 30:] = False  # the south-east tile of the padded grid left out, its neighbours' faces closed
    grid = TileGrid.from_mask(mask, 0.0, 0.0, dx)
    ny, nx = grid.shape
    # the bed starts three pixels in and stops short of the grid: reads clip at both ends
    z = _ground(7 + k, ny, nx, k)[:-5, 3:]
    bed = ArrayBed(z, dx / k, row_offset=0, col_offset=-3)
    cfg = Tables(rows=12, face_conveyance=mode)
    man = np.random.default_rng(3).uniform(0.02, 0.08, (ny, nx))
    print(f"device tables: rough ground k={k} {mode} tiles={grid.ntiles}", flush=True)
    host = build_tables(grid, bed, cfg, man, workers=4)
    dev = build_tables(grid, bed, cfg, man, device=True)
    _same(host, dev)


def test_device_and_host_tables_agree_with_crests_on_faces():
    k, dx = 5, 10.0
    grid = TileGrid.from_mask(np.ones((60, 60), bool), 0.0, 0.0, dx)
    ny, nx = grid.shape
    bed = ArrayBed(_ground(11, ny, nx, k), dx / k)
    crest = CrestFaces.empty((ny, nx))
    rng = np.random.default_rng(5)
    for a, ln, c in ((crest.xz, crest.xlen, crest.xc), (crest.yz, crest.ylen, crest.yc)):
        pick = rng.random((ny, nx)) < 0.05
        a[pick] = 100.0 + rng.normal(0.0, 3.0, pick.sum())
        ln[pick] = rng.uniform(0.3, 1.0, pick.sum())
        c[pick & (rng.random((ny, nx)) < 0.5)] = 1.55
    print("device tables: crests", flush=True)
    host = build_tables(grid, bed, Tables(), 0.035, crest=crest, workers=4)
    dev = build_tables(grid, bed, Tables(), 0.035, crest=crest, device=True)
    _same(host, dev)
    assert np.count_nonzero(host.cwx) > 0


def test_the_device_builder_refuses_what_the_host_refuses():
    grid = TileGrid.from_mask(np.ones((30, 30), bool), 0.0, 0.0, 10.0)
    bed = ArrayBed(np.zeros((300, 300)), 3.0)  # 10 m is not whole pixels of 3 m
    with pytest.raises(ValueError):
        build_tables(grid, bed, Tables(), 0.03, device=True)
