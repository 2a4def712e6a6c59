"""Seeded synthetic scenarios: the uniform-grid workloads this path is measured and checked on.

The reference defines none of these beds; they are this package's, fixed here and pinned by tests:

* ``plain_cpu_case`` -- 256 x 256 cells at 10 m (``c0_style_case``, which also makes the ensemble's members). Ground on a 1 m sub-grid (10 x 10 pixels per cell): a plane
  falling 0.5 % toward the east, plus 6-octave value noise (amplitude 2 m, base wavelength 640 m, persistence 0.5),
  plus a straight north-south wall 1.5 m high and 2 pixels thick at x = 1280 m with one 20 m gap centred on the
  domain. 8 table rows, Manning n 0.04, uniform rain 50 mm/h for 30 min then dry, 2 h, closed walls but a free
  (kind 1) outflow column on the east edge.
* ``precision_case`` -- n x n cells at 10 m (n = 2048), steep: 6-octave ridged noise (base wavelength 3 km,
  amplitude 100 m) on a 20 % fall to the east (cell-scale slopes: 5th percentile 11 %, median 21 %, 95th 30 %),
  on 5 x 5 sub-samples of 2 m; 40 seeded crest breaklines, straight ridges 0.5-3 m high, 10-50 m wide and
  200-1000 m long, the first running roughly north-south 1.5 km from the east edge. A reservoir of 1 km2 (a square of
  1000 m) lies on the uphill side of that first breakline, its surface 5 m above each cell's mean ground; that crest
  is gone from the ground the tables are built on (it broke at t = 0). No rain; all four edges free outflow;
  Manning n 0.035; 12 table rows; 1 h.
* ``full_extent_case`` -- n x n cells at 30 m (n = 4096: 16.8 M cells). Ground on 3 x 3 sub-samples of 10 m:
  8-octave value noise of amplitude 300 m (base wavelength 16 km, persistence 0.5), carved by 3 random-walk channels
  2 m deep and 60 m wide. 12 table rows; Manning n from a 4-class map (0.03/0.045/0.06/0.08, equal-area quartiles of a
  3-octave noise of base wavelength 4 km). Rain: a Gaussian storm cell centred on the domain, peak 80 mm/h, sigma
  20 km, on a triangular 3 h hyetograph peaking at 1.5 h. Four free outflow segments of 128 cells, one per edge,
  centred on that edge's lowest cell.

Noise is the host library's (csrc/synth.cpp); coordinates: the domain's north-west corner is (0, n dx).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from wadi2d_amd import _native
from wadi2d_amd.config import Numerics, Tables
from wadi2d_amd.engine.forcing import Boundary, triangular_storm, uniform_storm
from wadi2d_amd.grid import TileGrid
from wadi2d_amd.tables import ArrayBed, build_tables, table_workers


@dataclass
class Scenario:
    name: str
    grid: TileGrid
    bed: ArrayBed
    tables_cfg: Tables
    manning: float | np.ndarray
    boundaries: list[Boundary]
    cumulative_mm: object
    weight: np.ndarray | None
    duration_s: float
    numerics: Numerics
    tables: object = None
    curve_number: np.ndarray | None = None  # SCS curve numbers on the padded grid (NaN: no losses), or none at all

    def build(self, workers: int | None = None):
        if self.tables is None:
            self.tables = build_tables(self.grid, self.bed, self.tables_cfg, self.manning, workers=workers)
        return self.tables


def noise(out: np.ndarray, x0: float, y0: float, pixel: float, seed: int, octaves: int, base: float,
          persistence: float = 0.5, amplitude: float = 1.0, ridged: bool = False, threads: int | None = None) -> None:
    """Add value noise (or ridged noise) at the pixel centres of ``out`` (float64, C order), in place."""
    if out.dtype != np.float64 or not out.flags.c_contiguous:
        raise ValueError("noise writes into a C-contiguous float64 array")
    fn = _native.host().w2d_noise
    fn.restype = None
    fn.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                   ctypes.c_uint64, ctypes.c_int32, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int32,
                   ctypes.c_int32]  # fmt: skip
    fn(out.ctypes.data, out.shape[0], out.shape[1], x0, y0, pixel, seed & 0xFFFFFFFFFFFFFFFF, octaves, base,
       persistence, amplitude, int(ridged), table_workers(threads))  # fmt: skip


def _pixel_centres(grid: TileGrid, k: int):
    ny, nx = grid.shape
    p = grid.dx / k
    x = grid.x0 + (np.arange(nx * k) + 0.5) * p
    y = grid.y0 - (np.arange(ny * k) + 0.5) * p
    return x, y, p


def c0_style_case(n: int, dx: float, k: int, seed: int, storm, duration_s: float, name: str) -> Scenario:
    """n x n cells: a plane falling 0.5 % east, 6-octave value noise (2 m, base 640 m), a wall 1.5 m high and 2
    pixels thick across the middle (x = n dx / 2) with one 20 m gap centred on the domain; 8 rows, n 0.04, a free
    east column."""
    grid = TileGrid.from_mask(np.ones((n, n), bool), 0.0, n * dx, dx)
    x, y, p = _pixel_centres(grid, k)
    z = np.zeros((y.size, x.size))
    noise(z, grid.x0, grid.y0, p, seed, 6, 640.0, 0.5, 2.0)
    z += 0.005 * (n * dx - x)[None, :]
    wall = (x >= 0.5 * n * dx) & (x < 0.5 * n * dx + 2 * p)
    gap = np.abs(y - 0.5 * n * dx) < 10.0
    z[np.ix_(~gap, wall)] += 1.5
    east = [Boundary("free", np.arange(1, n + 1), np.full(n, n))]
    return Scenario(name, grid, ArrayBed(z, p), Tables(rows=8), 0.04, east, storm, None, duration_s, Numerics())


def plain_cpu_case(seed: int = 0) -> Scenario:
    return c0_style_case(256, 10.0, 10, seed, uniform_storm(25.0, 1800.0), 7200.0, "plain 256x256 dx10")


def _walks(rng, nsteps: int, extent: float, pixel: float, count: int) -> list[np.ndarray]:
    out = []
    for _ in range(count):
        pos = rng.uniform(0.3, 0.7, 2) * extent
        heading = rng.uniform(0.0, 2.0 * np.pi)
        turns = np.cumsum(rng.normal(0.0, 0.05, nsteps)) + heading
        pts = pos + np.cumsum(np.stack((np.cos(turns), np.sin(turns)), axis=1) * pixel, axis=0)
        inside = np.all((pts > 0.0) & (pts < extent), axis=1)
        stop = int(np.argmin(inside)) if not inside.all() else nsteps
        out.append(pts[:stop])
    return out


def full_extent_case(n: int = 4096, seed: int = 1) -> Scenario:
    dx, k = 30.0, 3
    grid = TileGrid.from_mask(np.ones((n, n), bool), 0.0, n * dx, dx)
    x, y, p = _pixel_centres(grid, k)
    z = np.full((y.size, x.size), 300.0)
    noise(z, grid.x0, grid.y0, p, seed, 8, 16000.0, 0.5, 300.0)
    # three random-walk channels, 2 m deep and 60 m wide, carved in the sub-grid
    rng = np.random.default_rng(seed)
    extent = n * dx
    carve = np.zeros(z.shape, bool)
    r = int(round(30.0 / p))
    offs = [(a, b) for a in range(-r, r + 1) for b in range(-r, r + 1) if (a * a + b * b) * p * p <= 30.0**2]
    for pts in _walks(rng, int(2 * extent / p), extent, p, 3):
        cols = np.floor((pts[:, 0] - grid.x0) / p).astype(np.int64)
        rows = np.floor((grid.y0 - pts[:, 1]) / p).astype(np.int64)
        for a, b in offs:
            rr = np.clip(rows + a, 0, z.shape[0] - 1)
            cc = np.clip(cols + b, 0, z.shape[1] - 1)
            carve[rr, cc] = True
    z[carve] -= 2.0
    del carve
    # four-class roughness
    ny, nx = grid.shape
    cls_noise = np.zeros((ny, nx))
    noise(cls_noise, grid.x0, grid.y0, dx, seed + 7, 3, 4000.0, 0.5, 1.0)
    inner = grid.mask
    q = np.quantile(cls_noise[inner], [0.25, 0.5, 0.75])
    manning = np.array([0.03, 0.045, 0.06, 0.08])[np.digitize(cls_noise, q)]
    # Gaussian storm cell on the cell centres
    cx = grid.x0 + (np.arange(nx) + 0.5) * dx
    cy = grid.y0 - (np.arange(ny) + 0.5) * dx
    sigma = 20000.0
    weight = np.exp(-((cx[None, :] - 0.5 * extent) ** 2 + (cy[:, None] - 0.5 * extent) ** 2) / (2.0 * sigma**2))
    scen = Scenario(f"full extent {n}x{n} dx30", grid, ArrayBed(z, p), Tables(rows=12), manning, [],
                    triangular_storm(80.0, 5400.0, 5400.0), weight, 3 * 3600.0, Numerics())  # fmt: skip
    scen.domain = n
    return scen


def _segments_distance(px: np.ndarray, py: np.ndarray, a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Distance of points (px, py) to the segment a-b."""
    d = b - a
    t = np.clip(((px - a[0]) * d[0] + (py - a[1]) * d[1]) / max(float(d @ d), 1e-12), 0.0, 1.0)
    return np.hypot(px - (a[0] + t * d[0]), py - (a[1] + t * d[1]))


def precision_case(n: int = 2048, seed: int = 3, k: int = 5, duration_s: float = 3600.0) -> Scenario:
    dx = 10.0
    grid = TileGrid.from_mask(np.ones((n, n), bool), 0.0, n * dx, dx)
    x, y, p = _pixel_centres(grid, k)
    extent = n * dx
    z = np.full((y.size, x.size), 500.0)
    noise(z, grid.x0, grid.y0, p, seed, 6, 3000.0, 0.5, 100.0, ridged=True)
    z += 0.2 * (extent - x)[None, :]
    rng = np.random.default_rng(seed)
    crests = []
    # the crest that breaks: north-south, 1.5 km from the east edge, so the flood reaches the outfall within the hour
    cy0 = rng.uniform(0.3, 0.7) * extent
    half = 0.5 * rng.uniform(200.0, 1000.0) * np.array([np.sin(0.2), np.cos(0.2)])
    c0 = np.array([extent - 1500.0, cy0])
    crests.append((c0 - half, c0 + half, rng.uniform(0.5, 3.0), rng.uniform(10.0, 50.0)))
    for _ in range(39):
        c = rng.uniform(0.1, 0.9, 2) * extent
        length = rng.uniform(200.0, 1000.0)
        ang = rng.uniform(0.0, np.pi)
        half = 0.5 * length * np.array([np.cos(ang), np.sin(ang)])
        crests.append((c - half, c + half, rng.uniform(0.5, 3.0), rng.uniform(10.0, 50.0)))
    # the reservoir: a 1000 m square on the uphill side of the first crest, touching it at its middle
    a0, b0, _, w0 = crests[0]
    mid = 0.5 * (a0 + b0)
    normal = np.array([-(b0 - a0)[1], (b0 - a0)[0]]) / np.hypot(*(b0 - a0))

    def ground_at(q):
        i = int(np.clip((q[0] - grid.x0) / p, 0, x.size - 1))
        j = int(np.clip((grid.y0 - q[1]) / p, 0, y.size - 1))
        return z[j, i]

    side = 1.0 if ground_at(mid + normal * 300.0) > ground_at(mid - normal * 300.0) else -1.0
    centre = mid + side * normal * (0.5 * w0 + 500.0)
    # every crest but the first (broken at t = 0) stands in the ground the tables are built on
    for a, b, height, width in crests[1:]:
        lo = np.minimum(a, b) - width
        hi = np.maximum(a, b) + width
        ci = np.nonzero((x >= lo[0]) & (x <= hi[0]))[0]
        rj = np.nonzero((y >= lo[1]) & (y <= hi[1]))[0]
        if ci.size == 0 or rj.size == 0:
            continue
        dist = _segments_distance(x[ci][None, :], y[rj][:, None], a, b)
        block = z[np.ix_(rj, ci)]
        block[dist <= 0.5 * width] += height
        z[np.ix_(rj, ci)] = block
    edges = [Boundary("free", np.full(n, 1), np.arange(1, n + 1)), Boundary("free", np.full(n, n), np.arange(1, n + 1)),
             Boundary("free", np.arange(2, n), np.full(n - 2, 1)), Boundary("free", np.arange(2, n), np.full(n - 2, n))]
    scen = Scenario(f"precision {n}x{n} dx10", grid, ArrayBed(z, p), Tables(rows=12), 0.035, edges,
                    uniform_storm(0.0, duration_s), None, duration_s, Numerics())  # fmt: skip
    ny, nx = grid.shape
    cx = grid.x0 + (np.arange(nx) + 0.5) * dx
    cy = grid.y0 - (np.arange(ny) + 0.5) * dx
    u = (cx[None, :] - centre[0]) * normal[1] - (cy[:, None] - centre[1]) * normal[0]
    v = (cx[None, :] - centre[0]) * normal[0] + (cy[:, None] - centre[1]) * normal[1]
    scen.reservoir = (np.abs(u) <= 500.0) & (np.abs(v) <= 500.0) & grid.mask
    scen.domain = n
    return scen


def crest_case(n: int, dx: float, k: int, seed: int, storm, duration_s: float, extent: float | None = None,
               name: str = "crested", density: float = 10.0, heights=(0.3, 1.5)) -> Scenario:
    """Ground defined in world coordinates over a square of side ``extent`` (default n dx), sampled on any grid of
    n x n cells (so a fine and a coarse grid of one seed see the same ground): a plane falling 0.5 % east, 6-octave
    value noise (2 m, base 640 m), and ``density`` crest ridges per km2, straight and flat-topped, 100-600 m long,
    10-80 m wide (0.5-4 cells of 20 m), ``heights`` high (uniform), at seeded places and angles. 8 rows, n 0.04, a
    free east column."""
    extent = n * dx if extent is None else extent
    grid = TileGrid.from_mask(np.ones((n, n), bool), 0.0, extent, dx)
    x, y, p = _pixel_centres(grid, k)
    z = np.zeros((y.size, x.size))
    noise(z, grid.x0, grid.y0, p, seed, 6, 640.0, 0.5, 2.0)
    z += 0.005 * (extent - x)[None, :]
    rng = np.random.default_rng(seed)
    count = int(round(density * (extent / 1000.0) ** 2))
    for _ in range(count):
        c = rng.uniform(0.1, 0.9, 2) * extent
        ang = rng.uniform(0.0, np.pi)
        half = 0.5 * rng.uniform(100.0, 600.0) * np.array([np.cos(ang), np.sin(ang)])
        width = rng.uniform(0.5, 4.0) * 20.0
        height = rng.uniform(*heights)
        a, b = c - half, c + half
        lo, hi = np.minimum(a, b) - width, np.maximum(a, b) + width
        ci = np.nonzero((x >= lo[0]) & (x <= hi[0]))[0]
        rj = np.nonzero((y >= lo[1]) & (y <= hi[1]))[0]
        if ci.size == 0 or rj.size == 0:
            continue
        dist = _segments_distance(x[ci][None, :], y[rj][:, None], a, b)
        block = z[np.ix_(rj, ci)]
        block[dist <= 0.5 * width] += height
        z[np.ix_(rj, ci)] = block
    east = [Boundary("free", np.arange(1, n + 1), np.full(n, n))]
    return Scenario(name, grid, ArrayBed(z, p), Tables(rows=8), 0.04, east, storm, None, duration_s, Numerics())


def reservoir_depth(scen: Scenario, above: float = 5.0) -> np.ndarray:
    """Mean depth of every reservoir cell with its water surface ``above`` metres over the cell's mean ground (the
    level read through the cell table, tables/builder.py:311-331), zero elsewhere."""
    from wadi2d_amd.tables import hbar_at_level

    t = scen.build()
    g = t.grid
    h = hbar_at_level(t.zbar + above, t.zmin, g.from_tiles(t.hc).astype(np.float64),
                      g.from_tiles(t.dtab).astype(np.float64))  # fmt: skip
    return np.where(scen.reservoir, h, 0.0)


def edge_outlets(scen: Scenario, length: int = 128) -> list[Boundary]:
    """Four free segments, one per edge, centred on the edge's lowest cell; needs the built tables (zmin)."""
    t = scen.build()
    n0, n1 = 1, scen.domain  # domain rows and columns of the padded grid
    out = []
    for edge in ("north", "south", "west", "east"):
        if edge in ("north", "south"):
            j = n0 if edge == "north" else n1
            line = t.zmin[j, n0 : n1 + 1]
            c = n0 + int(np.argmin(line))
            ii = np.arange(max(n0, c - length // 2), min(n1 + 1, c + length // 2))
            jj = np.full(ii.size, j)
        else:
            i = n0 if edge == "west" else n1
            line = t.zmin[n0 : n1 + 1, i]
            c = n0 + int(np.argmin(line))
            jj = np.arange(max(n0, c - length // 2), min(n1 + 1, c + length // 2))
            ii = np.full(jj.size, i)
        out.append(Boundary("free", jj, ii))
    # a corner cell named twice would be one cell with two kinds: keep the first
    seen = set()
    for b in out:
        keep = [q for q, cell in enumerate(zip(b.jj.tolist(), b.ii.tolist())) if cell not in seen]
        seen.update(zip(b.jj.tolist(), b.ii.tolist()))
        b.jj, b.ii = b.jj[keep], b.ii[keep]
    return out
