"""build_tables: bed pixels + grid -> every static array the step reads (wadi2d/tables/builder.py).

The arithmetic runs in the package's host library (``csrc/tables.cpp``, C++ with OpenMP over tiles); this
module holds the reference's data types and the call. Absolute elevations stop here: the step sees only heights
over a cell's lowest point and over a face invert, so float32 keeps its resolution at any datum.
"""

from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

from wadi2d_amd import _native
from wadi2d_amd.config import Tables
from wadi2d_amd.grid import TileGrid
from wadi2d_amd.layout import TILE


@dataclass
class ArrayBed:
    """A bed already in memory, pixels aligned to the grid origin; reads clip to the edge (builder.py:38-59)."""

    z: np.ndarray
    pixel: float
    row_offset: int = 0
    col_offset: int = 0

    def read(self, row0: int, col0: int, nrows: int, ncols: int) -> np.ndarray:
        r = np.clip(np.arange(row0, row0 + nrows) + self.row_offset, 0, self.z.shape[0] - 1)
        c = np.clip(np.arange(col0, col0 + ncols) + self.col_offset, 0, self.z.shape[1] - 1)
        return np.asarray(self.z[np.ix_(r, c)], dtype=np.float64)

    @property
    def span(self) -> int:
        return int(max(self.z.shape))


@dataclass
class CrestFaces:
    """Faces edited by terrain objects, as (ny, nx) grids; NaN level means untouched (builder.py:62-80)."""

    xz: np.ndarray
    xlen: np.ndarray
    xc: np.ndarray
    yz: np.ndarray
    ylen: np.ndarray
    yc: np.ndarray

    @classmethod
    def empty(cls, shape: tuple[int, int]) -> CrestFaces:
        nan = np.full(shape, np.nan)
        one = np.ones(shape)
        zero = np.zeros(shape)
        return cls(nan.copy(), one.copy(), zero.copy(), nan.copy(), one.copy(), zero.copy())


@dataclass
class GridTables:
    """Everything static (builder.py:83-114). Tile arrays are (ntiles, TILE, TILE, ...); x: west faces, y: north."""

    grid: TileGrid
    zmin: np.ndarray
    zbar: np.ndarray
    hc: np.ndarray
    dtab: np.ndarray
    flx: np.ndarray
    frx: np.ndarray
    glx: np.ndarray
    grx: np.ndarray
    efx: np.ndarray
    cwx: np.ndarray
    ftabx: np.ndarray
    fly: np.ndarray
    fry: np.ndarray
    gly: np.ndarray
    gry: np.ndarray
    efy: np.ndarray
    cwy: np.ndarray
    ftaby: np.ndarray
    zcx: np.ndarray
    zcy: np.ndarray

    @property
    def rows(self) -> int:
        return self.dtab.shape[-1]


class _Config(ctypes.Structure):
    _fields_ = [
        ("rows", ctypes.c_int32),
        ("local_radius", ctypes.c_int32),
        ("e_lo", ctypes.c_double),
        ("e_hi", ctypes.c_double),
        ("beta", ctypes.c_double),
        ("crest_tol", ctypes.c_double),
    ]


_OUT_NAMES = ("zmin", "zbar", "hc", "dtab", "flx", "frx", "glx", "grx", "efx", "cwx", "ftabx",
              "fly", "fry", "gly", "gry", "efy", "cwy", "ftaby", "zcx", "zcy")  # fmt: skip


class _Out(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in _OUT_NAMES]


def _ptr(a: np.ndarray | None) -> ctypes.c_void_p | None:
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def table_workers(workers: int | None = None) -> int:
    """Threads for the build: all cores unless told (tables/parallel.py:17-67)."""
    if workers:
        return int(workers)
    return int(os.environ.get("WADI2D_TABLE_WORKERS", os.cpu_count() or 1))


def build_tables(
    grid: TileGrid,
    bed: ArrayBed,
    cfg: Tables | None = None,
    manning: float | np.ndarray = 0.035,
    crest: CrestFaces | None = None,
    workers: int | None = None,
    device: bool = False,
) -> GridTables:
    """All tables of a uniform grid (builder.py:131-258). ``manning`` is one value or a (ny, nx) grid.

    ``device`` builds them on the GPU (w2d_build_tables_device: the same rows, one thread per cell; a row may
    differ from the host build by one float32 ulp, the device's cbrt and pow not being the host library's)."""
    cfg = cfg or Tables()
    k = round(grid.dx / bed.pixel)
    if abs(k * bed.pixel - grid.dx) > 1e-6 * grid.dx:
        raise ValueError(f"cell size {grid.dx} is not a multiple of the pixel {bed.pixel}")
    ny, nx = grid.shape
    nt, rows = grid.ntiles, cfg.rows
    z = np.ascontiguousarray(bed.z, dtype=np.float64)
    nman_grid = not np.isscalar(manning)
    nman = np.ascontiguousarray(np.broadcast_to(manning, (ny, nx)) if nman_grid else np.array([manning]),
                                dtype=np.float64)  # fmt: skip
    mask = np.ascontiguousarray(grid.mask, dtype=np.uint8)
    tiles = np.ascontiguousarray(grid.tiles, dtype=np.int32)
    crest_arrays = [None] * 6
    if crest is not None:
        crest_arrays = [np.ascontiguousarray(a, dtype=np.float64) for a in
                        (crest.xz, crest.xlen, crest.xc, crest.yz, crest.ylen, crest.yc)]  # fmt: skip
    out = {
        "zmin": np.zeros((ny, nx)),
        "zbar": np.zeros((ny, nx)),
        "hc": np.zeros((nt, TILE, TILE), np.float32),
        "dtab": np.zeros((nt, TILE, TILE, rows), np.float32),
        "ftabx": np.zeros((nt, TILE, TILE, rows, 2), np.float32),
        "ftaby": np.zeros((nt, TILE, TILE, rows, 2), np.float32),
    }
    for n in ("flx", "frx", "glx", "grx", "efx", "cwx", "fly", "fry", "gly", "gry", "efy", "cwy", "zcx", "zcy"):
        out[n] = np.zeros((nt, TILE, TILE), np.float32)
    c = _Config(rows, int(cfg.face_conveyance == "local_radius"), cfg.face_depth_min, cfg.face_depth_max,
                cfg.local_radius_beta, cfg.crest_tolerance)  # fmt: skip
    o = _Out(*[out[n].ctypes.data for n in _OUT_NAMES])
    I64, I32, DBL, VP = ctypes.c_int64, ctypes.c_int32, ctypes.c_double, ctypes.c_void_p
    argtypes = [VP, I64, I64, I64, I64, DBL, I32, I32, DBL, VP, I32, VP, VP, I32,
                VP, VP, VP, VP, VP, VP, ctypes.POINTER(_Config)]  # fmt: skip
    args = [
        _ptr(z), z.shape[0], z.shape[1], bed.row_offset, bed.col_offset, bed.pixel, ny, nx, grid.dx,
        _ptr(tiles), nt, _ptr(mask), _ptr(nman), int(nman_grid), *[_ptr(a) for a in crest_arrays],
        ctypes.byref(c),
    ]  # fmt: skip
    if device:
        from wadi2d_amd.solver import hip_backend

        hip_backend.check(hip_backend.api().w2d_build_tables_device(*args, ctypes.byref(o)))
        return GridTables(grid, **out)
    fn = _native.host().w2d_build_tables
    fn.restype = ctypes.c_int
    fn.argtypes = [*argtypes, I32, ctypes.POINTER(_Out)]
    status = fn(*args, table_workers(workers), ctypes.byref(o))
    if status != 0:
        raise ValueError("table build refused its arguments")
    return GridTables(grid, **out)


def hbar_at_level(level, zmin, hc, dtab) -> np.ndarray:
    """Mean depth that gives exactly this level through the table, the inverse of the step's lookup
    (builder.py:311-331)."""
    rows = dtab.shape[-1]
    last = rows - 1
    d = np.maximum(np.asarray(level, dtype=np.float64) - zmin, 0.0)
    out = np.where(hc <= 0.0, d, 0.0)
    top = dtab[..., last]
    above = (hc > 0.0) & (d >= top)
    out = np.where(above, hc + (d - top), out)
    inside = (hc > 0.0) & (d < top) & (d > 0.0)
    if inside.any():
        di = d[inside]
        tab = dtab[inside]
        r = np.clip(np.sum(tab <= di[:, None], axis=1) - 1, 0, last - 1)
        idx = np.arange(tab.shape[0])
        lo, hi = tab[idx, r], tab[idx, r + 1]
        s = r + (di - lo) / np.maximum(hi - lo, 1e-30)
        out[inside] = hc[inside] * (s / last) ** 2
    return out
