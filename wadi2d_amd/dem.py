"""Ground rasters for a run: what ``terrain.dem`` names, read into a bed and the uniform grid over it.

The reference reads a GeoTIFF with rasterio (terrain/raster.py) and snaps the grid origin to pixel corners, dx a
multiple of the pixel (schema.py:48-52). Two formats are read here, without GDAL:

* ESRI ASCII grid (``.asc``): the header keys ``ncols nrows xllcorner|xllcenter yllcorner|yllcenter cellsize
  [NODATA_value]``, then rows north to south;
* NumPy (``.npz``): ``z`` (rows north to south), ``x0`` and ``y0`` (the north-west corner) and ``pixel``; NaN for
  no data.

Cells with any no-data pixel under them are outside the domain, as the reference's mask of valid ground.
"""

from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path

import numpy as np

from wadi2d_amd.grid import TileGrid
from wadi2d_amd.tables import ArrayBed


@dataclass
class Raster:
    """Ground pixels, rows north to south; (x0, y0) is the north-west corner; NaN where there is no data."""

    z: np.ndarray
    x0: float
    y0: float
    pixel: float


def _read_asc(path: Path) -> Raster:
    header: dict[str, float] = {}
    with path.open("r", encoding="ascii") as f:
        while len(header) < 6:
            pos = f.tell()
            line = f.readline()
            if not line:
                break
            parts = line.split()
            if len(parts) != 2 or not parts[0][0].isalpha():
                f.seek(pos)
                break
            header[parts[0].lower()] = float(parts[1])
        data = np.loadtxt(f, dtype=np.float64, ndmin=2)
    for key in ("ncols", "nrows", "cellsize"):
        if key not in header:
            raise ValueError(f"{path.name}: ESRI ASCII header has no {key}")
    ncols, nrows, p = int(header["ncols"]), int(header["nrows"]), header["cellsize"]
    if data.shape != (nrows, ncols):
        raise ValueError(f"{path.name}: {data.shape[0]} x {data.shape[1]} values for a {nrows} x {ncols} header")
    if "xllcorner" in header:
        x0 = header["xllcorner"]
    elif "xllcenter" in header:
        x0 = header["xllcenter"] - 0.5 * p
    else:
        raise ValueError(f"{path.name}: ESRI ASCII header has no xllcorner or xllcenter")
    if "yllcorner" in header:
        yll = header["yllcorner"]
    elif "yllcenter" in header:
        yll = header["yllcenter"] - 0.5 * p
    else:
        raise ValueError(f"{path.name}: ESRI ASCII header has no yllcorner or yllcenter")
    if "nodata_value" in header:
        data[data == header["nodata_value"]] = np.nan
    return Raster(data, x0, yll + nrows * p, p)


def _read_npz(path: Path) -> Raster:
    with np.load(path, allow_pickle=False) as f:
        for key in ("z", "x0", "y0", "pixel"):
            if key not in f:
                raise ValueError(f"{path.name}: no '{key}' array")
        return Raster(np.asarray(f["z"], np.float64), float(f["x0"]), float(f["y0"]), float(f["pixel"]))


def read_raster(path: str | Path, stride: int = 1) -> Raster:
    """The ground of ``path``; ``stride`` reads every n-th pixel (terrain.pixel_stride)."""
    p = Path(path)
    if not p.is_file():
        raise FileNotFoundError(f"raster not found: {p}")
    suffix = p.suffix.lower()
    if suffix == ".asc":
        r = _read_asc(p)
    elif suffix == ".npz":
        r = _read_npz(p)
    else:
        raise ValueError(f"{p.name}: this path reads .asc (ESRI ASCII grid) or .npz ground, not '{suffix}'")
    if r.z.ndim != 2 or r.z.size == 0 or not r.pixel > 0.0:
        raise ValueError(f"{p.name}: not a 2-D raster with a positive pixel")
    if stride > 1:
        r = Raster(r.z[::stride, ::stride].copy(), r.x0, r.y0, r.pixel * stride)
    return r


def grid_over(r: Raster, dx: float, extent: tuple[float, float, float, float] | None = None):
    """The tile grid at ``dx`` over the raster (or the part of it inside ``extent``, snapped inward to pixel
    corners), its bed, and the domain mask; a cell is in the domain when every pixel under it has ground."""
    k = round(dx / r.pixel)
    if k < 1 or abs(k * r.pixel - dx) > 1e-6 * dx:
        raise ValueError(f"grid.dx {dx} is not a multiple of the DEM pixel {r.pixel}")
    ny_pix, nx_pix = r.z.shape
    c0, r0 = 0, 0
    c1, r1 = nx_pix, ny_pix
    if extent is not None:
        xmin, ymin, xmax, ymax = extent
        if not (xmax > xmin and ymax > ymin):
            raise ValueError("grid.extent must be xmin, ymin, xmax, ymax with xmax > xmin and ymax > ymin")
        c0 = max(int(np.ceil((xmin - r.x0) / r.pixel - 1e-9)), 0)
        c1 = min(int(np.floor((xmax - r.x0) / r.pixel + 1e-9)), nx_pix)
        r0 = max(int(np.ceil((r.y0 - ymax) / r.pixel - 1e-9)), 0)
        r1 = min(int(np.floor((r.y0 - ymin) / r.pixel + 1e-9)), ny_pix)
    ny, nx = (r1 - r0) // k, (c1 - c0) // k
    if ny < 1 or nx < 1:
        raise ValueError("the grid extent holds no whole cell of the DEM")
    sub = r.z[r0 : r0 + ny * k, c0 : c0 + nx * k]
    valid = np.isfinite(sub).reshape(ny, k, nx, k).all(axis=(1, 3))
    if not valid.any():
        raise ValueError("no cell of the grid lies on ground with data")
    x0, y0 = r.x0 + c0 * r.pixel, r.y0 - r0 * r.pixel
    grid = TileGrid.from_mask(valid, x0, y0, dx)
    # pixels of the padded grid: one cell of margin (k pixels) before the domain; reads clip to the raster edge
    z = np.where(np.isfinite(r.z), r.z, np.nan)
    fill = np.nanmax(z) if np.isfinite(z).any() else 0.0
    z = np.where(np.isfinite(z), z, fill)
    bed = ArrayBed(z, r.pixel, row_offset=r0 - k, col_offset=c0 - k)
    return grid, bed


def raster_at_cells(path: str | Path, grid) -> np.ndarray:
    """A raster read at the centres of the grid's cells (terrain/sampling.py:28-70): the value of the pixel holding
    each mask cell's centre, NaN off the raster, on no-data and off the mask. The raster must share the grid's
    coordinate system (this path reads no CRS)."""
    r = read_raster(path)
    out = np.full(grid.shape, np.nan)
    jj, ii = np.nonzero(grid.mask)
    if jj.size == 0:
        return out
    x, y = grid.xy_of(jj, ii)
    cols = np.floor((x - r.x0) / r.pixel).astype(np.int64)
    rows = np.floor((r.y0 - y) / r.pixel).astype(np.int64)
    on = (cols >= 0) & (cols < r.z.shape[1]) & (rows >= 0) & (rows < r.z.shape[0])
    out[jj[on], ii[on]] = np.asarray(r.z, np.float64)[rows[on], cols[on]]
    return out
