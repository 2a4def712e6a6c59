"""Cells and faces under drawn lines and polygons, on the uniform grid (wadi2d/forcing.py:227-296,411-423;
wadi2d/terrain/crest.py:58-153), without shapely.

* ``cells_along``: the active cells a polyline passes over, sampled ``samples`` times per cell length; a sample off
  the domain takes an active cell next to it, else the nearest active cells within ``snap_m`` (with the staircase
  behind the nearest, 1.5 cells deep), as the reference opens a boundary drawn along the domain's edge.
* ``inside``: cell centres inside a polygon ring (even-odd rule; a centre on an edge counts as the reference's
  ``contains_xy`` would not, which matters only for centres exactly on a drawn edge).
* ``line_faces``: the lattice faces a polyline follows, each vertex snapped to its nearest cell corner and the
  corners between joined by the path nearest the straight segment; a discharge through them is positive to the
  line's right.
"""

from __future__ import annotations

import numpy as np

from wadi2d_amd.grid import TileGrid
from wadi2d_amd.solver.state import state_index

SNAP_DEPTH = 1.5
"""Cells behind the nearest one that a snapped sample also opens (forcing.py:57)."""

NORTH, WEST = 0, 1
"""Axis of a lattice edge: a cell's north face (a y face) or its west face (an x face)."""


def _points_along(coordinates, step: float) -> np.ndarray:
    pts = np.asarray(coordinates, dtype=np.float64)
    if pts.ndim != 2 or pts.shape[0] < 2:
        raise ValueError("a line needs at least two points")
    seg = np.hypot(*np.diff(pts, axis=0).T)
    at = np.concatenate(([0.0], np.cumsum(seg)))
    n = max(int(np.ceil(at[-1] / step)), 2)
    s = np.linspace(0.0, at[-1], n)
    return np.stack((np.interp(s, at, pts[:, 0]), np.interp(s, at, pts[:, 1])), axis=1)


def _nearest_active(mask: np.ndarray, j: int, i: int, reach: int) -> list[tuple[int, int]]:
    ny, nx = mask.shape
    top, left = max(j - reach, 0), max(i - reach, 0)
    bottom, right = min(j + reach + 1, ny), min(i + reach + 1, nx)
    if top >= bottom or left >= right:
        return []
    rows, cols = np.nonzero(mask[top:bottom, left:right])
    if rows.size == 0:
        return []
    rows, cols = rows + top, cols + left
    gap = np.hypot(rows - j, cols - i)
    nearest = gap.min()
    if nearest > reach:
        return []
    keep = gap <= nearest + SNAP_DEPTH
    return list(zip(rows[keep].tolist(), cols[keep].tolist(), strict=True))


def cells_along(grid: TileGrid, coordinates, samples: int = 4, snap_m: float = 0.0) -> tuple[np.ndarray, np.ndarray]:
    """Rows and columns (ascending) of the active cells under a polyline."""
    pts = _points_along(coordinates, grid.dx / samples)
    jj = np.floor((grid.y0 - pts[:, 1]) / grid.dx).astype(np.int64)
    ii = np.floor((pts[:, 0] - grid.x0) / grid.dx).astype(np.int64)
    ny, nx = grid.shape
    reach = int(snap_m / grid.dx)
    found: set[tuple[int, int]] = set()
    order = ((0, 0), (0, 1), (0, -1), (1, 0), (-1, 0), (1, 1), (1, -1), (-1, 1), (-1, -1))
    for j, i in zip(jj.tolist(), ii.tolist(), strict=True):
        for dj, di in order:
            a, b = j + dj, i + di
            if 0 <= a < ny and 0 <= b < nx and grid.mask[a, b]:
                found.add((a, b))
                break
        else:
            found.update(_nearest_active(grid.mask, j, i, reach))
    if not found:
        raise ValueError(f"line from {tuple(coordinates[0])} does not touch the domain")
    cells = np.array(sorted(found), np.int64)
    return cells[:, 0], cells[:, 1]


def inside(grid: TileGrid, ring) -> np.ndarray:
    """(ny, nx) mask of the active cells whose centre lies inside the polygon ``ring`` (even-odd rule)."""
    poly = np.asarray(ring, dtype=np.float64)
    if poly.ndim != 2 or poly.shape[0] < 3:
        raise ValueError("a polygon needs at least three points")
    jj, ii = np.nonzero(grid.mask)
    x, y = grid.xy_of(jj, ii)
    hit = np.zeros(x.size, bool)
    xa, ya = poly[:, 0], poly[:, 1]
    xb, yb = np.roll(xa, -1), np.roll(ya, -1)
    for x1, y1, x2, y2 in zip(xa, ya, xb, yb, strict=True):
        if y1 == y2:
            continue
        crosses = (y1 > y) != (y2 > y)
        xc = x1 + (y - y1) * (x2 - x1) / (y2 - y1)
        hit ^= crosses & (x < xc)
    out = np.zeros(grid.shape, bool)
    out[jj[hit], ii[hit]] = True
    return out


def _corner(grid: TileGrid, x: float, y: float) -> tuple[int, int]:
    return int(round((grid.y0 - y) / grid.dx)), int(round((x - grid.x0) / grid.dx))


def _path(a: tuple[int, int], b: tuple[int, int]) -> list[tuple[int, int]]:
    """Corners after ``a`` up to ``b``, one lattice step each, staying nearest the straight segment: a step across
    (in i) is taken while its midpoint lies before the next step down (in j) along the segment."""
    (cj, ci), (dj, di) = a, (b[0] - a[0], b[1] - a[1])
    nj, ni, sj, si = abs(dj), abs(di), (dj > 0) - (dj < 0), (di > 0) - (di < 0)
    taken_i = taken_j = 0
    out = []
    while taken_i < ni or taken_j < nj:
        if taken_i < ni and (taken_j >= nj or (2 * taken_i + 1) * nj < (2 * taken_j + 1) * ni):
            ci += si
            taken_i += 1
        else:
            cj += sj
            taken_j += 1
        out.append((cj, ci))
    return out


def line_faces(grid: TileGrid, coordinates):
    """State indices and signs of the x faces and y faces under a polyline; sign * q is the discharge to the
    line's right, m3/s. Faces outside the grid or of inactive cells are dropped."""
    pts = np.asarray(coordinates, dtype=np.float64)
    corners = [_corner(grid, *pts[0])]
    for q in pts[1:]:
        for c in _path(corners[-1], _corner(grid, *q)):
            if len(corners) > 1 and c == corners[-2]:
                corners.pop()  # a step straight back cancels the one before
            else:
                corners.append(c)
    ny, nx = grid.shape
    xs, ys = [], []
    for (aj, ai), (bj, bi) in zip(corners[:-1], corners[1:], strict=True):
        if aj == bj:  # along a row of corners: the north face of the cell below the edge
            j, i, side = aj, (ai if bi > ai else ai - 1), (1 if bi > ai else -1)
            if 0 <= j < ny and 0 <= i < nx and grid.mask[j, i]:
                ys.append((j, i, side))
        else:  # down or up a column of corners: the west face of the cell right of the edge
            j, i, side = (aj if bj > aj else aj - 1), ai, (-1 if bj > aj else 1)
            if 0 <= j < ny and 0 <= i < nx and grid.mask[j, i]:
                xs.append((j, i, side))

    def index(rows):
        if not rows:
            return np.zeros(0, np.int64), np.zeros(0)
        a = np.array(rows, np.int64)
        return np.asarray(state_index(grid, a[:, 0], a[:, 1]), np.int64), a[:, 2].astype(np.float64)

    xi, xsgn = index(xs)
    yi, ysgn = index(ys)
    return xi, xsgn, yi, ysgn
