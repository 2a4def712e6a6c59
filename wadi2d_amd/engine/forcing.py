"""Boundary cells, their open faces, and the rain source of the interior cells (wadi2d/forcing.py).

``boundary_set`` takes boundary cells as (row, column) lists of the padded grid; the reference finds them along
drawn lines with shapely (forcing.py:248-296), which this path does not need: its scenarios name their cells.
Kinds, references, prescribed values and the open-face set follow forcing.py:299-381.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from wadi2d_amd.config import MM_PER_M
from wadi2d_amd.engine.gridded import CellPixels, RainFrames
from wadi2d_amd.engine.losses import cell_excess
from wadi2d_amd.solver.backend import BoundaryCells
from wadi2d_amd.solver.state import state_index
from wadi2d_amd.tables import hbar_at_level

KIND_FIXED, KIND_FREE, KIND_FOLLOW = 0, 1, 2


@dataclass
class Boundary:
    """One boundary: ``kind`` "free", "stage" (level ``stage``) or "normal" (follows its inner cell, ``slope``)."""

    kind: str
    jj: np.ndarray
    ii: np.ndarray
    stage: float = 0.0
    slope: float = 0.001


@dataclass
class BoundarySet:
    """Boundary cells, and the faces where water crosses into them (forcing.py:207-222)."""

    cells: BoundaryCells
    mask: np.ndarray
    xfaces: np.ndarray
    xsign: np.ndarray
    yfaces: np.ndarray
    ysign: np.ndarray

    @property
    def size(self) -> int:
        return int(self.cells.cell.size)


def _one_boundary(b: Boundary, tables, taken: np.ndarray):
    """Kind, follow reference and prescribed values of one boundary's cells (forcing.py:299-332)."""
    grid = tables.grid
    j, i = np.asarray(b.jj, np.int64), np.asarray(b.ii, np.int64)
    n = j.size
    zmin = tables.zmin[j, i]
    kind = np.full(n, KIND_FIXED, np.int32)
    ref = np.array(state_index(grid, j, i), np.int64)
    d0 = np.zeros(n)
    h0 = np.zeros(n)
    if b.kind == "stage":
        hc = grid.from_tiles(tables.hc)[j, i]
        dtab = grid.from_tiles(tables.dtab)[j, i]
        d0 = np.maximum(b.stage - zmin, 0.0)
        h0 = np.asarray(hbar_at_level(b.stage, zmin, hc, dtab))
    elif b.kind == "free":
        kind[:] = KIND_FREE
    else:
        kind[:] = KIND_FOLLOW
        ny, nx = grid.shape
        for m in range(n):
            options = [
                (a, c)
                for a, c in ((j[m] - 1, i[m]), (j[m] + 1, i[m]), (j[m], i[m] - 1), (j[m], i[m] + 1))
                if 0 <= a < ny and 0 <= c < nx and grid.mask[a, c] and not taken[a, c]
            ]
            if not options:
                kind[m] = KIND_FREE
                continue
            a, c = max(options, key=lambda cell: tables.zbar[cell])
            ref[m] = state_index(grid, np.array(a), np.array(c))
            d0[m] = min(tables.zmin[a, c] - zmin[m] - b.slope * grid.dx, 0.0)
    return kind, ref, d0, h0


def open_faces(grid, inner: np.ndarray, taken: np.ndarray):
    """The x and y faces between ``inner`` and ``taken`` cells, with the sign of a flow into ``taken``
    (forcing.py:335-356)."""
    faces = []
    for axis in (1, 0):
        out_low = inner & np.roll(taken, -1, axis=axis)
        j, i = np.nonzero(out_low)
        if axis == 1:
            i = i + 1
        else:
            j = j + 1
        into = inner & np.roll(taken, 1, axis=axis)
        ja, ia = np.nonzero(into)
        index = np.concatenate((state_index(grid, j, i), state_index(grid, ja, ia))).astype(np.int64)
        sign = np.concatenate((np.ones(j.size), -np.ones(ja.size)))
        faces.append((index, sign))
    return faces


def boundary_set(boundaries: list[Boundary], tables) -> BoundarySet:
    """All boundary cells, and the faces between them and the interior with the sign of an outflow
    (forcing.py:359-381)."""
    grid = tables.grid
    taken = np.zeros(grid.shape, bool)
    for b in boundaries:
        taken[np.asarray(b.jj), np.asarray(b.ii)] = True
    parts = []
    for b in boundaries:
        kind, ref, d0, h0 = _one_boundary(b, tables, taken)
        parts.append((np.array(state_index(grid, np.asarray(b.jj), np.asarray(b.ii)), np.int32),
                      ref.astype(np.int32), kind, d0, h0))  # fmt: skip
    if not parts:
        cells = BoundaryCells.empty()
    else:
        cat = [np.concatenate(a) for a in zip(*parts, strict=True)]
        cells = BoundaryCells(cat[0], cat[1], cat[2].astype(np.int32), cat[3].astype(np.float32),
                              cat[4].astype(np.float32))  # fmt: skip
    faces = open_faces(grid, grid.mask & ~taken, taken)
    return BoundarySet(cells, taken, faces[0][0], faces[0][1], faces[1][0], faces[1][1])


class Rain:
    """A storm as a cumulative depth curve over the interior cells, in one of two forms.

    * weight form: the curve times a per-cell weight (1 for a uniform storm). ``rate(t0, t1)`` is the window's mean
      rate per unit weight, ``scale``, times the weight, taken to float32 per cell.
    * excess form (``curve_number``, ``depth_grid``, ``frames`` or ``gauges`` given): the reference's own
      (forcing.py:644-678). A cell's cumulative depth is the curve's times its depth factor (storm.depth_grid; the
      curve alone without one), with gridded rain its pixel's (``frames``, times ``area_factor``), with gauge rain
      its gauge weights times the gauges' depths (``gauges``, an engine.gauges.GaugeStorm), its cumulative excess that depth
      through the SCS loss of its curve number (NaN: none), and its rate the window's difference of excess over
      1000 (t1 - t0), taken to float32.

    Either way the window's rain volume is the FP64 sum of the float32 rates times the cell area and the window.
    """

    def __init__(self, tables, boundary: BoundarySet, cumulative_mm, weight: np.ndarray | None = None,
                 curve_number: np.ndarray | None = None, depth_grid: np.ndarray | None = None,
                 frames: RainFrames | None = None, area_factor: float = 1.0, gauges=None) -> None:  # fmt: skip
        grid = tables.grid
        interior = grid.mask & ~boundary.mask
        self.jj, self.ii = np.nonzero(interior)
        self.state = np.asarray(state_index(grid, self.jj, self.ii), np.int64)
        self.area = grid.dx**2
        self.cumulative_mm = cumulative_mm
        if weight is not None and (curve_number is not None or depth_grid is not None or frames is not None):
            raise ValueError("a weighted storm takes no curve numbers, depth grid or frames: give the excess form")
        if (frames is not None or gauges is not None) and depth_grid is not None:
            raise ValueError("gridded or gauge rain brings its own depths: no depth grid with it")
        if frames is not None and gauges is not None:
            raise ValueError("rain comes from frames or from gauges, not both")
        if weight is not None and gauges is not None:
            raise ValueError("a weighted storm takes no gauges: give the excess form")
        self.pixels = None
        if frames is not None:
            x, y = grid.xy_of(self.jj, self.ii)
            self.pixels = CellPixels(frames, x, y, area_factor)
        if gauges is not None:  # gauge rain through the same per-cell index (engine/gauges.py)
            from wadi2d_amd.engine.gauges import GaugeCells

            x, y = grid.xy_of(self.jj, self.ii)
            self.pixels = GaugeCells(gauges, x, y)
        self.weight = None if weight is None else np.asarray(weight, np.float64)[self.jj, self.ii]
        self.cn = None if curve_number is None else np.asarray(curve_number, np.float64)[self.jj, self.ii]
        self.depths = None if depth_grid is None else np.asarray(depth_grid, np.float64)[self.jj, self.ii]
        if self.cn is not None and np.any(np.isfinite(self.cn) & ~((self.cn > 0.0) & (self.cn <= 100.0))):
            raise ValueError("a curve number lies outside (0, 100]")
        if self.depths is not None and not np.all(np.isfinite(self.depths)):
            raise ValueError("a depth-grid value is missing on a rained cell")

    @property
    def excess_form(self) -> bool:
        return self.cn is not None or self.depths is not None or self.pixels is not None

    def scale(self, t0: float, t1: float) -> float:
        """The window's mean rate per unit weight, m/s (float64)."""
        return (float(self.cumulative_mm(t1)) - float(self.cumulative_mm(t0))) / (MM_PER_M * (t1 - t0))

    def terms(self, t0: float, t1: float) -> tuple[float, float, float]:
        """The excess form's window as the device takes it: the storm depth at both ends (0 with gridded rain, whose
        depths are the pixels'), and 1000 (t1 - t0)."""
        if self.pixels is not None:
            return 0.0, 0.0, MM_PER_M * (t1 - t0)
        return float(self.cumulative_mm(t0)), float(self.cumulative_mm(t1)), MM_PER_M * (t1 - t0)

    def pixel_terms(self, t0: float, t1: float) -> np.ndarray:
        """Gridded rain: the used pixels' cumulative depths at the window's start and end, (2, npix)."""
        return np.stack((self.pixels.pixel_depths(t0), self.pixels.pixel_depths(t1)))

    def excess_mm(self, t: float) -> np.ndarray:
        """Cumulative excess per interior cell at time t (forcing.py:644-665)."""
        if self.pixels is not None:  # cell_rain.py:52-58
            rain = np.array(self.pixels.cell_depths(t), dtype=np.float64)
        else:
            depth = float(self.cumulative_mm(t))
            rain = depth * self.depths if self.depths is not None else np.full(self.state.size, depth)
        if self.cn is None:
            return rain
        return cell_excess(rain, self.cn)

    def rate(self, t0: float, t1: float) -> tuple[np.ndarray, float]:
        step = t1 - t0
        if self.excess_form:  # forcing.py:667-672
            out = ((self.excess_mm(t1) - self.excess_mm(t0)) / (MM_PER_M * step)).astype(np.float32)
        else:
            s = self.scale(t0, t1)
            if self.weight is None:
                out = np.full(self.state.size, np.float32(s), np.float32)
            else:
                out = (self.weight * s).astype(np.float32)
        rain = float(out.sum(dtype=np.float64)) * self.area * step
        return out, rain


def uniform_storm(depth_mm: float, duration_s: float):
    """Cumulative depth of a constant-rate storm, mm (storm/uniform.py:32-35)."""

    def cumulative(t: float) -> float:
        return depth_mm * min(max(t / duration_s, 0.0), 1.0)

    return cumulative


def triangular_storm(peak_mm_h: float, rise_s: float, fall_s: float):
    """Cumulative depth of a triangular hyetograph, 0 -> peak at ``rise_s`` -> 0 at ``rise_s + fall_s``, mm."""
    p = peak_mm_h / 3600.0

    def cumulative(t: float) -> float:
        if t <= 0.0:
            return 0.0
        if t <= rise_s:
            return p * t * t / (2.0 * rise_s)
        up = p * rise_s / 2.0
        if t <= rise_s + fall_s:
            x = t - rise_s
            return up + p * (x - x * x / (2.0 * fall_s))
        return up + p * fall_s / 2.0

    return cumulative
