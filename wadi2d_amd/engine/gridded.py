"""Rain that varies over the ground and in time: a stack of rain grids, one grid per period, sampled per 2D cell.

Behaviour followed: wadi2d/storm/rain_frames.py:74-185 (the frames, a point's pixel, per-pixel totals and the
per-pixel cumulative curve), storm/gridded.py:73-101 (the area reduction) and hydrology/cell_rain.py:19-58 (each cell
takes the pixel its centre falls in). A pixel's cumulative depth rises at a constant rate through each period and
holds still through any gap between periods; a NaN in a grid is "no data" and adds nothing.

Files: the reference reads NetCDF, DSS and GeoTIFF stacks through libraries this image does not have, so here a rain
file is an ``.npz`` holding ``depths`` (periods, rows, columns; mm fallen in each period, NaN for no data),
``starts_min`` and ``ends_min`` (each period's bounds, minutes), ``west`` and ``north`` (the outer corner of pixel
(0, 0)) and ``dx``, ``dy`` (pixel width and height, m). Row 0 is the northern row, column 0 the western column.
"""

from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path

import numpy as np

NPZ_KEYS = ("depths", "starts_min", "ends_min", "west", "north", "dx", "dy")


def _problem(depths: np.ndarray, starts: np.ndarray, ends: np.ndarray, dx: float, dy: float) -> str | None:
    """What is wrong with a stack of rain grids, or None when nothing is."""
    if depths.ndim != 3:
        return f"rain depths need three axes (period, row, column), not {depths.ndim}"
    n = depths.shape[0]
    if n == 0:
        return "rain holds no period"
    if starts.size != n or ends.size != n:
        return f"{n} rain periods but {starts.size} starts and {ends.size} ends"
    if not np.all(ends > starts):
        return "a rain period ends at or before its start"
    if not np.all(starts[1:] >= ends[:-1]):
        return "rain periods must come in time order without overlapping"
    if not (dx > 0.0 and dy > 0.0):
        return f"rain pixels need a positive size, not {dx} x {dy}"
    return None


@dataclass(frozen=True, eq=False)
class RainFrames:
    """A stack of rain grids (the fields of rain_frames.py:87-98): ``depths`` (periods, rows, columns) in mm per period
    over [``starts_min``, ``ends_min``); pixel (0, 0)'s outer corner at (``west``, ``north``); pixels ``dx`` by ``dy``."""

    depths: np.ndarray
    starts_min: np.ndarray
    ends_min: np.ndarray
    west: float
    north: float
    dx: float
    dy: float

    def __post_init__(self) -> None:
        why = _problem(self.depths, self.starts_min, self.ends_min, self.dx, self.dy)
        if why:
            raise ValueError(why)

    @property
    def count(self) -> int:
        """Number of periods."""
        return int(self.depths.shape[0])

    @property
    def rows(self) -> int:
        return int(self.depths.shape[1])

    @property
    def columns(self) -> int:
        return int(self.depths.shape[2])

    def pixel_of(self, x, y) -> np.ndarray:
        """Row-major pixel number under each point (x, y); -1 for a point beyond the grid's edges
        (rain_frames.py:119-125)."""
        fx = (np.asarray(x, np.float64) - self.west) / self.dx
        fy = (self.north - np.asarray(y, np.float64)) / self.dy
        ok = (fx >= 0.0) & (fx < self.columns) & (fy >= 0.0) & (fy < self.rows)  # False for NaN too
        col = np.floor(np.where(ok, fx, 0.0)).astype(np.int64)
        row = np.floor(np.where(ok, fy, 0.0)).astype(np.int64)
        return np.where(ok, row * self.columns + col, -1)

    def totals(self) -> np.ndarray:
        """(rows, columns): each pixel's depth summed over every period it has data for; NaN for a pixel with data in
        no period (rain_frames.py:127-130)."""
        has = ~np.isnan(self.depths)
        total = np.where(has, self.depths, 0.0).sum(axis=0)
        return np.where(has.any(axis=0), total, np.nan)


class PixelSeries:
    """The cumulative depth of chosen pixels as a piecewise-linear function of time (rain_frames.py:167-185).

    The curve runs through two knots per period: at its start the depth of the periods before it, at its end that
    plus its own. Between one period's end and the next one's start both knots hold the same depth, so the curve is
    level there; before the first start it is 0, after the last end the event's total.
    """

    def __init__(self, frames: RainFrames, pixels: np.ndarray) -> None:
        per_period = np.nan_to_num(frames.depths.reshape(frames.count, -1)[:, np.asarray(pixels)])
        after = np.cumsum(per_period, axis=0)  # depth once each period has ended
        before = np.concatenate((np.zeros((1, after.shape[1])), after[:-1]))
        self.times = np.column_stack((frames.starts_min, frames.ends_min)).ravel()  # start 0, end 0, start 1, ...
        self.depths = np.stack((before, after), axis=1).reshape(2 * frames.count, -1)

    def at(self, minutes: float) -> np.ndarray:
        """Depth fallen by ``minutes``, one value per chosen pixel."""
        t = self.times
        if minutes <= t[0]:
            return self.depths[0]
        if minutes >= t[-1]:
            return self.depths[-1]
        k = int(np.searchsorted(t, minutes, side="right")) - 1  # t[k] <= minutes < t[k + 1]
        lo, hi = self.depths[k], self.depths[k + 1]
        return lo + (hi - lo) * ((minutes - t[k]) / (t[k + 1] - t[k]))


def read_frames(path: str | Path) -> RainFrames:
    """A stack of rain grids from an ``.npz`` file holding the arrays named in ``NPZ_KEYS``."""
    p = Path(path)
    if not p.is_file():
        raise FileNotFoundError(f"no rain file at {p}")
    if p.suffix.lower() != ".npz":
        raise ValueError(f"{p.name}: rain is read from .npz files only (NetCDF, DSS and GeoTIFF readers are absent)")
    with np.load(p, allow_pickle=False) as f:
        absent = [k for k in NPZ_KEYS if k not in f]
        if absent:
            raise ValueError(f"{p.name} lacks {', '.join(absent)}")
        depths, starts, ends = (np.asarray(f[k], np.float64) for k in NPZ_KEYS[:3])
        west, north, dx, dy = (float(f[k]) for k in NPZ_KEYS[3:])
    return RainFrames(depths, starts.ravel(), ends.ravel(), west, north, dx, dy)


class CellPixels:
    """Each rained cell's pixel (cell_rain.py:19-44 with forcing.py:571-578), as a column of a ``PixelSeries`` over
    the pixels used; a rained cell off the grid or on a pixel that never had data is refused."""

    def __init__(self, frames: RainFrames, x: np.ndarray, y: np.ndarray, factor: float = 1.0) -> None:
        pixel = frames.pixel_of(x, y)
        known = pixel >= 0
        known[known] = np.isfinite(frames.totals().ravel()[pixel[known]])
        holes = int(np.count_nonzero(~known))
        if holes:
            raise ValueError(
                f"the rain grid gives nothing on {holes} rained cells of {pixel.size}: it does not cover the domain"
            )
        self.used, self.column = np.unique(pixel, return_inverse=True)
        self.series = PixelSeries(frames, self.used)
        self.factor = float(factor)

    def reduced(self, depth_mm):
        """Depths after the area reduction; untouched when there is none (gridded.py:97-100)."""
        return depth_mm if self.factor == 1.0 else depth_mm * self.factor

    def pixel_depths(self, t: float) -> np.ndarray:
        """Each used pixel's (reduced) cumulative depth at time t (s), mm."""
        return self.reduced(self.series.at(t / 60.0))

    def cell_depths(self, t: float) -> np.ndarray:
        """Each rained cell's cumulative depth at time t (s), mm."""
        return self.pixel_depths(t)[self.column]
