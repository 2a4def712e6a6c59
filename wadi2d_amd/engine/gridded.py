"""Rain that varies by cell and by time (wadi2d/storm/rain_frames.py:74-185, storm/gridded.py:73-101,
hydrology/cell_rain.py:19-58): frames of rain depth on a pixel grid; each 2D cell takes the pixel its centre falls in,
and a pixel's rain accumulates linearly through each frame's period and stays level across a gap.

The reference reads NetCDF, DSS and GeoTIFF stacks (io/rain/); none of their libraries is in this image, so frames
come as ``.npz``: ``depths`` (frames, rows, columns) in mm per frame, NaN for no data; ``starts_min`` and
``ends_min`` (one per frame); ``west``, ``north`` (the grid's outer corner), ``dx``, ``dy`` (pixel size, m).
"""

from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path

import numpy as np


@dataclass(frozen=True, eq=False)
class RainFrames:
    """``depths`` is (frames, rows, columns) in mm; ``west`` and ``north`` are the outer corner of the grid
    (rain_frames.py:87-117)."""

    depths: np.ndarray
    starts_min: np.ndarray
    ends_min: np.ndarray
    west: float
    north: float
    dx: float
    dy: float

    def __post_init__(self) -> None:
        if self.depths.ndim != 3 or self.depths.shape[0] == 0:
            raise ValueError("rain frames must be (frames, rows, columns), with at least one frame")
        if self.starts_min.size != self.depths.shape[0] or self.ends_min.size != self.depths.shape[0]:
            raise ValueError("rain frames need one start and one end per frame")
        if np.any(self.ends_min <= self.starts_min) or np.any(self.starts_min[1:] < self.ends_min[:-1]):
            raise ValueError("rain periods must run forward and must not overlap")
        if self.dx <= 0.0 or self.dy <= 0.0:
            raise ValueError("rain pixels must have a size")

    @property
    def count(self) -> int:
        return self.depths.shape[0]

    @property
    def columns(self) -> int:
        return self.depths.shape[2]

    @property
    def rows(self) -> int:
        return self.depths.shape[1]

    def pixel_of(self, x: np.ndarray, y: np.ndarray) -> np.ndarray:
        """Flat index of the pixel each point falls in, -1 off the grid (rain_frames.py:119-125)."""
        column = np.floor((np.asarray(x) - self.west) / self.dx)
        row = np.floor((self.north - np.asarray(y)) / self.dy)
        inside = (column >= 0) & (column < self.columns) & (row >= 0) & (row < self.rows)
        flat = row * self.columns + column
        return np.where(inside, np.nan_to_num(flat), -1.0).astype(np.int64)

    def totals(self) -> np.ndarray:
        """Depth over the whole event per pixel; NaN where a pixel never had data (rain_frames.py:127-130)."""
        seen = np.isfinite(self.depths).any(axis=0)
        return np.where(seen, np.nansum(self.depths, axis=0), np.nan)


class PixelSeries:
    """Cumulative rain of some pixels against time, linear inside a period and level across a gap between periods
    (rain_frames.py:167-185)."""

    def __init__(self, frames: RainFrames, pixels: np.ndarray) -> None:
        self.edges = np.unique(np.concatenate((frames.starts_min, frames.ends_min)))
        steps = np.zeros((self.edges.size - 1, pixels.size))
        at = np.searchsorted(self.edges, frames.starts_min)
        steps[at] = np.nan_to_num(frames.depths.reshape(frames.count, -1)[:, pixels])
        self.cumulative = np.vstack((np.zeros(pixels.size), np.cumsum(steps, axis=0)))

    def at(self, minutes: float) -> np.ndarray:
        """Depth fallen by this time, one value per pixel."""
        if minutes <= self.edges[0]:
            return self.cumulative[0]
        if minutes >= self.edges[-1]:
            return self.cumulative[-1]
        k = int(np.searchsorted(self.edges, minutes, side="right")) - 1
        share = (minutes - self.edges[k]) / (self.edges[k + 1] - self.edges[k])
        return self.cumulative[k] + (self.cumulative[k + 1] - self.cumulative[k]) * share


def read_frames(path: str | Path) -> RainFrames:
    """Rain frames from an ``.npz`` file (the keys in this module's docstring)."""
    p = Path(path)
    if not p.is_file():
        raise FileNotFoundError(f"rain file not found: {p}")
    if p.suffix.lower() != ".npz":
        raise ValueError(f"{p.name}: this path reads rain frames as .npz, not '{p.suffix}' (no NetCDF/DSS/GDAL here)")
    with np.load(p, allow_pickle=False) as f:
        for key in ("depths", "starts_min", "ends_min", "west", "north", "dx", "dy"):
            if key not in f:
                raise ValueError(f"{p.name}: no '{key}' array")
        return RainFrames(np.asarray(f["depths"], np.float64), np.asarray(f["starts_min"], np.float64).ravel(),
                          np.asarray(f["ends_min"], np.float64).ravel(), float(f["west"]), float(f["north"]),
                          float(f["dx"]), float(f["dy"]))  # fmt: skip


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
