"""Rain from gauges spread over the 2D cells (behaviour of wadi2d/storm/gauges.py:38-135,
wadi2d/storm/gauge_weights.py:32-90 and wadi2d/hydrology/gauge_rain.py:27-60).

Every gauge's cumulative depth is known at the same times (``edges``, minutes), linear between them and level before
the first and after the last. A cell takes the gauges' depths through its centre's weights: under Thiessen its nearest
gauge alone (a tie goes to the gauge listed first), under inverse distance ``d^-p / sum d^-p`` over the gauges that
count (within the search radius; with ``quadrants``, only the nearest in each quadrant about the cell; a cell on a
gauge, or with none in reach, takes its nearest gauge alone).

On the device gauge rain travels the path gridded rain takes (a per-cell index into a short list of depths per window,
``w2d_source_pixels``): under Thiessen the list is the gauges a cell can take, under inverse distance every rained
cell is its own entry, its depth the weighted sum formed on the host. Observed rain has no area reduction.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

THIESSEN = "thiessen"
INVERSE_DISTANCE = "inverse_distance"
QUADRANTS = ((True, True), (False, True), (False, False), (True, False))  # (east of the cell, north of it)
CHUNK = 500_000  # cells weighed at once


def _offsets(x: np.ndarray, y: np.ndarray, gauges: np.ndarray):
    east = gauges[None, :, 0] - x[:, None]
    north = gauges[None, :, 1] - y[:, None]
    return np.hypot(east, north), east, north


def nearest(x, y, gauges: np.ndarray) -> np.ndarray:
    """Index of each point's nearest gauge; the first listed wins a tie."""
    x, y = np.asarray(x, np.float64).ravel(), np.asarray(y, np.float64).ravel()
    out = np.empty(x.size, np.int64)
    for a in range(0, x.size, CHUNK):
        out[a : a + CHUNK] = np.argmin(_offsets(x[a : a + CHUNK], y[a : a + CHUNK], gauges)[0], axis=1)
    return out


def _weights(x, y, gauges, power: float, radius: float | None, quadrants: bool) -> tuple[np.ndarray, int]:
    dist, east, north = _offsets(x, y, gauges)
    counts = np.ones(dist.shape, bool) if radius is None else dist <= radius
    none_in_reach = ~counts.any(axis=1)
    if quadrants:
        picked = np.zeros_like(counts)
        rows = np.arange(x.size)
        for is_east, is_north in QUADRANTS:
            inside = counts & ((east >= 0.0) == is_east) & ((north >= 0.0) == is_north)
            has = inside.any(axis=1)
            closest_here = np.argmin(np.where(inside, dist, np.inf), axis=1)
            picked[rows[has], closest_here[has]] = True
        counts = picked
    closest = np.argmin(dist, axis=1)
    alone = none_in_reach | (dist[np.arange(x.size), closest] == 0.0)
    with np.errstate(divide="ignore"):
        raw = np.where(counts, dist**-power, 0.0)
    raw[alone] = 0.0
    raw[alone, closest[alone]] = 1.0
    return raw / raw.sum(axis=1, keepdims=True), int(np.count_nonzero(none_in_reach))


def inverse_distance(x, y, gauges: np.ndarray, power: float = 2.0, radius: float | None = None,
                     quadrants: bool = True) -> tuple[np.ndarray, int]:  # fmt: skip
    """(weights, points with no gauge in reach): one row of gauge weights per point, summing to 1."""
    x, y = np.asarray(x, np.float64).ravel(), np.asarray(y, np.float64).ravel()
    out = np.empty((x.size, gauges.shape[0]), np.float64)
    lost = 0
    for a in range(0, x.size, CHUNK):
        out[a : a + CHUNK], n = _weights(x[a : a + CHUNK], y[a : a + CHUNK], gauges, power, radius, quadrants)
        lost += n
    return out, lost


def fill_gaps(depths: np.ndarray, gauges: np.ndarray, rule: str, weigh) -> np.ndarray:
    """``depths`` (stretches, gauges), NaN for no value. ``zero``: a gap rains nothing. ``neighbours``: a gauge's gap
    takes what ``weigh(x, y, gauges that have a value)`` reads at the gauge from those gauges, stretch pattern by
    stretch pattern; a stretch that no gauge covers stays dry (gauges.py:59-77)."""
    missing = np.isnan(depths)
    out = np.where(missing, 0.0, depths)
    if rule != "neighbours":
        return out
    for pattern in np.unique(missing, axis=0):
        have = np.flatnonzero(~pattern)
        if not pattern.any() or have.size == 0:
            continue
        rows = np.all(missing == pattern, axis=1)
        for g in np.flatnonzero(pattern):
            w = weigh(gauges[g : g + 1, 0], gauges[g : g + 1, 1], gauges[have])[0]
            out[np.ix_(rows, [g])] = (depths[np.ix_(rows, have)] @ w.T).reshape(-1, 1)
    return out


@dataclass(eq=False)
class GaugeStorm:
    """``cumulative`` (edges, gauges): the depth each gauge has had by each edge (minutes), mm; ``xy`` (gauges, 2) their
    positions in the grid's system, m."""

    names: tuple[str, ...]
    xy: np.ndarray
    edges: np.ndarray
    cumulative: np.ndarray
    method: str = THIESSEN
    power: float = 2.0
    search_radius_m: float | None = None
    quadrants: bool = True

    def __post_init__(self) -> None:
        n = len(self.names)
        self.xy = np.asarray(self.xy, np.float64)
        self.edges = np.asarray(self.edges, np.float64)
        self.cumulative = np.asarray(self.cumulative, np.float64)
        if n == 0:
            raise ValueError("a gauges storm needs one rain gauge at least")
        if self.xy.shape != (n, 2) or self.cumulative.shape != (self.edges.size, n):
            raise ValueError("a gauges storm needs one position and one cumulative curve per gauge")
        if self.method not in (THIESSEN, INVERSE_DISTANCE):
            raise ValueError(f"unknown gauge method {self.method!r}")
        if np.any(np.diff(self.edges) <= 0.0) or np.any(np.diff(self.cumulative, axis=0) < 0.0):
            raise ValueError("gauge times must rise and cumulative depths must not fall")

    @property
    def totals(self) -> np.ndarray:
        return self.cumulative[-1]

    def depths_at(self, minutes: float) -> np.ndarray:
        """Each gauge's depth by this time, mm (gauges.py:124-132)."""
        e = self.edges
        if minutes <= e[0]:
            return self.cumulative[0]
        if minutes >= e[-1]:
            return self.cumulative[-1]
        k = int(np.searchsorted(e, minutes, side="right")) - 1
        share = (minutes - e[k]) / (e[k + 1] - e[k])
        return self.cumulative[k] + (self.cumulative[k + 1] - self.cumulative[k]) * share

    def weights_at(self, x, y) -> tuple[np.ndarray, int]:
        """(weights, points with no gauge in reach) at points under the storm's method."""
        if self.method == THIESSEN:
            index = nearest(x, y, self.xy)
            w = np.zeros((index.size, len(self.names)))
            w[np.arange(index.size), index] = 1.0
            return w, 0
        return inverse_distance(x, y, self.xy, self.power, self.search_radius_m, self.quadrants)


class GaugeCells:
    """The rained cells' gauge rain in the form gridded rain takes (``engine.gridded.CellPixels``): ``used`` the
    entries of the depth list, ``column`` each cell's entry, ``pixel_depths(t)`` the entries' cumulative depths."""

    factor = 1.0  # observed rain: no area reduction

    def __init__(self, storm: GaugeStorm, x: np.ndarray, y: np.ndarray) -> None:
        self.storm = storm
        self.beyond = 0
        if storm.method == THIESSEN:
            self.gauge_of = nearest(x, y, storm.xy)
            self.used, self.column = np.unique(self.gauge_of, return_inverse=True)
            self.weights = None
        else:
            self.weights, self.beyond = storm.weights_at(x, y)
            self.used = np.arange(np.asarray(x).size)
            self.column = self.used.copy()

    def pixel_depths(self, t: float) -> np.ndarray:
        """Each entry's cumulative depth at time t (s), mm (gauge_rain.py:52-57)."""
        depths = self.storm.depths_at(t / 60.0)
        if self.weights is None:
            return np.asarray(depths[self.used], np.float64)
        return np.asarray(self.weights @ depths, np.float64)

    def cell_depths(self, t: float) -> np.ndarray:
        return self.pixel_depths(t)[self.column]
