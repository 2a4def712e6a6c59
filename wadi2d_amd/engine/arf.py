"""Area reduction of a design storm, and the depth-area study that runs one storm at many storm areas.

Behaviour followed: the factor table and the areas it places (wadi2d/storm/arf.py:13-98), the study's areas, its
reading of the runs' peaks at a design area and its composite maps (wadi2d/arf.py:79-247), and the study itself
(wadi2d/arf.py:328-388), which runs one full run per storm area, one after another. Here every storm area is a member
of one device-resident batched launch over copies of the same grid (engine/ensemble.py): the members differ only in
their storm's factor. Files, the job schema and the lumped network's hand-backs stay with the caller.
"""

from __future__ import annotations

import dataclasses
import time
from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True, eq=False)
class ArfTable:
    """A table of (storm area km2, factor), read linearly in area; the end factors hold beyond it, except that below
    the first area the depth stays whole when ``reduces_below_table`` is False (arf.py:13-36)."""

    areas_km2: np.ndarray
    factors: np.ndarray
    reduces_below_table: bool = True

    def __post_init__(self) -> None:
        a, f = np.asarray(self.areas_km2, np.float64), np.asarray(self.factors, np.float64)
        if a.ndim != 1 or a.size < 1 or f.shape != a.shape:
            raise ValueError("an ARF table needs one factor per area")
        if np.any(np.diff(a) <= 0.0):
            raise ValueError("ARF table areas must increase")
        object.__setattr__(self, "areas_km2", a)
        object.__setattr__(self, "factors", f)

    def factor(self, area_km2: float) -> float:
        """The factor at one storm area."""
        if not self.reduces_below_table and area_km2 < self.areas_km2[0]:
            return 1.0
        return float(np.interp(area_km2, self.areas_km2, self.factors))

    def step_areas(self, low_km2: float, high_km2: float, largest_change: float) -> list[float]:
        """Storm areas from ``low_km2`` to ``high_km2`` (both kept) whose factors, read as ``factor`` reads them, change
        by at most ``largest_change`` between neighbours, with as few areas as that allows (arf.py:38-68).

        The factor's whole path over the span (every rise and fall, and the jump to the first factor when the depth is
        whole below the table) is cut into equal parts no larger than the limit; each cut is placed by linear reading
        between the knots of that path.
        """
        if largest_change <= 0.0:
            raise ValueError("the largest change of the area-reduction factor must be above zero")
        if high_km2 <= low_km2 * (1.0 + 1e-9):
            return [float(low_km2)]
        inner = [float(a) for a in self.areas_km2 if low_km2 < a < high_km2]
        knots = sorted({float(low_km2), float(high_km2), *inner})
        values = [self.factor(a) for a in knots]
        first = float(self.areas_km2[0])
        if not self.reduces_below_table and low_km2 < first < high_km2:  # whole just below the first area
            at = knots.index(first)
            knots.insert(at, first)
            values.insert(at, 1.0)
        walked = np.concatenate([[0.0], np.cumsum(np.abs(np.diff(values)))])
        total = float(walked[-1])
        parts = max(1, int(np.ceil(total / largest_change - 1e-9)))
        areas = [float(low_km2)]
        for goal in total * np.arange(1, parts) / parts:
            k = min(max(int(np.searchsorted(walked, goal, side="left")), 1), len(knots) - 1)
            span = walked[k] - walked[k - 1]
            share = (goal - walked[k - 1]) / span if span > 0.0 else 1.0
            areas.append(float(knots[k - 1] + share * (knots[k] - knots[k - 1])))
        areas.append(float(high_km2))
        return sorted(set(areas))


@dataclass(frozen=True)
class GivenFactor:
    """One factor at every storm area (arf.py:71-87); a study refuses it, every run would be the same."""

    value: float

    def __post_init__(self) -> None:
        if not 0.0 < self.value <= 1.0:
            raise ValueError("a given area-reduction factor lies above 0 and at most 1")

    def factor(self, _area_km2: float | None = None) -> float:
        return self.value


def factor_at(arf: ArfTable | GivenFactor | None, area_km2: float | None) -> float:
    """The multiplier of a storm's depths: 1 without a reduction or, for a table, without a storm area; the given
    factor; the table read at the area (arf.py:90-98)."""
    if arf is None:
        return 1.0
    if isinstance(arf, GivenFactor):
        return arf.value
    return 1.0 if area_km2 is None else arf.factor(area_km2)


class ReducedStorm:
    """A cumulative storm curve (mm against s) times a factor."""

    def __init__(self, curve, factor: float) -> None:
        self.curve = curve
        self.factor = float(factor)

    def __call__(self, t: float) -> float:
        return self.factor * float(self.curve(t))


# ---- the study's areas and its reading of the runs (arf.py:79-247)


def bracket_areas(wanted: list[float], ratio: float, most: int) -> list[float]:
    """Log-spaced areas from the least to the greatest wanted, neighbours at most ``ratio`` apart unless ``most``
    caps their number; one area when all are the same (arf.py:79-86)."""
    low, high = min(wanted), max(wanted)
    if high <= low * (1.0 + 1e-9):
        return [low]
    count = int(np.ceil(np.log(high / low) / np.log(ratio) - 1e-9)) + 1
    return list(np.geomspace(low, high, min(max(count, 2), most)))


def interpolate_peak(areas: np.ndarray, peaks: np.ndarray, area: float) -> float:
    """A peak read between the runs linearly in (ln A, ln Q); the end runs hold beyond them; runs with no flow are
    passed over, and none at all reads 0 (arf.py:125-130)."""
    areas, peaks = np.asarray(areas, np.float64), np.asarray(peaks, np.float64)
    flowing = peaks > 0.0
    if not flowing.any():
        return 0.0
    return float(np.exp(np.interp(np.log(area), np.log(areas[flowing]), np.log(peaks[flowing]))))


def log_area_weights(areas: np.ndarray, area):
    """(low, high, w): the runs either side of ``area`` and the weight of the higher, linear in ln A, the value read
    as (1 - w) low + w high; beyond the runs the end run holds (arf.py:228-237). ``area`` may be an array."""
    ln_runs = np.log(np.asarray(areas, np.float64))
    ln_a = np.log(np.clip(area, areas[0], areas[-1]))
    high = np.clip(np.searchsorted(ln_runs, ln_a), 1, ln_runs.size - 1)
    low = high - 1
    return low, high, (ln_a - ln_runs[low]) / (ln_runs[high] - ln_runs[low])


def composite_map(maps: list[np.ndarray], areas: np.ndarray, cell_area: np.ndarray) -> np.ndarray:
    """Each cell read from the runs either side of its own drainage area, linear in ln A (arf.py:240-247)."""
    if np.asarray(areas).size == 1:
        return np.array(maps[0], dtype=np.float64)
    low, high, w = log_area_weights(np.asarray(areas, np.float64), cell_area)
    stack = np.stack(maps)
    rows, cols = np.indices(cell_area.shape, sparse=True)
    return (1.0 - w) * stack[low, rows, cols] + w * stack[high, rows, cols]


def design_row(name: str, area: float, areas: np.ndarray, peaks: np.ndarray, times: np.ndarray,
               kind: str = "") -> dict[str, object]:  # fmt: skip
    """One design point read from the runs either side of its drainage area; its time is the nearer run's (in ln A)
    (arf.py:133-156)."""
    areas = np.asarray(areas, np.float64)
    high = int(np.clip(np.searchsorted(areas, area), 0, areas.size - 1))
    low = max(high - 1, 0)
    nearer = low if abs(np.log(area / areas[low])) <= abs(np.log(area / areas[high])) else high
    return {
        "design_point": name,
        "drainage_area_km2": float(area),
        "storm_area_low_km2": float(areas[low]),
        "storm_area_high_km2": float(areas[high]),
        "peak_low_m3s": float(peaks[low]),
        "peak_high_m3s": float(peaks[high]),
        "design_peak_m3s": interpolate_peak(areas, peaks, area),
        "time_of_peak_s": float(times[nearer]),
        "arf_table_type": kind,
        "read_as": "line",
    }


# ---- the study as one batched launch


@dataclass
class StudyResult:
    areas_km2: list[float]
    factors: list[float]
    peak_q: list[float]  # m3/s out through the outflow boundary, largest after any step
    t_peak: list[float]
    max_depth: list[np.ndarray]
    mass_error: list[float]
    steps: list[int]
    wall_s: float
    rows: list[dict[str, object]]


def depth_area_study(scen, arf: ArfTable, areas_km2: list[float], backend: str = "hip",
                     design: list[tuple[str, float]] = (), graph_steps: int = 64, maxima_every: int = 1,
                     kind: str = "") -> StudyResult:  # fmt: skip
    """Run scenario ``scen``'s storm at every storm area in one batched launch (the serial loop of arf.py:347-371 as
    members of one device run): member k takes the storm times the table's factor at area k; then each design point
    (name, drainage area km2) is read from the runs either side of it."""
    from wadi2d_amd.engine.ensemble import Ensemble, run_batched

    if isinstance(arf, GivenFactor):
        raise ValueError("a depth-area study needs an area table: a given factor makes every run the same")
    areas = sorted({float(a) for a in areas_km2})
    factors = [arf.factor(a) for a in areas]
    if len(areas) > 1 and len({round(f, 9) for f in factors}) == 1:
        raise ValueError(f"the table gives one factor at all the storm areas {areas}: the runs would be identical")
    tables = scen.build()
    members = [dataclasses.replace(scen, cumulative_mm=ReducedStorm(scen.cumulative_mm, f), tables=tables,
                                   name=f"{scen.name} at {a:g} km2") for a, f in zip(areas, factors)]  # fmt: skip
    ens = Ensemble.from_scenarios(members, keep_beds=True)
    t0 = time.perf_counter()
    got = run_batched(ens, backend, graph_steps=graph_steps, tune=False, maxima_every=maxima_every)
    wall = time.perf_counter() - t0
    rows = [design_row(name, a, np.asarray(areas), np.asarray(got.peak_q), np.asarray(got.t_peak), kind)
            for name, a in design]  # fmt: skip
    return StudyResult(areas, factors, got.peak_q, got.t_peak, got.max_depth, got.mass_error, got.steps, wall, rows)
