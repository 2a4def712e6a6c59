"""Rain lost before it runs off (behaviour of wadi2d/hydrology/losses.py:1-77).

Two methods: none (all rain runs off) and the SCS curve number in its cumulative form, in which a period's runoff is
the rise over the period of the runoff the curve gives for the rain fallen so far. For a curve number CN the
retention is S = (1000 / CN - 10) inch, the initial abstraction Ia = 0.2 S unless given, and rain P fallen so far
yields (P - Ia)^2 / (P - Ia + S) of runoff once P passes Ia, none before. An impervious share passes its rain whole.

``excess_kernel`` in csrc/step.hip forms the per-cell numbers on the device with the same float64 statements.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

MM_PER_INCH = 25.4  # hydrology/constants.py
PERCENT = 100.0
CN_NUMERATOR = 1000.0  # S in inches is 1000 / CN - 10 (losses.py:18-20)
CN_OFFSET = 10.0
DEFAULT_IA_RATIO = 0.2


@dataclass(frozen=True)
class NoLoss:
    """Every millimetre of rain runs off (losses.py:37-44)."""

    def excess(self, precipitation_mm: np.ndarray) -> np.ndarray:
        return np.array(precipitation_mm, dtype=np.float64, copy=True)


@dataclass(frozen=True)
class ScsCurveNumber:
    """The SCS curve-number loss (losses.py:47-77). ``curve_number`` may be one number or an array of them (one
    independent loss per entry); ``initial_abstraction_mm`` is a depth, or None for 0.2 S; ``percent_impervious``
    is the share of rain that passes whole."""

    curve_number: float | np.ndarray
    initial_abstraction_mm: float | None = None
    percent_impervious: float = 0.0

    @property
    def retention_mm(self):
        """S in mm."""
        inches = CN_NUMERATOR / np.asarray(self.curve_number, np.float64) - CN_OFFSET
        return inches * MM_PER_INCH

    @property
    def abstraction_mm(self):
        """Ia in mm: the given depth, else 0.2 S."""
        given = self.initial_abstraction_mm
        return DEFAULT_IA_RATIO * self.retention_mm if given is None else given

    def cumulative_excess(self, cumulative_mm) -> np.ndarray:
        """Runoff (mm) from the pervious part once ``cumulative_mm`` of rain has fallen."""
        over = np.maximum(np.asarray(cumulative_mm, np.float64) - self.abstraction_mm, 0.0)
        return over**2 / (over + self.retention_mm)

    def excess(self, precipitation_mm: np.ndarray) -> np.ndarray:
        """Runoff of each period from the rain of each period: the pervious part's rise in cumulative runoff, plus the
        impervious part's rain."""
        rain = np.asarray(precipitation_mm, np.float64)
        pervious = np.diff(self.cumulative_excess(np.cumsum(rain)), prepend=0.0)
        share = self.percent_impervious / PERCENT
        return (1.0 - share) * pervious + share * rain


def cell_excess(depth_mm, curve_number: np.ndarray) -> np.ndarray:
    """Cumulative excess per cell of cumulative depths ``depth_mm`` (a number or one per cell): the depth where the
    cell has no curve number (NaN), else the SCS excess read through nan_to_num (forcing.py:644-665)."""
    cn = np.asarray(curve_number, np.float64)
    out = np.array(np.broadcast_to(np.asarray(depth_mm, np.float64), cn.shape), dtype=np.float64)
    has = np.isfinite(cn)
    if has.any():
        with np.errstate(divide="ignore", invalid="ignore"):
            out[has] = np.nan_to_num(ScsCurveNumber(cn[has]).cumulative_excess(out[has]))
    return out
