"""Rainfall losses (wadi2d/hydrology/losses.py:1-77): the SCS curve-number method in HMS's cumulative-difference
form, and no loss.

The device forms the same numbers per cell (``excess_kernel`` in csrc/step.hip): the statements below are the
ones both follow, in float64.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

MM_PER_INCH = 25.4  # hydrology/constants.py
PERCENT = 100.0
CN_NUMERATOR = 1000.0  # losses.py:18-20
CN_OFFSET = 10.0
DEFAULT_IA_RATIO = 0.2


@dataclass(frozen=True)
class NoLoss:
    """All rainfall is excess (losses.py:37-44)."""

    def excess(self, precipitation_mm: np.ndarray) -> np.ndarray:
        return np.array(precipitation_mm, dtype=np.float64, copy=True)


@dataclass(frozen=True)
class ScsCurveNumber:
    """SCS curve number (losses.py:47-77); ``curve_number`` may be an array (one loss per entry).
    ``initial_abstraction_mm`` None means 0.2 S."""

    curve_number: float | np.ndarray
    initial_abstraction_mm: float | None = None
    percent_impervious: float = 0.0

    @property
    def retention_mm(self):
        """Potential maximum retention S = (1000 / CN - 10) 25.4 mm."""
        return (CN_NUMERATOR / np.asarray(self.curve_number, np.float64) - CN_OFFSET) * MM_PER_INCH

    @property
    def abstraction_mm(self):
        """Initial abstraction Ia, mm."""
        if self.initial_abstraction_mm is None:
            return DEFAULT_IA_RATIO * self.retention_mm
        return self.initial_abstraction_mm

    def cumulative_excess(self, cumulative_mm) -> np.ndarray:
        """Pervious excess Pe(P) = (P - Ia)^2 / (P - Ia + S), zero up to Ia."""
        effective = np.maximum(np.asarray(cumulative_mm, np.float64) - self.abstraction_mm, 0.0)
        return effective**2 / (effective + self.retention_mm)

    def excess(self, precipitation_mm: np.ndarray) -> np.ndarray:
        """The excess of each period: the difference of cumulative excess, impervious rain passing whole."""
        p = np.asarray(precipitation_mm, np.float64)
        pervious = np.diff(self.cumulative_excess(np.cumsum(p)), prepend=0.0)
        impervious = self.percent_impervious / PERCENT
        return (1.0 - impervious) * pervious + impervious * p


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
