"""Design storms for the ensemble: the NRCS 24-hour Type II distribution as an HMS Hypothetical Storm, reduced by an
area-reduction factor (wadi2d/storm/patterns.py:33-44, storm/hypothetical.py:43-69, storm/arf.py:89-98).

The 24-hour curve is held the way HEC-HMS holds it: the NRCS ordinates every 0.1 h, carried to every minute on a
natural cubic spline through them, the ordinates themselves kept to the digit; between minutes it is linear. A storm
of duration D stretches the 24-hour curve over D (percent of duration against percent of depth), and its depth is
the point depth times the factor.
"""

from __future__ import annotations

from functools import cache

import numpy as np

from wadi2d_amd.config import MM_PER_M

KNOT_HOURS = 0.1
SPAN_MIN = 1440.0

TYPE_II = (
    0.0000, 0.0010, 0.0020, 0.0030, 0.0041, 0.0051, 0.0062, 0.0072, 0.0083, 0.0094,
    0.0105, 0.0116, 0.0127, 0.0138, 0.0150, 0.0161, 0.0173, 0.0184, 0.0196, 0.0208,
    0.0220, 0.0232, 0.0244, 0.0257, 0.0269, 0.0281, 0.0294, 0.0306, 0.0319, 0.0332,
    0.0345, 0.0358, 0.0371, 0.0384, 0.0398, 0.0411, 0.0425, 0.0439, 0.0452, 0.0466,
    0.0480, 0.0494, 0.0508, 0.0523, 0.0538, 0.0553, 0.0568, 0.0583, 0.0598, 0.0614,
    0.0630, 0.0646, 0.0662, 0.0679, 0.0696, 0.0712, 0.0730, 0.0747, 0.0764, 0.0782,
    0.0800, 0.0818, 0.0836, 0.0855, 0.0874, 0.0892, 0.0912, 0.0931, 0.0950, 0.0970,
    0.0990, 0.1010, 0.1030, 0.1051, 0.1072, 0.1093, 0.1114, 0.1135, 0.1156, 0.1178,
    0.1200, 0.1222, 0.1246, 0.1270, 0.1296, 0.1322, 0.1350, 0.1379, 0.1408, 0.1438,
    0.1470, 0.1502, 0.1534, 0.1566, 0.1598, 0.1630, 0.1663, 0.1697, 0.1733, 0.1771,
    0.1810, 0.1851, 0.1895, 0.1941, 0.1989, 0.2040, 0.2094, 0.2152, 0.2214, 0.2280,
    0.2350, 0.2427, 0.2513, 0.2609, 0.2715, 0.2830, 0.3068, 0.3544, 0.4308, 0.5679,
    0.6630, 0.6820, 0.6986, 0.7130, 0.7252, 0.7350, 0.7434, 0.7514, 0.7588, 0.7656,
    0.7720, 0.7780, 0.7836, 0.7890, 0.7942, 0.7990, 0.8036, 0.8080, 0.8122, 0.8162,
    0.8200, 0.8237, 0.8273, 0.8308, 0.8342, 0.8376, 0.8409, 0.8442, 0.8474, 0.8505,
    0.8535, 0.8565, 0.8594, 0.8622, 0.8649, 0.8676, 0.8702, 0.8728, 0.8753, 0.8777,
    0.8800, 0.8823, 0.8845, 0.8868, 0.8890, 0.8912, 0.8934, 0.8955, 0.8976, 0.8997,
    0.9018, 0.9038, 0.9058, 0.9078, 0.9097, 0.9117, 0.9136, 0.9155, 0.9173, 0.9192,
    0.9210, 0.9228, 0.9245, 0.9263, 0.9280, 0.9297, 0.9313, 0.9330, 0.9346, 0.9362,
    0.9377, 0.9393, 0.9408, 0.9423, 0.9438, 0.9452, 0.9466, 0.9480, 0.9493, 0.9507,
    0.9520, 0.9533, 0.9546, 0.9559, 0.9572, 0.9584, 0.9597, 0.9610, 0.9622, 0.9635,
    0.9647, 0.9660, 0.9672, 0.9685, 0.9697, 0.9709, 0.9722, 0.9734, 0.9746, 0.9758,
    0.9770, 0.9782, 0.9794, 0.9806, 0.9818, 0.9829, 0.9841, 0.9853, 0.9864, 0.9876,
    0.9887, 0.9899, 0.9910, 0.9922, 0.9933, 0.9944, 0.9956, 0.9967, 0.9978, 0.9989,
    1.0000,
)
"""Cumulative fraction of the 24-hour depth at every 0.1 h, 0 to 24 h: the NRCS (TR-55) Type II distribution."""


@cache
def scs_curve() -> tuple[np.ndarray, np.ndarray]:
    """Minutes 0..1440 and the fraction of the depth fallen by each (natural cubic spline through the knots)."""
    from scipy.interpolate import CubicSpline

    knots = np.asarray(TYPE_II, np.float64)
    at = np.arange(knots.size) * KNOT_HOURS * 60.0
    minutes = np.arange(0.0, SPAN_MIN + 0.5, 1.0)
    fraction = CubicSpline(at, knots, bc_type="natural")(minutes)
    step = int(round(KNOT_HOURS * 60.0))
    fraction[::step] = knots
    return minutes, fraction


class ScsStorm:
    """``depth_mm`` (point depth) over ``duration_h`` hours on the Type II curve, times ``arf``."""

    def __init__(self, depth_mm: float, duration_h: float, arf: float = 1.0) -> None:
        if not 0.0 < arf <= 1.0:
            raise ValueError("an area-reduction factor lies above 0 and at most 1")
        self.depth_mm = float(depth_mm)
        self.duration_s = float(duration_h) * 3600.0
        self.arf = float(arf)
        minutes, fraction = scs_curve()
        self._pct_t = minutes / SPAN_MIN * 100.0
        self._pct_d = fraction * 100.0

    @property
    def total_mm(self) -> float:
        return self.depth_mm * self.arf

    def __call__(self, t: float) -> float:
        """Depth fallen by ``t`` seconds, mm."""
        pct = np.interp(float(t) / self.duration_s * 100.0, self._pct_t, self._pct_d)
        return self.total_mm * float(pct) / 100.0

    def mean_rate(self, t0: float, t1: float) -> float:
        """Mean rate over [t0, t1], m/s."""
        return (self(t1) - self(t0)) / (MM_PER_M * (t1 - t0))
