"""The compute interface a backend implements, and the choice of one (wadi2d/solver/backend.py:1-75)."""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from wadi2d_amd.config import G, Numerics


@dataclass(frozen=True)
class NumericsRecord:
    """The run-time numbers kernels read; built once from the configuration (backend.py:14-38)."""

    g: float
    dx: float
    e_wet: float
    h_cfl: float
    u_max: float
    lead: float
    a_min: float
    second: bool
    sheet: bool
    mirror: bool
    carry: bool = True

    @classmethod
    def from_config(cls, num: Numerics, dx: float) -> NumericsRecord:
        lead = float(max(num.dt_every, 1 << num.lts_levels) if num.local_time_stepping else num.dt_every)
        return cls(G, dx, num.e_wet, num.h_cfl, num.u_max, lead, num.plan_min, num.limiter != "none",
                   num.sheet_flow, num.mirror_dry_faces, num.face_depth_carry)  # fmt: skip


@dataclass
class BoundaryCells:
    """Cells whose level is prescribed after every update. kind: 0 fixed, 1 free, 2 follows ``ref`` plus d0."""

    cell: np.ndarray
    ref: np.ndarray
    kind: np.ndarray
    d0: np.ndarray
    h0: np.ndarray

    @classmethod
    def empty(cls) -> BoundaryCells:
        i = np.zeros(0, np.int32)
        f = np.zeros(0, np.float32)
        return cls(i, i.copy(), i.copy(), f, f.copy())


BACKENDS = {
    "hip": ("f32", "fast"),
    "hip-exact": ("f32", "exact"),
    "hip-f64": ("f64", "exact"),
}
"""Backend names this package answers: float32 with the hardware's fast reciprocal/sqrt (the reference GPU's
``--use_fast_math`` counterpart), float32 with IEEE division/sqrt (bit for bit the float32 statement order of the
reference's CPU twin), and float64."""


class BackendUnavailable(RuntimeError):
    """No device, or the HIP library is not built."""


def select_backend(name: str, tables, num: NumericsRecord, **options):
    """``auto``, ``gpu`` and ``hip`` take the HIP step; ``hip-exact`` and ``hip-f64`` its other arithmetic. There is no
    silent CPU fallback: the CPU backend is the reference's own (solver/cpu), so a run without a device fails here."""
    from wadi2d_amd.solver.hip_backend import HipBackend, device_count

    key = "hip" if name in ("auto", "gpu") else name
    if key not in BACKENDS:
        raise BackendUnavailable(f"backend {name!r}: this package provides {', '.join(BACKENDS)} (and auto, gpu)")
    if device_count() == 0:
        raise BackendUnavailable("no HIP device available")
    precision, variant = BACKENDS[key]
    return HipBackend(tables, num, precision, variant, **options)
