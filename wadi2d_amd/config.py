"""The part of the reference's job schema this path reads (wadi2d/config/schema.py, constants.py).

Field names, defaults and bounds follow the reference so a ``job.json`` written for it validates here for the
fields this path knows. Models are frozen and forbid unknown keys, as there (schema.py:38-39).
"""

from __future__ import annotations

from typing import Literal

from pydantic import BaseModel, ConfigDict, Field

G = 9.80665
"""Standard gravity, m/s2 (constants.py:3)."""

CRITICAL_WEIR_COEFF = 1.7046
"""(2/3) * sqrt(2 g / 3): ideal critical-flow discharge coefficient over a crest, m^0.5/s (constants.py:6)."""

SECONDS_PER_HOUR = 3600.0
MM_PER_M = 1000.0


class _Part(BaseModel):
    model_config = ConfigDict(extra="forbid", frozen=True)


class Tables(_Part):
    """Sub-grid tables (schema.py:422-454)."""

    rows: int = Field(16, ge=4, le=64)
    face_depth_max: float = Field(16.0, ge=1.0, le=100.0)
    face_depth_min: float = Field(8.0, ge=0.5, le=50.0)
    face_conveyance: Literal["strips", "local_radius"] = "local_radius"
    local_radius_beta: float = Field(9.0, gt=0.0, le=50.0)
    crest_tolerance: float = Field(0.05, ge=0.0, le=1.0)


class Numerics(_Part):
    """Step numerics (schema.py:457-516)."""

    cfl: float = Field(0.45, ge=0.05, le=0.7)
    limiter: Literal["none", "minmod"] = "minmod"
    sheet_flow: bool = True
    mirror_dry_faces: bool = True
    face_depth_carry: bool = True
    e_wet: float = Field(1e-4, ge=1e-6, le=0.1)
    h_cfl: float = Field(0.01, ge=1e-4, le=1.0)
    u_max: float = Field(25.0, ge=1.0, le=100.0)
    dt_max: float = Field(5.0, ge=1e-3, le=600.0)
    dt_min: float = Field(1e-3, ge=1e-6, le=1.0)
    dt_every: int = Field(8, ge=1, le=64)
    wave_floor: float = Field(3.0, ge=0.1, le=30.0)
    plan_min: float = Field(0.02, ge=1e-4, le=1.0)
    active_every: int = Field(16, ge=1, le=32)
    local_time_stepping: bool = False
    lts_levels: int = Field(3, ge=0, le=5)
    lts_wave_floor: float = Field(1.0, ge=0.1, le=30.0)
    lts_buffer: float = Field(1.0, ge=0.5, le=1.0)
