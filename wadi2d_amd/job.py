"""job.json: the part of the reference's job schema a uniform-grid run reads (wadi2d/config/schema.py).

Section and field names, defaults and bounds follow the reference, so a ``job.json`` written for it validates here
for the parts this path runs: the grid (schema.py:48-72), the terrain's DEM and Manning n (:315-420), the tables
and numerics (:422-516), the storm kinds that are a depth curve (:1069-1170), open boundaries (:587-614), initial
levels (:617-622), outputs (:745-760, 988-1046) and the run (:1576-1600).

What a job may name that this path does not run -- a quadtree or strip mesh, terrain objects, lumped hydrology,
transmission loss, breaches, dam levels, flow lines, gridded or gauge rain -- is refused with ``InvalidJob``
(exit 2) when the job switches it on, and otherwise read and left alone, so a reference job that keeps those
sections at their defaults loads unchanged. The reference reads its DEM with rasterio; this path reads an ESRI
ASCII grid (``.asc``) or a NumPy file (``.npz`` holding ``z``, ``x0``, ``y0``, ``pixel``), see ``wadi2d_amd.dem``.
"""

from __future__ import annotations

import json
from pathlib import Path
from typing import Any, Literal

from pydantic import BaseModel, ConfigDict, Field, ValidationError, model_validator

from wadi2d_amd.config import Numerics, Tables


class JobError(Exception):
    """A failure the run protocol names; ``code`` is the process exit code (run.py:130-146)."""

    code = 1


class InvalidJob(JobError):
    code = 2


class InputMissing(JobError):
    code = 3


class BackendUnavailable(JobError):
    code = 4


class _Part(BaseModel):
    """A section: unknown keys are kept (``extra``) so the reference's other fields load; ``_inert`` checks them."""

    model_config = ConfigDict(extra="allow", frozen=True)


def _falsy(value: Any) -> bool:
    return value is None or value is False or value == 0 or value == [] or value == {} or value == ""


class GridConfig(_Part):
    dx: float = Field(10.0, ge=0.5, le=1000.0, description="Cell size, m")
    domain: str | None = Field(None, description="Polygon file; None uses extent")
    extent: tuple[float, float, float, float] | None = Field(None, description="xmin, ymin, xmax, ymax, m")
    line_samples: int = Field(4, ge=1, le=32, description="Points per cell when a boundary line picks its cells")


class Terrain(_Part):
    dem: str | None = Field(None, description="DEM path: .asc (ESRI ASCII grid) or .npz (z, x0, y0, pixel)")
    pixel_stride: int = Field(1, ge=1, le=64, description="Read every n-th DEM pixel")
    manning: float = Field(0.035, ge=0.008, le=0.5, description="Default Manning n, s/m^(1/3)")
    manning_raster: str | None = None
    cn_raster: str | None = Field(None, description="Curve numbers read at the cell centres (.asc or .npz); they "
                                  "beat hydrology.curve_number where they have a value (forcing.py:186-203)")


class Storm(_Part):
    """Rain as a cumulative depth curve over the interior cells: 'none', 'uniform' (depth_mm over duration_min),
    'hypothetical' (time_percent / depth_percent), 'hyetograph' (hyetograph_min / hyetograph_mm) or 'depth' with
    no pattern (uniform in time). ``arf_factor`` multiplies every depth but a hyetograph's (schema.py:1069-1170)."""

    kind: Literal["none", "uniform", "hypothetical", "hyetograph", "gridded", "gauges", "depth", "depth_grid"] = "none"
    depth_mm: float = Field(0.0, ge=0.0, le=5000.0)
    duration_min: float = Field(60.0, gt=0.0, le=20160.0)
    time_percent: list[float] | None = None
    depth_percent: list[float] | None = None
    arf_factor: float | None = Field(None, gt=0.0, le=1.0)
    hyetograph_min: list[float] | None = None
    hyetograph_mm: list[float] | None = None
    depth_grid: str | None = Field(None, description="Storm depths per cell, mm (.asc or .npz), for kind "
                                   "'depth_grid': each cell takes its depth over duration_min (forcing.py:594-601)")
    rain_file: str | None = Field(None, description="Rain frames for kind 'gridded' (.npz: depths mm per frame, "
                                  "starts_min, ends_min, west, north, dx, dy); each cell reads the pixel under its "
                                  "centre (storm/gridded.py, hydrology/cell_rain.py)")

    @model_validator(mode="after")
    def _complete(self) -> Storm:
        # schema.py:1158-1161
        if self.kind == "depth" and (self.depth_grid is not None or self.depth_mm <= 0.0):
            raise ValueError("a depth storm needs storm.depth_mm above zero; a raster of depths is kind 'depth_grid'")
        if self.kind == "depth_grid" and (self.depth_grid is None or self.depth_mm != 0.0):
            raise ValueError("a depth_grid storm needs storm.depth_grid, and takes its depths from it, not depth_mm")
        if self.kind == "gridded" and self.rain_file is None:
            raise ValueError("a gridded storm needs storm.rain_file")
        if self.kind == "hypothetical":
            t, d = self.time_percent, self.depth_percent
            if not t or not d or len(t) != len(d):
                raise ValueError("a hypothetical storm needs time_percent and depth_percent of the same length")
        if self.kind == "hyetograph":
            t, d = self.hyetograph_min, self.hyetograph_mm
            if not t or not d or len(t) != len(d):
                raise ValueError("a hyetograph needs hyetograph_min and hyetograph_mm of the same length")
        return self


class Boundary(_Part):
    kind: Literal["stage", "normal", "free", "flow"]
    coordinates: list[tuple[float, float]] = Field(description="Polyline along the boundary cells, m")
    stage: float = Field(0.0, description="Level for kind 'stage', m")
    slope: float = Field(0.005, ge=1e-5, le=0.5, description="Energy slope for kind 'normal'")
    snap_m: float = Field(150.0, ge=0.0, le=5000.0)
    min_share: float = Field(0.6, ge=0.0, le=1.0)
    name: str | None = None


class InitialLevel(_Part):
    coordinates: list[tuple[float, float]] = Field(description="Polygon ring, m", min_length=3)
    level: float = Field(description="Water level on the datum of the DEM, m")


class DesignPoint(_Part):
    name: str
    x: float
    y: float


class DesignLine(_Part):
    name: str
    coordinates: list[tuple[float, float]] = Field(min_length=2)


class Outputs(_Part):
    map_every_s: float = Field(900.0, ge=0.0, le=86400.0)
    series_every_s: float = Field(60.0, ge=1.0, le=3600.0)
    maxima_every: int = Field(4, ge=1, le=64)
    arrival_depth: float = Field(0.1, ge=0.0, le=10.0)
    speed_depth: float = Field(0.05, ge=0.0, le=1.0)
    wet_depth: float = Field(0.1, ge=0.0, le=10.0)
    progress_every_s: float = Field(2.0, ge=0.05, le=600.0)
    face_discharge: bool = True
    points: list[DesignPoint] = []
    lines: list[DesignLine] = []


class Run(_Part):
    duration_s: float = Field(3600.0, ge=1.0, le=2.6e6)
    backend: Literal["auto", "gpu", "cpu"] = "auto"
    source_every_s: float = Field(30.0, ge=1.0, le=600.0)
    checkpoint_every_s: float = Field(300.0, ge=0.0, le=86400.0)


# (section, key): settings that change the water and that this path does not run; a job must leave them off
SWITCHED_OFF = (
    ("grid", "domain"), ("grid", "exclude_lumped"), ("grid", "exclude_subbasins"),
    ("terrain", "manning_raster"), ("terrain", "objects"), ("terrain", "roughness"),
    ("terrain", "fill_pits"),
    ("storm", "gauges"), ("storm", "area_reduction"),
    ("storm", "arf"), ("storm", "pattern"),
    ("numerics", "local_time_stepping"),
    ("mesh", "enabled"), ("strips", "enabled"), ("transmission", "enabled"),
    ("hydrology", "subbasins"), ("hydrology", "hms_basin"), ("hydrology", "basin_json"), ("hydrology", "handbacks"),
    ("run", "lumped_only"),
)  # fmt: skip
TOP_LISTS_OFF = ("initial_points", "dam_levels", "breaches")
"""Top-level lists this path has no part for: they must be empty."""
READ_ONLY_SECTIONS = ("project", "mesh", "strips", "breaklines", "hydrology", "transmission", "arf", "basin",
                      "data_tables")  # fmt: skip
"""Sections of the reference's job read only for the switches above."""


class RunConfig(BaseModel):
    """The whole job (schema.py:1955-1984), the part a uniform-grid run reads."""

    model_config = ConfigDict(extra="forbid", frozen=True)

    grid: GridConfig = GridConfig()
    terrain: Terrain = Terrain()
    tables: Tables = Tables()
    numerics: Numerics = Numerics()
    storm: Storm = Storm()
    boundaries: list[Boundary] = []
    initial_levels: list[InitialLevel] = []
    outputs: Outputs = Outputs()
    run: Run = Run()
    project: dict = {}
    mesh: dict = {}
    strips: dict = {}
    breaklines: dict = {}
    hydrology: dict = {}
    transmission: dict = {}
    arf: dict = {}
    basin: dict = {}
    data_tables: dict = {}
    initial_points: list = []
    dam_levels: list = []
    breaches: list = []

    @model_validator(mode="before")
    @classmethod
    def _tables_numerics_extra(cls, data: Any) -> Any:
        """Tables and Numerics forbid unknown keys; the reference's extra fields in them are dropped here when inert."""
        if not isinstance(data, dict):
            return data
        data = dict(data)
        for name, model in (("tables", Tables), ("numerics", Numerics)):
            part = data.get(name)
            if isinstance(part, dict):
                known = {k: v for k, v in part.items() if k in model.model_fields}
                for k, v in part.items():
                    if k not in model.model_fields and not _falsy(v) and (name, k) in SWITCHED_OFF:
                        raise ValueError(f"{name}.{k} is not run by this path")
                data[name] = known
        return data

    @model_validator(mode="after")
    def _curve_number(self) -> RunConfig:
        """hydrology.curve_number: CN of 2D cells without a better source, in (0, 100] (schema.py:1435-1437)."""
        cn = self.hydrology.get("curve_number")
        if cn is not None and (isinstance(cn, bool) or not isinstance(cn, (int, float)) or not 0.0 < cn <= 100.0):
            raise ValueError("hydrology.curve_number must lie in (0, 100]")
        return self


def unsupported(cfg: RunConfig) -> list[str]:
    """What the job switches on that this path does not run, in words; empty for a job it runs as written."""
    out = []
    for section, key in SWITCHED_OFF:
        part = getattr(cfg, section)
        value = part.get(key) if isinstance(part, dict) else getattr(part, key, None)
        if value is None and not isinstance(part, dict):
            value = (part.model_extra or {}).get(key)
        if not _falsy(value):
            out.append(f"{section}.{key}")
    for name in TOP_LISTS_OFF:
        if getattr(cfg, name):
            out.append(name)
    if cfg.storm.kind == "gauges":
        out.append(f"storm.kind '{cfg.storm.kind}'")
    out += [f"boundaries: kind 'flow' ({b.name or 'unnamed'})" for b in cfg.boundaries if b.kind == "flow"]
    return out


def ignored(cfg: RunConfig) -> list[str]:
    """Fields of the job that this path reads past: the reference's settings with no effect on a uniform-grid run
    here (live rasters, products, estimates ...) and the read-only sections, for the run's 'warn' events."""
    out = []
    for name in ("grid", "terrain", "storm", "outputs", "run"):
        extra = getattr(cfg, name).model_extra or {}
        out += [f"{name}.{k}" for k, v in extra.items() if not _falsy(v)]
    out += [name for name in READ_ONLY_SECTIONS if name != "hydrology" and getattr(cfg, name)]
    out += [f"hydrology.{k}" for k, v in cfg.hydrology.items() if k not in RUN_HYDROLOGY and not _falsy(v)]
    return out


RUN_HYDROLOGY = ("curve_number",)
"""Keys of the hydrology section this path runs: the 2D cells' default curve number."""

PATH_FIELDS = (("terrain", "dem"), ("terrain", "cn_raster"), ("storm", "depth_grid"), ("storm", "rain_file"))
"""Paths of a job this path reads; relative ones are relative to the job file (run.py:147-163, 209-220)."""


def load_job(path: str | Path) -> RunConfig:
    """Parse and validate a job file; relative paths become absolute (run.py:197-206)."""
    source = Path(path)
    if not source.is_file():
        raise InputMissing(f"job file not found: {source}")
    try:
        cfg = RunConfig.model_validate(json.loads(source.read_text(encoding="utf-8")))
    except (ValidationError, json.JSONDecodeError, ValueError) as error:
        raise InvalidJob(str(error)) from error
    updates = {}
    for part, name in PATH_FIELDS:
        value = getattr(getattr(cfg, part), name)
        if value is not None and not Path(value).is_absolute():
            updates.setdefault(part, {})[name] = str((source.parent / value).resolve())
    if updates:
        cfg = cfg.model_copy(update={p: getattr(cfg, p).model_copy(update=u) for p, u in updates.items()})
    problems = unsupported(cfg)
    if problems:
        raise InvalidJob("this build runs the uniform-grid step only; the job switches on: " + ", ".join(problems))
    return cfg


def job_hash(cfg: RunConfig) -> str:
    """What a checkpoint must match: every setting that changes the water (run.py:384-490, NOT_WATER left out)."""
    import hashlib

    data = cfg.model_dump(mode="json")
    data["run"].pop("backend", None)
    data["run"].pop("checkpoint_every_s", None)
    for key in ("progress_every_s", "map_every_s", "series_every_s", "points", "lines", "face_discharge"):
        data["outputs"].pop(key, None)
    for name in READ_ONLY_SECTIONS:
        data.pop(name, None)
    dem = cfg.terrain.dem
    if dem and Path(dem).is_file():
        st = Path(dem).stat()
        data["dem_identity"] = [st.st_size, int(st.st_mtime)]
    return hashlib.sha256(json.dumps(data, sort_keys=True).encode()).hexdigest()[:16]
