"""One job end to end on the uniform grid: ``run_job`` (wadi2d/run.py:1445-1759,1862-1901).

It reads the DEM, builds the tables, places the boundaries and the initial water, and runs the window loop of the
reference: every ``run.source_every_s`` the rain source is refreshed and the model advanced to the window's end
exactly; every step the boundary tally counts in FP64, every ``outputs.maxima_every`` steps the maxima are sampled
on the device, every ``outputs.series_every_s`` the points and lines are recorded, every ``outputs.map_every_s`` a
map is saved, and every ``outputs.progress_every_s`` of wall time a ``step`` event is written and the stop file
looked for. A stop, or the wall-clock checkpoint interval, saves ``checkpoint.npz`` at a window's end, so a resume
carries on the unbroken run's windows.

``run_job`` always returns; ``RunSummary.exit_code`` is the process exit code (run.py:130-133): 0 finished,
1 unexpected, 2 invalid job, 3 input missing, 4 backend unavailable, 5 numerical failure, 6 stopped, 7 out of
memory; ``progress.jsonl`` records what happened whatever the outcome.
"""

from __future__ import annotations

import time
import traceback
from collections.abc import Iterable
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from wadi2d_amd.dem import grid_over, raster_at_cells, read_raster
from wadi2d_amd.engine import forcing
from wadi2d_amd.engine.driver import Balance, Driver
from wadi2d_amd.engine.gridded import read_frames
from wadi2d_amd.engine.lines import cells_along, inside, line_faces
from wadi2d_amd.io import (
    SAVED_FIELDS,
    Progress,
    Results,
    load_checkpoint,
    safe_name,
    save_checkpoint,
    write_series_csv,
)
from wadi2d_amd.job import BackendUnavailable, InvalidJob, JobError, RunConfig, ignored, job_hash
from wadi2d_amd.solver.state import from_state, state_index
from wadi2d_amd.solver.stepper import NumericalFailure
from wadi2d_amd.tables import build_tables, hbar_at_level

UNIT_DEPTH = 1.0  # forcing.py:54: a depth grid's storm falls one millimetre, each cell scaling it by its own depth

EXIT_DONE, EXIT_UNEXPECTED, EXIT_INVALID, EXIT_INPUT, EXIT_BACKEND = 0, 1, 2, 3, 4
EXIT_NUMERICAL, EXIT_STOPPED, EXIT_MEMORY = 5, 6, 7
MAXIMA = ("depth", "mean_depth", "speed", "hazard", "arrival")
"""The running maxima, in the order of the device's arrays (engine/maxima.py:23)."""
M2_PER_KM2 = 1.0e6


class StopRequested(Exception):
    """A stop file was found (engine/monitor.py)."""


@dataclass
class RunSummary:
    """What a run reports to its caller (run.py:1080-1101)."""

    exit_code: int = EXIT_UNEXPECTED
    status: str = "failed"
    message: str = ""
    t_end: float = 0.0
    steps: int = 0
    tables_s: float = 0.0
    loop_s: float = 0.0
    wet_area_km2: float = 0.0
    dt_min: float = 0.0
    balance: Balance = field(default_factory=Balance)
    peaks: dict[str, tuple[float, float]] = field(default_factory=dict)
    backend: str = ""
    cells: int = 0
    done: dict = field(default_factory=dict)


# ---------------------------------------------------------------------------------------------------------------
# the job's inputs


def storm_curve(cfg: RunConfig):
    """Cumulative depth in mm at time t (s) of the job's storm, or None for no rain (storm/uniform.py:32-35,
    storm/hyetograph.py:19-48, the hypothetical storm's percent table read linearly)."""
    s = cfg.storm
    factor = s.arf_factor if s.arf_factor is not None else 1.0
    if s.kind == "none":
        return None
    if s.kind in ("uniform", "depth"):
        return forcing.uniform_storm(s.depth_mm * factor, s.duration_min * 60.0)
    if s.kind == "depth_grid":  # a unit depth, each cell's own from the grid (forcing.py:109)
        return forcing.uniform_storm(UNIT_DEPTH * factor, s.duration_min * 60.0)
    if s.kind == "gridded":  # each cell reads its pixel's frames (build_rain): no one curve
        return None
    if s.kind == "hypothetical":
        tp = np.concatenate(([0.0], np.asarray(s.time_percent, np.float64)))
        dp = np.concatenate(([0.0], np.asarray(s.depth_percent, np.float64)))
        if np.any(np.diff(tp) < 0.0) or np.any(np.diff(dp) < 0.0):
            raise ValueError("a hypothetical storm's time and depth percents must not fall")
        total = s.depth_mm * factor
        duration = s.duration_min * 60.0

        def hypothetical(t: float) -> float:
            return total * float(np.interp(100.0 * t / duration, tp, dp)) / 100.0

        return hypothetical
    if s.kind == "hyetograph":
        minutes = np.concatenate(([0.0], np.asarray(s.hyetograph_min, np.float64)))
        depth = np.concatenate(([0.0], np.cumsum(np.asarray(s.hyetograph_mm, np.float64))))
        if np.any(np.diff(minutes) <= 0.0) or np.any(np.diff(depth) < 0.0):
            raise ValueError("a hyetograph's times must rise and its depths must not fall")

        def hyetograph(t: float) -> float:
            return float(np.interp(t / 60.0, minutes, depth))

        return hyetograph
    raise InvalidJob(f"storm.kind {s.kind!r} is not run by this path")


def build_rain(cfg: RunConfig, tables, bset, curve) -> forcing.Rain:
    """The job's rain over the interior cells (forcing.py:527-550): with a curve number (hydrology.curve_number,
    beaten by terrain.cn_raster where it has a value, forcing.py:186-203) or a depth grid, the excess form, whose
    rates the device forms; else the storm curve alone."""
    grid = tables.grid
    interior = grid.mask & ~bset.mask
    default = cfg.hydrology.get("curve_number")
    cn = None
    if default is not None or cfg.terrain.cn_raster is not None:
        cn = np.full(grid.shape, np.nan if default is None else float(default))
        if cfg.terrain.cn_raster is not None:
            raster = raster_at_cells(cfg.terrain.cn_raster, grid)
            cn = np.where(np.isfinite(raster), raster, cn)
            holes = int(np.count_nonzero(np.isnan(cn) & interior))
            if holes:  # forcing.py:582-592
                raise ValueError(f"terrain.cn_raster gives no CN on {holes} rained cells of {int(interior.sum())}, "
                                 "and nothing else does: extend the raster or set hydrology.curve_number")
        if np.any(np.isfinite(cn) & interior & ~((cn > 0.0) & (cn <= 100.0))):
            raise ValueError("terrain.cn_raster holds curve numbers outside (0, 100]")
    frames = None
    if cfg.storm.kind == "gridded":
        frames = read_frames(cfg.storm.rain_file)
    depths = None
    if cfg.storm.kind == "depth_grid":  # forcing.py:594-601
        values = raster_at_cells(cfg.storm.depth_grid, grid)
        holes = int(np.count_nonzero(~np.isfinite(values) & interior))
        if holes:
            raise ValueError(f"storm.depth_grid has no depth on {holes} rained cells of {int(interior.sum())}")
        depths = np.nan_to_num(values)
    zero = lambda t: 0.0  # noqa: E731
    factor = cfg.storm.arf_factor if cfg.storm.arf_factor is not None else 1.0
    return forcing.Rain(tables, bset, curve if curve is not None else zero, curve_number=cn, depth_grid=depths,
                        frames=frames, area_factor=factor)  # fmt: skip


def build_grid(cfg: RunConfig):
    """DEM -> grid -> tables (run.py prepare_grid, tables/builder.py:131-258)."""
    if cfg.terrain.dem is None:
        raise InvalidJob("terrain.dem is required: the run needs ground")
    raster = read_raster(cfg.terrain.dem, cfg.terrain.pixel_stride)
    grid, bed = grid_over(raster, cfg.grid.dx, cfg.grid.extent)
    return build_tables(grid, bed, cfg.tables, cfg.terrain.manning)


def boundary_list(cfg: RunConfig, tables) -> list[forcing.Boundary]:
    """Each open boundary's cells along its line (forcing.py:248-296)."""
    out = []
    for b in cfg.boundaries:
        jj, ii = cells_along(tables.grid, b.coordinates, cfg.grid.line_samples, b.snap_m)
        out.append(forcing.Boundary(b.kind, jj, ii, stage=b.stage, slope=b.slope))
    return out


def initial_depth(cfg: RunConfig, tables) -> np.ndarray:
    """Mean depth at the start: each polygon fills its cells to its level, later ones over earlier ones
    (forcing.py:227-246)."""
    grid = tables.grid
    out = np.zeros(grid.shape)
    if not cfg.initial_levels:
        return out
    hc = grid.from_tiles(tables.hc).astype(np.float64)
    dtab = grid.from_tiles(tables.dtab).astype(np.float64)
    for item in cfg.initial_levels:
        jj, ii = np.nonzero(inside(grid, item.coordinates))
        depth = np.asarray(hbar_at_level(item.level, tables.zmin[jj, ii], hc[jj, ii], dtab[jj, ii]))
        if not (depth > 0.0).any():
            raise ValueError(f"initial level {item.level} m wets no cell of its polygon: check the datum")
        out[jj, ii] = depth
    return out


# ---------------------------------------------------------------------------------------------------------------
# the monitor: what the run records as it steps (engine/monitor.py:89-300)


class Monitor:
    def __init__(self, cfg: RunConfig, tables, model, drv: Driver, results: Results, progress: Progress, out: Path,
                 stop_files: Iterable[Path] = ()) -> None:  # fmt: skip
        o = cfg.outputs
        self.cfg, self.tables, self.model, self.drv = cfg, tables, model, drv
        self.results, self.progress = results, progress
        self.stop_file = out / "stop"
        self.stop_files = (self.stop_file, *stop_files)
        self.started = time.monotonic()
        self.next_tick = self.started + o.progress_every_s
        self.next_check = self.started + cfg.run.checkpoint_every_s
        self.next_series = 0.0
        self.next_map = 0.0
        self.times: list[float] = []
        self.dt_min = np.inf
        self.checkpoint_due = False
        self.stop_due = False
        grid = tables.grid
        self.points, self.point_series = [], []
        for p in o.points:
            j, i = grid.cell_of(p.x, p.y)
            if not (0 <= j < grid.shape[0] and 0 <= i < grid.shape[1] and grid.mask[j, i]):
                progress.emit("warn", message=f"point {p.name} is outside the domain")
                continue
            k = np.asarray([state_index(grid, np.array(j), np.array(i))], np.int64)
            self.points.append((k, float(tables.zmin[j, i])))
            self.point_series.append((p.name, {"depth_m": [], "mean_depth_m": [], "level_m": []}))
        self.lines, self.line_series = [], []
        for ln in o.lines:
            xi, xs, yi, ys = line_faces(grid, ln.coordinates)
            if xi.size + yi.size == 0:
                progress.emit("warn", message=f"line {ln.name} crosses no active face")
            self.lines.append((xi, xs, yi, ys))
            self.line_series.append((ln.name, {"discharge_m3s": []}))

    def stop_requested(self) -> bool:
        return any(path.exists() for path in self.stop_files)

    def record(self, t: float) -> None:
        m = self.model
        self.times.append(t)
        for (index, zmin), (_, cols) in zip(self.points, self.point_series, strict=True):
            d = float(m.read_cells("d", index)[0])
            cols["depth_m"].append(d)
            cols["mean_depth_m"].append(float(m.read_cells("h", index)[0]))
            cols["level_m"].append(zmin + d)
        for (xi, xs, yi, ys), (_, cols) in zip(self.lines, self.line_series, strict=True):
            q = 0.0
            if xi.size:
                q += float(np.dot(m.read_cells("qx", xi).astype(np.float64), xs))
            if yi.size:
                q += float(np.dot(m.read_cells("qy", yi).astype(np.float64), ys))
            cols["discharge_m3s"].append(q)

    def snapshot(self, t: float) -> None:
        m = self.model
        fields = {"depth": m.field("d"), "mean_depth": m.field("h"), "speed": m.cell_speed()}
        if self.cfg.outputs.face_discharge:
            fields.update({"qx": m.field("qx"), "qy": m.field("qy")})
        self.results.snapshot(t, fields)
        self.progress.emit("output", kind="map", t=t)

    def tick(self, st, now: float) -> None:
        self.next_tick = now + self.cfg.outputs.progress_every_s
        bal = self.drv.accounts()
        self.progress.emit(
            "step", t=round(st.t, 3), dt=round(st.dt, 4), steps=st.nsteps, active_tiles=int(st.backend.nact),
            volume=round(bal.storage, 1), mass_error=float(f"{bal.relative:.3e}"), wall=round(now - self.started, 2),
        )  # fmt: skip
        if self.stop_requested():
            self.stop_due = True

    def on_step(self, st) -> None:
        self.drv.on_step(st)
        if st.dt < self.dt_min:
            self.dt_min = st.dt
        self.events(st)

    def events(self, st) -> None:
        """What a step's end may call for: a series sample, a map, a progress event, a checkpoint at the window's
        end (run.py:1731-1736)."""
        t = st.t
        if t >= self.next_series:
            self.record(t)
            self.next_series = t + self.cfg.outputs.series_every_s
        every = self.cfg.outputs.map_every_s
        if every > 0.0 and t >= self.next_map:
            self.snapshot(t)
            self.next_map = t + every
        now = time.monotonic()
        if now >= self.next_tick:
            self.tick(st, now)
        if self.cfg.run.checkpoint_every_s > 0.0 and now >= self.next_check:
            self.next_check = now + self.cfg.run.checkpoint_every_s
            self.checkpoint_due = True

    def maxima(self) -> dict[str, np.ndarray]:
        b = self.model.backend
        b.synchronize()
        return {name: from_state(self.tables.grid, b.maxima_get(k)) for k, name in enumerate(MAXIMA)}

    def finish(self, t: float) -> None:
        self.model.backend.maxima_update(t)
        self.record(t)
        self.drv.accounts()

    # checkpoint (io/checkpoint.py)
    def save(self, path: Path, key: str, t: float) -> None:
        b, st = self.model.backend, self.model.stepper
        b.synchronize()
        payload = {
            "job_hash": key, "t": t, "t_stepper": st.t, "dt": st.dt, "nsteps": st.nsteps, "refresh_due": st.refresh_due,
            "next_series": self.next_series, "next_map": self.next_map, "last_t": self.drv.last_t, "dry": self.drv.dry,
            "settled": self.drv.settled, "rain": self.drv.balance.rain, "storage0": self.drv.balance.storage0,
            "dt_min": float(self.dt_min), "times": np.asarray(self.times, np.float64),
            "idle": st.idle.copy(), "keep_active": st.keep_active.copy(), "active": np.asarray(b.active_tiles, np.int32),
            "tally": b.tally_totals(),
        }  # fmt: skip
        for name in SAVED_FIELDS:
            payload[f"state_{name}"] = b.get_field(name)
        for k, name in enumerate(MAXIMA):
            payload[f"maxima_{name}"] = b.maxima_get(k)
        for kind, series in (("points", self.point_series), ("lines", self.line_series)):
            for n, (_, cols) in enumerate(series):
                for q, values in cols.items():
                    payload[f"series_{kind}_{n}_{q}"] = np.asarray(values, np.float64)
        save_checkpoint(path, payload)

    def load(self, saved: dict) -> float:
        b, st = self.model.backend, self.model.stepper
        for name in SAVED_FIELDS:
            b.set_field(name, saved[f"state_{name}"])
        for k, name in enumerate(MAXIMA):
            b.maxima_put(k, saved[f"maxima_{name}"])
        b.tally_restore(saved["tally"])
        st.t, st.dt, st.nsteps = float(saved["t_stepper"]), float(saved["dt"]), int(saved["nsteps"])
        st.refresh_due = bool(saved["refresh_due"])
        st.idle[:] = saved["idle"]
        st.keep_active[:] = saved["keep_active"]
        b.set_active(saved["active"])
        from wadi2d_amd.solver.model import tiles_touched

        b.set_wet(tiles_touched(self.model.grid, self.model.field("h") > 0.0).astype(np.int32))
        self.next_series, self.next_map = float(saved["next_series"]), float(saved["next_map"])
        self.dt_min = float(saved["dt_min"])
        self.times = list(saved["times"])
        d = self.drv
        d.last_t, d.dry, d.settled = float(saved["last_t"]), bool(saved["dry"]), float(saved["settled"])
        d.balance.rain, d.balance.storage0 = float(saved["rain"]), float(saved["storage0"])
        for kind, series in (("points", self.point_series), ("lines", self.line_series)):
            for n, (_, cols) in enumerate(series):
                for q in cols:
                    cols[q] = list(saved[f"series_{kind}_{n}_{q}"])
        return float(saved["t"])


# ---------------------------------------------------------------------------------------------------------------
# the run


class _DeviceLoop:
    """The run's steps on the device (engine/device_run.py) in chunks that hold at the next series sample, map or
    window end: the device stops stepping the member once its clock has reached that time (w2d_run_pause), so the
    step that reaches it is the last one taken, and a chunk sized from the current dt rarely wastes a launch. The
    device keeps the books of every step (tally, maxima every maxima_every steps, rain, the smallest dt); after each
    chunk the host stepper and the driver take its state, so the monitor, the checkpoint and the summary read what
    the host loop would have left."""

    GRAPH_STEPS = 16
    MAX_CHUNK = 64

    def __init__(self, cfg: RunConfig, model, drv: Driver, monitor: Monitor, rain, bset, end: float) -> None:
        from wadi2d_amd.engine.device_run import DeviceRun, MemberPlan

        if not hasattr(model.backend, "run_setup"):
            raise InvalidJob("the device loop needs a HIP backend")
        if cfg.numerics.local_time_stepping:
            raise InvalidJob("the device loop steps uniformly: local time stepping runs with the host loop")
        self.cfg, self.model, self.drv, self.monitor = cfg, model, drv, monitor
        st = model.stepper
        plan = MemberPlan(1, model.grid.ntiles, rain, bset, st.t, end, st.dt, st.nsteps, st.refresh_due, rain.state,
                          dry=drv.dry)  # fmt: skip
        self.run = DeviceRun(model.backend, [plan], model.num, cfg.run.source_every_s,
                             maxima=cfg.outputs.maxima_every, idle=st.idle)  # fmt: skip
        self.rain0 = drv.balance.rain

    def chunk(self, target: float) -> None:
        st, drv = self.model.stepper, self.drv
        # the host loop looks after a step: an output due already is taken at the next step's end
        self.model.backend.run_pause([max(target, float(np.nextafter(st.t, np.inf)))])
        k = min(self.MAX_CHUNK, max(1, int(np.ceil((target - st.t) / max(st.dt, 1e-9))) + 1))
        s = self.run.steps(k, self.GRAPH_STEPS)[0]
        st.t, st.dt, st.nsteps, st.refresh_due = float(s.t), float(s.dt), int(s.nsteps), bool(s.refresh_due)
        st.idle[:] = self.model.backend.run_idle()
        drv.balance.rain = self.rain0 + float(s.rain)
        # the device has credited the rain: the driver's own window adds nothing when it settles
        drv.window, drv.settled, drv.last_t, drv.dry = (st.t, st.t + 1.0, 0.0), st.t, st.t, bool(s.dry)
        self.monitor.dt_min = min(self.monitor.dt_min, float(s.dt_min))
        if s.failed == 1:
            raise NumericalFailure(f"non-finite wave speed at t = {st.t:.3f} s")
        if s.failed:
            raise NumericalFailure(f"time step {st.dt:.2e} s below the minimum at t = {st.t:.3f} s")
        self.monitor.events(st)


def _run(cfg: RunConfig, out: Path, backend: str | None, resume: bool, progress: Progress, summary: RunSummary,
         stop_files, stop_at_s: float | None = None, loop: str = "host") -> None:  # fmt: skip
    from wadi2d_amd.solver.model import Model

    for name in ignored(cfg):
        progress.emit("warn", message=f"{name}: read past, it has no part in this path's run")
    curve = storm_curve(cfg)
    checkpoint = out / "checkpoint.npz"
    key = job_hash(cfg)
    saved = None
    if resume:
        saved = load_checkpoint(checkpoint) if checkpoint.exists() else None
        if saved is None:
            progress.emit("warn", message="no checkpoint to resume from: the run starts from the beginning")
        elif saved.get("job_hash") != key:
            raise InvalidJob("the checkpoint belongs to another job: it cannot continue this one")
    (out / "config.resolved.json").write_text(cfg.model_dump_json(indent=2), encoding="utf-8")
    quiet = curve is None and cfg.storm.kind != "gridded"
    if quiet and not cfg.initial_levels and not any(b.kind == "stage" for b in cfg.boundaries):
        progress.emit("warn", message="no rain, no inflow and no stage boundary: nothing will happen")
    t0 = time.monotonic()
    tables = build_grid(cfg)
    summary.tables_s = time.monotonic() - t0
    grid = tables.grid
    summary.cells = int(grid.mask.sum())
    name = backend or cfg.run.backend
    try:
        model = Model(tables, cfg.numerics, name)
    except (RuntimeError, ImportError, OSError) as error:
        raise BackendUnavailable(str(error)) from error
    summary.backend = model.backend.name if hasattr(model.backend, "name") else name
    bset = forcing.boundary_set(boundary_list(cfg, tables), tables)
    model.set_depth(initial_depth(cfg, tables))
    rain = build_rain(cfg, tables, bset, curve)
    drv = Driver(model, bset, rain, cfg.run.source_every_s, maxima_every=cfg.outputs.maxima_every)
    model.backend.maxima_set(cfg.outputs.arrival_depth, cfg.outputs.speed_depth)
    results = Results(out / "results", grid, cfg.model_dump_json(), append=saved is not None)
    monitor = Monitor(cfg, tables, model, drv, results, progress, out, stop_files)
    st = model.stepper
    t = 0.0
    if saved is not None:
        t = monitor.load(saved)
        progress.emit("resume", t=t, steps=st.nsteps)
    progress.emit("start", backend=summary.backend, cells=summary.cells, tiles=grid.ntiles - 1, dx=grid.dx,
                  duration_s=cfg.run.duration_s, tables_s=round(summary.tables_s, 2), t=t)  # fmt: skip
    t1 = time.monotonic()
    end = cfg.run.duration_s
    try:
        if monitor.stop_requested():
            raise StopRequested
        device = _DeviceLoop(cfg, model, drv, monitor, rain, bset, end) if loop == "device" else None
        while t < end - 1e-9:
            t_next = min(t + cfg.run.source_every_s, end)
            if device is None:
                drv.open_window(t, t_next)
                model.advance(t_next, monitor.on_step)
            else:
                while st.t < t_next - 1e-9:  # the window's end as the host loop's stepper reaches it
                    maps = monitor.next_map if cfg.outputs.map_every_s > 0.0 else np.inf
                    device.chunk(min(monitor.next_series, maps, t_next - 1e-9))
            drv.settle(t_next)
            t = t_next
            if monitor.checkpoint_due:
                monitor.checkpoint_due = False
                monitor.save(checkpoint, key, t)
            if (monitor.stop_due or (stop_at_s is not None and t >= stop_at_s - 1e-9)) and t < end - 1e-9:
                raise StopRequested
        summary.exit_code, summary.status = EXIT_DONE, "finished"
    except StopRequested:
        drv.settle(st.t)
        monitor.save(checkpoint, key, t)
        summary.exit_code, summary.status = EXIT_STOPPED, "stopped"
    except NumericalFailure as error:
        summary.exit_code, summary.status = EXIT_NUMERICAL, "numerical failure"
        summary.message = str(error)
        progress.emit("error", code=EXIT_NUMERICAL, message=str(error))
    _finish(cfg, tables, model, monitor, results, out, progress, summary, t1)


def _finish(cfg, tables, model, monitor: Monitor, results: Results, out: Path, progress: Progress, summary, t1) -> None:
    """Final samples, the maxima, the series and the books (run.py:1821-1859)."""
    st = model.stepper
    monitor.finish(st.t)
    monitor.stop_file.unlink(missing_ok=True)
    maxima = monitor.maxima()
    results.maxima(maxima)
    progress.emit("output", kind="maxima", names=list(maxima))
    times = np.asarray(monitor.times)
    for kind, series in (("points", monitor.point_series), ("lines", monitor.line_series)):
        for name, cols in series:
            results.series(kind, name, times, cols)
            write_series_csv(out / "series" / f"{kind}_{safe_name(name)}.csv", times, cols)
    summary.loop_s = time.monotonic() - t1
    summary.t_end, summary.steps = st.t, st.nsteps
    summary.dt_min = float(monitor.dt_min) if np.isfinite(monitor.dt_min) else 0.0
    wet = int(np.count_nonzero(maxima["mean_depth"] >= cfg.outputs.wet_depth))
    summary.wet_area_km2 = wet * tables.grid.dx**2 / M2_PER_KM2
    for name, cols in monitor.line_series:
        q = np.abs(np.asarray(cols["discharge_m3s"]))
        if q.size:
            summary.peaks[name] = (float(q.max()), float(times[int(np.argmax(q))]))
    b = monitor.drv.balance
    summary.balance = b
    balance = {"rain_m3": b.rain, "inflow_m3": b.inflow, "outflow_m3": b.outflow, "storage_start_m3": b.storage0,
               "storage_end_m3": b.storage, "error_m3": b.error, "relative_error": b.relative}  # fmt: skip
    results.attributes({"status": summary.status, "t_end_s": st.t, "steps": st.nsteps, **balance})
    summary.done = dict(
        status=summary.status, code=summary.exit_code, partial=summary.status != "finished", t=st.t,
        duration_s=cfg.run.duration_s, steps=st.nsteps, loop_s=round(summary.loop_s, 2),
        tables_s=round(summary.tables_s, 2), wet_area_km2=round(summary.wet_area_km2, 4), dt_min=summary.dt_min,
        balance=balance, peaks={k: {"q_m3s": v[0], "t_s": v[1]} for k, v in summary.peaks.items()},
    )  # fmt: skip


def run_job(cfg: RunConfig, out: str | Path, backend: str | None = None, resume: bool = False, echo: bool = False,
            stop_files: Iterable[Path] = (), stop_at_s: float | None = None, loop: str = "host") -> RunSummary:  # fmt: skip
    """Run one job into folder ``out`` (run.py:1862-1901). ``out/stop`` stops it at a window's end, and so does any
    of ``stop_files``; a stop that was there first stops it at once. ``stop_at_s`` (no counterpart in the reference)
    stops it as a stop file would, at the first window end at or after that simulated time: a scheduler's slice.
    ``loop`` "device" (no counterpart) steps on the device between outputs (``_DeviceLoop``): the same steps and
    results as the host loop, fewer host round trips."""
    folder = Path(out)
    folder.mkdir(parents=True, exist_ok=True)
    summary = RunSummary()
    with Progress(folder / "progress.jsonl", echo=echo, append=resume) as progress:
        try:
            _run(cfg, folder, backend, resume, progress, summary, tuple(Path(p) for p in stop_files), stop_at_s, loop)
            if summary.done:
                progress.emit("done", **summary.done)
        except JobError as error:
            summary.exit_code, summary.message = error.code, str(error)
            progress.emit("error", code=error.code, message=str(error))
        except FileNotFoundError as error:
            summary.exit_code, summary.message = EXIT_INPUT, str(error)
            progress.emit("error", code=EXIT_INPUT, message=str(error))
        except ValueError as error:
            summary.exit_code, summary.message = EXIT_INVALID, str(error)
            progress.emit("error", code=EXIT_INVALID, message=str(error))
        except Exception as error:  # the protocol wants a code and an event, never a traceback alone
            memory = isinstance(error, MemoryError)
            summary.exit_code = EXIT_MEMORY if memory else EXIT_UNEXPECTED
            summary.message = str(error)
            progress.emit("error", code=summary.exit_code, message=str(error), trace=traceback.format_exc())
    return summary


__all__ = ["RunSummary", "run_job", "storm_curve"]
