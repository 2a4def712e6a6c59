"""``wadi2d run``: the job file, the DEM, the lines, the exit codes and progress events (wadi2d/run.py,
wadi2d/engine/cli.py, tests/engine/test_run.py and test_resume_exact.py of the reference are the model).

CPU tests cover what needs no device: the job schema (a reference-shaped job loads; what this path does not run is
refused with exit 2), the DEM readers, the cells and faces under lines, the storm curves, and the exit codes of a run
that cannot start (3 input missing, 4 no device). GPU tests run a small job end to end against the oracle driven
through the same window loop (bit for bit in the exact float32 variant) and check that a stopped and resumed run
equals the unbroken one bit for bit.
"""

from __future__ import annotations

import json
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from wadi2d_amd import job as jobmod
from wadi2d_amd.dem import grid_over, read_raster
from wadi2d_amd.engine.lines import cells_along, inside, line_faces
from wadi2d_amd.grid import TileGrid

ROOT = Path(__file__).resolve().parent.parent
X0, Y0 = 1000.0, 2000.0  # south-west corner of the test DEM
NY, NX, K, DX = 64, 96, 2, 10.0  # cells, pixels per cell, cell size


def ground(seed: int = 3) -> np.ndarray:
    """Pixels (rows north to south) of a plane falling 1 % east with roughness and a wall across x = 480 m with a
    30 m gap: NY*K x NX*K pixels of DX/K m."""
    rng = np.random.default_rng(seed)
    p = DX / K
    x = (np.arange(NX * K) + 0.5) * p
    y = (np.arange(NY * K) + 0.5) * p
    z = 100.0 - 0.01 * x[None, :] + 0.05 * rng.standard_normal((y.size, x.size))
    wall = (x >= 480.0) & (x < 490.0)
    gap = np.abs(y - 320.0) < 15.0
    z[np.ix_(~gap, wall)] += 1.0
    return z


def write_asc(path: Path, z: np.ndarray, pixel: float, nodata: float | None = None, center: bool = False) -> None:
    ny, nx = z.shape
    with path.open("w", encoding="ascii") as f:
        f.write(f"ncols {nx}\nnrows {ny}\n")
        if center:
            f.write(f"xllcenter {X0 + 0.5 * pixel}\nyllcenter {Y0 + 0.5 * pixel}\n")
        else:
            f.write(f"xllcorner {X0}\nyllcorner {Y0}\n")
        f.write(f"cellsize {pixel}\n")
        if nodata is not None:
            f.write(f"NODATA_value {nodata}\n")
        np.savetxt(f, z, fmt="%.4f")


def job_dict(dem: str = "ground.asc", duration: float = 600.0) -> dict:
    xe, yn = X0 + NX * DX, Y0 + NY * DX
    return {
        "project": {"name": "test", "description": ""},
        "grid": {"dx": DX},
        "terrain": {"dem": dem, "manning": 0.04},
        "tables": {"rows": 8},
        "numerics": {},
        "storm": {"kind": "uniform", "depth_mm": 40.0, "duration_min": 20.0},
        "boundaries": [{"kind": "free", "coordinates": [[xe - 1.0, yn], [xe - 1.0, Y0]], "name": "east"}],
        "initial_levels": [{"coordinates": [[X0, yn], [X0 + 200.0, yn], [X0 + 200.0, yn - 200.0], [X0, yn - 200.0]],
                            "level": 100.2}],
        "outputs": {"maxima_every": 1, "series_every_s": 30.0, "map_every_s": 300.0, "progress_every_s": 0.5,
                    "points": [{"name": "p1", "x": X0 + 300.0, "y": Y0 + 320.0}],
                    "lines": [{"name": "gap", "coordinates": [[X0 + 470.0, Y0], [X0 + 470.0, yn]]}]},
        "run": {"duration_s": duration, "source_every_s": 30.0, "checkpoint_every_s": 0.0},
    }  # fmt: skip


@pytest.fixture()
def job_folder(tmp_path):
    write_asc(tmp_path / "ground.asc", ground(), DX / K)
    (tmp_path / "job.json").write_text(json.dumps(job_dict()), encoding="utf-8")
    return tmp_path


# ---- the job file


def test_a_job_loads_with_its_dem_path_made_absolute(job_folder):
    cfg = jobmod.load_job(job_folder / "job.json")
    assert Path(cfg.terrain.dem) == (job_folder / "ground.asc").resolve()
    assert cfg.tables.rows == 8 and cfg.numerics.cfl == 0.45 and cfg.run.source_every_s == 30.0


def test_a_reference_shaped_job_with_inert_sections_loads(job_folder):
    d = job_dict()
    d["mesh"] = {"enabled": False, "coarse_dx": 40.0}
    d["strips"] = {"enabled": False}
    d["hydrology"] = {"subbasins": [], "hms_basin": None, "time_step_min": 5.0}
    d["transmission"] = {"enabled": False, "rate_mm_h": 10.0}
    d["outputs"]["live_every_s"] = 5.0
    d["outputs"]["pixels"] = {"enabled": False}
    d["numerics"] = {"local_time_stepping": False, "lts_levels": 3}
    d["terrain"]["join_cells"] = 2.0
    d["initial_points"], d["dam_levels"], d["breaches"] = [], [], []
    (job_folder / "job.json").write_text(json.dumps(d), encoding="utf-8")
    cfg = jobmod.load_job(job_folder / "job.json")
    said = jobmod.ignored(cfg)
    assert "outputs.live_every_s" in said and "terrain.join_cells" in said and "hydrology.time_step_min" in said
    assert not any(x.startswith("hydrology.subbasins") or x.startswith("hydrology.hms") for x in said)


@pytest.mark.parametrize(
    "edit, named",
    [
        (lambda d: d["terrain"].update(objects=[{"kind": "crest"}]), "terrain.objects"),
        (lambda d: d.update(mesh={"enabled": True}), "mesh.enabled"),
        (lambda d: d["boundaries"].append({"kind": "flow", "coordinates": [[0, 0], [1, 1]], "name": "in",
                                           "flow": {"t": [0], "q": [1]}}), "flow"),  # fmt: skip
        (lambda d: d["storm"].update(kind="gridded"), "gridded"),
        (lambda d: d["storm"].update(kind="gauges"), "gauges"),
        (lambda d: d["numerics"].update(local_time_stepping=True), "numerics.local_time_stepping"),
        (lambda d: d.update(breaches=[{"name": "b"}]), "breaches"),
        (lambda d: d["grid"].update(domain="outline.gpkg"), "grid.domain"),
    ],
)
def test_what_this_path_does_not_run_is_refused_as_an_invalid_job(job_folder, edit, named):
    d = job_dict()
    edit(d)
    (job_folder / "job.json").write_text(json.dumps(d), encoding="utf-8")
    with pytest.raises(jobmod.InvalidJob) as info:
        jobmod.load_job(job_folder / "job.json")
    assert info.value.code == 2 and named in str(info.value)


def test_a_bad_value_or_a_missing_file_has_its_exit_code(tmp_path):
    with pytest.raises(jobmod.InputMissing) as info:
        jobmod.load_job(tmp_path / "nothing.json")
    assert info.value.code == 3
    d = job_dict()
    d["numerics"]["cfl"] = 5.0  # above the schema's bound
    (tmp_path / "job.json").write_text(json.dumps(d), encoding="utf-8")
    with pytest.raises(jobmod.InvalidJob):
        jobmod.load_job(tmp_path / "job.json")
    (tmp_path / "job.json").write_text("{not json", encoding="utf-8")
    with pytest.raises(jobmod.InvalidJob):
        jobmod.load_job(tmp_path / "job.json")


def test_the_job_hash_follows_the_water_and_not_the_outputs(job_folder):
    a = jobmod.load_job(job_folder / "job.json")
    b = a.model_copy(update={"outputs": a.outputs.model_copy(update={"map_every_s": 60.0})})
    c = a.model_copy(update={"terrain": a.terrain.model_copy(update={"manning": 0.05})})
    assert jobmod.job_hash(a) == jobmod.job_hash(b)
    assert jobmod.job_hash(a) != jobmod.job_hash(c)


# ---- the DEM


def test_asc_and_npz_give_the_same_grid_and_no_data_leaves_cells_out(tmp_path):
    z = ground()
    write_asc(tmp_path / "a.asc", z, DX / K)
    write_asc(tmp_path / "c.asc", z, DX / K, center=True)
    np.savez(tmp_path / "b.npz", z=z, x0=X0, y0=Y0 + NY * DX, pixel=DX / K)
    ra, rb, rc = read_raster(tmp_path / "a.asc"), read_raster(tmp_path / "b.npz"), read_raster(tmp_path / "c.asc")
    for r in (rb, rc):
        assert np.allclose(r.z, ra.z, atol=1e-4) and r.x0 == pytest.approx(ra.x0) and r.y0 == pytest.approx(ra.y0)
    grid, bed = grid_over(ra, DX)
    assert grid.mask.sum() == NY * NX
    assert grid.x0 == pytest.approx(X0 - DX) and grid.y0 == pytest.approx(Y0 + NY * DX + DX)
    # the bed's pixel (r, c) of the padded grid is raster pixel (r - K, c - K)
    assert np.array_equal(bed.read(K, K, 3, 3), ra.z[:3, :3])
    zn = z.copy()
    zn[5, 7] = -9999.0  # cell (2, 3)
    write_asc(tmp_path / "n.asc", zn, DX / K, nodata=-9999.0)
    grid_n, _ = grid_over(read_raster(tmp_path / "n.asc"), DX)
    assert grid_n.mask.sum() == NY * NX - 1 and not grid_n.mask[1 + 2, 1 + 3]


def test_an_extent_snaps_inward_and_dx_must_be_a_pixel_multiple(tmp_path):
    write_asc(tmp_path / "a.asc", ground(), DX / K)
    r = read_raster(tmp_path / "a.asc")
    grid, _ = grid_over(r, DX, (X0 + 101.0, Y0 + 52.0, X0 + 400.0, Y0 + 300.0))
    assert grid.x0 == pytest.approx(X0 + 105.0 - DX)
    assert int(grid.mask.sum()) == 29 * 24
    with pytest.raises(ValueError):
        grid_over(r, 7.0)
    with pytest.raises(FileNotFoundError):
        read_raster(tmp_path / "none.asc")
    (tmp_path / "g.tif").write_bytes(b"II*\x00")
    with pytest.raises(ValueError):
        read_raster(tmp_path / "g.tif")


# ---- lines


def _grid(ny: int = 40, nx: int = 50) -> TileGrid:
    return TileGrid.from_mask(np.ones((ny, nx), bool), 0.0, ny * 10.0, 10.0)


def test_a_line_along_the_edge_opens_the_edge_cells():
    g = _grid()
    jj, ii = cells_along(g, [(505.0, 400.0), (505.0, 0.0)], samples=4, snap_m=150.0)
    assert set(ii.tolist()) == {50} and np.array_equal(np.unique(jj), np.arange(1, 41))


def test_a_polygon_takes_the_cells_whose_centre_it_holds():
    g = _grid()
    m = inside(g, [(0.0, 400.0), (100.0, 400.0), (100.0, 300.0), (0.0, 300.0)])
    assert m.sum() == 100 and m[1:11, 1:11].all()
    tri = inside(g, [(0.0, 0.0), (200.0, 0.0), (0.0, 200.0)])
    j, i = np.nonzero(tri)
    x, y = g.xy_of(j, i)
    assert np.all(x + y < 200.0) and tri.sum() == 190


def test_line_faces_follow_the_line_with_the_discharge_to_its_right():
    g = _grid()
    # drawn south to north on x = 200 m: west faces of column 21 (x 200..210 m), positive to the east
    xi, xs, yi, ys = line_faces(g, [(200.0, 0.0), (200.0, 400.0)])
    assert yi.size == 0 and xi.size == 40 and np.all(xs == 1.0)
    xi2, xs2, _, _ = line_faces(g, [(200.0, 400.0), (200.0, 0.0)])
    assert np.array_equal(np.sort(xi2), np.sort(xi)) and np.all(xs2 == -1.0)
    # drawn west to east on y = 200 m: north faces of row 21, positive to the south
    xi, xs, yi, ys = line_faces(g, [(0.0, 200.0), (500.0, 200.0)])
    assert xi.size == 0 and yi.size == 50 and np.all(ys == 1.0)
    # a diagonal is a staircase of both kinds with one face per lattice step
    xi, xs, yi, ys = line_faces(g, [(0.0, 0.0), (300.0, 300.0)])
    assert xi.size == 30 and yi.size == 30


# ---- storms


def test_storm_curves():
    from wadi2d_amd.run import storm_curve

    def cfg(**storm):
        return jobmod.RunConfig.model_validate({"storm": storm})

    u = storm_curve(cfg(kind="uniform", depth_mm=60.0, duration_min=60.0, arf_factor=0.5))
    assert u(0.0) == 0.0 and u(1800.0) == pytest.approx(15.0) and u(7200.0) == pytest.approx(30.0)
    h = storm_curve(cfg(kind="hyetograph", hyetograph_min=[10.0, 20.0], hyetograph_mm=[5.0, 15.0]))
    assert h(300.0) == pytest.approx(2.5) and h(900.0) == pytest.approx(12.5) and h(5000.0) == pytest.approx(20.0)
    p = storm_curve(cfg(kind="hypothetical", depth_mm=100.0, duration_min=100.0, time_percent=[50.0, 100.0],
                        depth_percent=[80.0, 100.0]))  # fmt: skip
    assert p(1500.0) == pytest.approx(40.0) and p(4500.0) == pytest.approx(90.0) and p(9000.0) == pytest.approx(100.0)
    assert storm_curve(cfg(kind="none")) is None


# ---- exit codes without a device


def _no_device() -> bool:
    from wadi2d_amd.solver.hip_backend import device_count

    return device_count() == 0


def test_a_run_without_a_device_ends_in_exit_4_with_an_error_event(job_folder):
    if not _no_device():
        pytest.skip("a HIP device is present")
    from wadi2d_amd.run import run_job

    cfg = jobmod.load_job(job_folder / "job.json")
    s = run_job(cfg, job_folder / "out")
    assert s.exit_code == 4
    events = [json.loads(x) for x in (job_folder / "out" / "progress.jsonl").read_text().splitlines()]
    assert events[-1]["ev"] == "error" and events[-1]["code"] == 4
    assert (job_folder / "out" / "config.resolved.json").exists()


def test_the_command_line_returns_the_protocol_codes(job_folder):
    cmd = [sys.executable, "-m", "wadi2d_amd", "run"]
    r = subprocess.run([*cmd, str(job_folder / "none.json"), "--out", str(job_folder / "o")], capture_output=True,
                       text=True, cwd=ROOT)  # fmt: skip
    assert r.returncode == 3 and json.loads(r.stderr)["code"] == 3
    d = job_dict()
    d["terrain"]["objects"] = [{"kind": "wall"}]
    (job_folder / "bad.json").write_text(json.dumps(d), encoding="utf-8")
    r = subprocess.run([*cmd, str(job_folder / "bad.json"), "--out", str(job_folder / "o")], capture_output=True,
                       text=True, cwd=ROOT)  # fmt: skip
    assert r.returncode == 2
    r = subprocess.run([*cmd, str(job_folder / "job.json"), "--out", str(job_folder / "o"), "--dry-run"],
                       capture_output=True, text=True, cwd=ROOT)  # fmt: skip
    assert r.returncode == 0 and json.loads(r.stdout)["cells"] == NY * NX
    if _no_device():
        r = subprocess.run([*cmd, str(job_folder / "job.json"), "--out", str(job_folder / "o")], capture_output=True,
                           text=True, cwd=ROOT)  # fmt: skip
        assert r.returncode == 4


# ---- runs on the device


def _oracle_run(cfg, full: bool = False):
    """The same job through the oracle, in the window loop of run._run (run.py:1712-1739); with ``full`` its five
    running maxima sampled every ``outputs.maxima_every`` steps are returned too."""
    from oracle.oracle import Ledger, OracleModel
    from oracle.oracle import Numerics as ONumerics
    from wadi2d_amd.engine.forcing import boundary_set
    from wadi2d_amd.run import boundary_list, build_grid, build_rain, initial_depth, storm_curve
    from wadi2d_amd.solver.state import from_state

    tables = build_grid(cfg)
    bset = boundary_set(boundary_list(cfg, tables), tables)
    rain = build_rain(cfg, tables, bset, storm_curve(cfg))
    if cfg.numerics.local_time_stepping:
        from oracle.levels import OracleLevelModel as OracleModel  # noqa: F811
    m = OracleModel(tables, ONumerics.like(cfg.numerics), prec=4, threads=4)
    m.set_depth(initial_depth(cfg, tables))
    if bset.size:
        m.set_boundary(bset.cells)
    ledger = Ledger(m, tables.grid.mask & ~bset.mask, bset.xfaces, bset.xsign, bset.yfaces, bset.ysign,
                    maxima_every=cfg.outputs.maxima_every if full else 0, arrive=cfg.outputs.arrival_depth,
                    deep=cfg.outputs.speed_depth)  # fmt: skip
    t, dry, end = 0.0, False, cfg.run.duration_s
    while t < end - 1e-9:
        t_next = min(t + cfg.run.source_every_s, end)
        rate, vol = rain.rate(t, t_next)
        if vol > 0.0 or not dry:
            m.set_source_cells(rain.state, rate)
        dry = vol <= 0.0
        ledger.open_window(t, t_next, vol, 0.0)
        m.advance(t_next, ledger.on_step)
        ledger.settle(t_next)
        t = t_next
    if full:
        ledger.maxima.update(t)  # the final sample a run takes as it finishes (monitor.py:298-300)
        maxima = {n: from_state(tables.grid, a) for n, a in zip(ledger.maxima.NAMES, ledger.maxima.arrays, strict=True)}
        return maxima, ledger.accounts(t), m.stepper.nsteps
    return from_state(tables.grid, ledger.dmax), ledger.accounts(t), m.stepper.nsteps


def _kinds_job(folder: Path) -> dict:
    """The test job with a stage and a normal boundary beside the free one and a hyetograph storm."""
    xe, yn = X0 + NX * DX, Y0 + NY * DX
    d = job_dict(duration=600.0)
    d["storm"] = {"kind": "hyetograph", "hyetograph_min": [2.0, 4.0, 12.0], "hyetograph_mm": [10.0, 20.0, 25.0]}
    d["boundaries"] = [
        {"kind": "free", "coordinates": [[xe - 1.0, yn], [xe - 1.0, Y0 + 320.0]], "name": "east top"},
        {"kind": "normal", "slope": 0.01, "coordinates": [[xe - 1.0, Y0 + 310.0], [xe - 1.0, Y0]], "name": "east"},
        {"kind": "stage", "stage": 100.5, "coordinates": [[X0 + 1.0, Y0 + 100.0], [X0 + 1.0, Y0]], "name": "west"},
    ]
    d["outputs"]["maxima_every"] = 4
    (folder / "job.json").write_text(json.dumps(d), encoding="utf-8")
    return d


@pytest.mark.gpu
@pytest.mark.parametrize("loop", ["host", "device"])
def test_stage_and_normal_boundaries_and_every_maximum_equal_the_oracle_bit_for_bit(job_folder, loop):
    """Boundary kinds 0 (stage) and 2 (normal, the inner cell's level lowered by S dx; forcing.py:299-332,
    cells.cu:95-111) beside a free one, through the exact float32 HIP path and the oracle's float32 statement order;
    the five running maxima (maxima.py:29-51), sampled every 4 steps as a job does, equal bit for bit."""
    from wadi2d_amd.run import MAXIMA, run_job

    _kinds_job(job_folder)
    cfg = jobmod.load_job(job_folder / "job.json")
    s = run_job(cfg, job_folder / "kinds", backend="hip-exact", loop=loop)
    assert s.exit_code == 0, s.message
    want, bal, steps = _oracle_run(cfg, full=True)
    print(f"{loop}: {s.steps} steps (oracle {steps}), mass {s.balance.relative:.2e} / {bal.relative:.2e}", flush=True)
    assert s.steps == steps
    for name in MAXIMA:
        got = np.load(job_folder / "kinds" / "results" / "maxima" / f"{name}.npy")
        ref = want[name].astype(np.float32)
        assert np.array_equal(got, ref), f"{name}: {np.abs(got - ref).max()}"
    assert (want["speed"] > 0.1).sum() > 100 and (want["arrival"] > 0).sum() > 100  # the maxima are not trivial
    assert abs(s.balance.relative - bal.relative) <= 1e-9


@pytest.mark.gpu
def test_a_job_runs_end_to_end_bit_for_bit_the_oracle(job_folder):
    from wadi2d_amd.run import run_job

    cfg = jobmod.load_job(job_folder / "job.json")
    s = run_job(cfg, job_folder / "out", backend="hip-exact")
    print(f"run: exit {s.exit_code} steps {s.steps} mass {s.balance.relative:.2e} peaks {s.peaks}", flush=True)
    assert s.exit_code == 0, s.message
    dmax, bal, steps = _oracle_run(cfg)
    got = np.load(job_folder / "out" / "results" / "maxima" / "depth.npy")
    assert s.steps == steps
    assert np.array_equal(got, dmax.astype(np.float32))
    assert (got > 0.05).sum() > 500  # the reservoir and the storm wet the domain
    assert abs(s.balance.relative) <= 1e-6 and abs(s.balance.relative - bal.relative) <= 1e-9
    events = [json.loads(x) for x in (job_folder / "out" / "progress.jsonl").read_text().splitlines()]
    kinds = [e["ev"] for e in events]
    assert kinds[0] in ("warn", "start") and "start" in kinds and kinds[-1] == "done"
    assert events[-1]["code"] == 0 and events[-1]["steps"] == s.steps
    assert (job_folder / "out" / "series" / "lines_gap.csv").exists()
    maps = sorted((job_folder / "out" / "results" / "maps").iterdir())
    assert len(maps) >= 2 and (maps[0] / "depth.npy").exists()


@pytest.mark.gpu
@pytest.mark.skip(reason="its one GPU run faulted the device (an illegal address reported at the cycle's wave read); "
                  "the cause is not found, so a job refuses local time stepping and this case stays off the GPU")
def test_a_job_with_local_time_stepping_runs_bit_for_bit_the_oracle_level_stepper(job_folder):
    """numerics.local_time_stepping: the run takes the level stepper (model.py:23-29), its boundary tiles pinned and
    its tally counted every smallest step; the maxima equal the oracle's level stepper's bit for bit."""
    from wadi2d_amd.run import run_job

    d = json.loads((job_folder / "job.json").read_text())
    d.setdefault("numerics", {}).update(local_time_stepping=True, lts_levels=2)
    (job_folder / "job.json").write_text(json.dumps(d))
    cfg = jobmod.load_job(job_folder / "job.json")
    s = run_job(cfg, job_folder / "out", backend="hip-exact")
    print(f"LTS run: exit {s.exit_code} steps {s.steps} mass {s.balance.relative:.2e}", flush=True)
    assert s.exit_code == 0, s.message
    dmax, _bal, steps = _oracle_run(cfg)
    got = np.load(job_folder / "out" / "results" / "maxima" / "depth.npy")
    assert s.steps == steps
    assert np.array_equal(got, dmax.astype(np.float32))
    assert abs(s.balance.relative) <= 1e-6


def _loss_job(folder: Path, holes: bool = False, gridded: bool = False) -> dict:
    """The test job with a curve number (80), a CN raster over the west half at cell resolution (a no-data patch
    falling back to 80, unless ``holes`` drops the default) and a depth-grid storm rising from 20 to 70 mm eastward,
    or with ``gridded`` rain frames of 100 m pixels: four 5-minute periods of 0-12 mm with a 5-minute gap."""
    cn = np.full((NY, NX), -9999.0)
    cn[:, : NX // 2] = 60.0 + 30.0 * np.random.default_rng(4).random((NY, NX // 2))
    cn[10:20, 5:15] = -9999.0
    write_asc(folder / "cn.asc", cn, DX, nodata=-9999.0)
    depth = np.tile(np.linspace(20.0, 70.0, NX), (NY, 1))
    write_asc(folder / "depth.asc", depth, DX)
    d = job_dict()
    d["terrain"]["cn_raster"] = "cn.asc"
    d["storm"] = {"kind": "depth_grid", "depth_grid": "depth.asc", "duration_min": 20.0}
    if gridded:
        rng = np.random.default_rng(8)
        cols, rows = NX * DX // 100 + 2, NY * DX // 100 + 2
        np.savez(folder / "rain.npz", depths=rng.uniform(0.0, 12.0, (4, int(rows), int(cols))),
                 starts_min=np.array([0.0, 5.0, 15.0, 20.0]), ends_min=np.array([5.0, 10.0, 20.0, 25.0]),
                 west=X0 - 50.0, north=Y0 + NY * DX + 50.0, dx=100.0, dy=100.0)  # fmt: skip
        d["storm"] = {"kind": "gridded", "rain_file": "rain.npz", "arf_factor": 0.8}
    if not holes:
        d["hydrology"] = {"curve_number": 80.0}
    (folder / "job.json").write_text(json.dumps(d), encoding="utf-8")
    return d


def test_curve_numbers_and_a_depth_grid_reach_the_rain(job_folder):
    from wadi2d_amd.engine.forcing import boundary_set
    from wadi2d_amd.run import boundary_list, build_grid, build_rain, storm_curve

    _loss_job(job_folder)
    cfg = jobmod.load_job(job_folder / "job.json")
    assert Path(cfg.terrain.cn_raster).is_absolute() and Path(cfg.storm.depth_grid).is_absolute()
    assert not [x for x in jobmod.ignored(cfg) if x.startswith("hydrology")]
    tables = build_grid(cfg)
    bset = boundary_set(boundary_list(cfg, tables), tables)
    rain = build_rain(cfg, tables, bset, storm_curve(cfg))
    assert rain.excess_form
    x, _ = tables.grid.xy_of(rain.jj, rain.ii)
    west = x < X0 + NX * DX / 2
    assert np.all((rain.cn[west] >= 60.0) & (rain.cn[west] <= 90.0) | (rain.cn[west] == 80.0))
    assert np.all(rain.cn[~west] == 80.0) and np.any(rain.cn[west] != 80.0)
    col = np.floor((x - X0) / DX).astype(int)
    assert np.array_equal(rain.depths, np.linspace(20.0, 70.0, NX)[col].round(4))
    assert storm_curve(cfg)(600.0) == pytest.approx(0.5)  # the unit storm, half way


def test_gridded_rain_reaches_the_rain_and_bad_files_are_refused(job_folder):
    from wadi2d_amd.engine.forcing import boundary_set
    from wadi2d_amd.run import boundary_list, build_grid, build_rain, storm_curve

    _loss_job(job_folder, gridded=True)
    cfg = jobmod.load_job(job_folder / "job.json")
    assert Path(cfg.storm.rain_file).is_absolute() and storm_curve(cfg) is None
    tables = build_grid(cfg)
    bset = boundary_set(boundary_list(cfg, tables), tables)
    rain = build_rain(cfg, tables, bset, storm_curve(cfg))
    assert rain.pixels is not None and rain.pixels.factor == 0.8 and rain.pixels.used.size > 50
    assert rain.rate(0.0, 300.0)[1] > 0.0 and rain.rate(600.0, 900.0)[1] == 0.0  # the gap
    (job_folder / "rain.nc").write_bytes(b"CDF\x01")
    for name, error in (("rain.nc", ValueError), ("absent.npz", FileNotFoundError)):
        d = json.loads((job_folder / "job.json").read_text())
        d["storm"]["rain_file"] = name
        (job_folder / "job.json").write_text(json.dumps(d), encoding="utf-8")
        cfg = jobmod.load_job(job_folder / "job.json")
        with pytest.raises(error):
            build_rain(cfg, tables, bset, None)


def test_a_loss_job_with_holes_or_bad_settings_is_refused(job_folder):
    from wadi2d_amd.engine.forcing import boundary_set
    from wadi2d_amd.run import boundary_list, build_grid, build_rain, storm_curve

    _loss_job(job_folder, holes=True)
    cfg = jobmod.load_job(job_folder / "job.json")
    tables = build_grid(cfg)
    bset = boundary_set(boundary_list(cfg, tables), tables)
    with pytest.raises(ValueError, match="cn_raster gives no CN"):
        build_rain(cfg, tables, bset, storm_curve(cfg))
    for edit in (lambda d: d.update(hydrology={"curve_number": 0.0}),
                 lambda d: d["storm"].update(depth_mm=10.0),
                 lambda d: d.update(storm={"kind": "depth", "depth_mm": 0.0})):  # fmt: skip
        d = _loss_job(job_folder)
        edit(d)
        (job_folder / "job.json").write_text(json.dumps(d), encoding="utf-8")
        with pytest.raises(jobmod.InvalidJob):
            jobmod.load_job(job_folder / "job.json")


@pytest.mark.gpu
@pytest.mark.parametrize("gridded", [False, True])
def test_a_job_with_losses_and_a_depth_grid_or_gridded_rain_runs_bit_for_bit_the_oracle(job_folder, gridded):
    from wadi2d_amd.run import run_job

    _loss_job(job_folder, gridded=gridded)
    cfg = jobmod.load_job(job_folder / "job.json")
    s = run_job(cfg, job_folder / "out", backend="hip-exact")
    print(f"loss run (gridded {gridded}): exit {s.exit_code} steps {s.steps} mass {s.balance.relative:.2e} "
          f"rain {s.balance.rain:.1f}", flush=True)
    assert s.exit_code == 0, s.message
    dmax, bal, steps = _oracle_run(cfg)
    got = np.load(job_folder / "out" / "results" / "maxima" / "depth.npy")
    assert s.steps == steps
    assert np.array_equal(got, dmax.astype(np.float32))
    assert abs(s.balance.relative) <= 1e-6 and s.balance.rain == pytest.approx(bal.rain, rel=1e-9)
    assert 0.0 < s.balance.rain


@pytest.mark.gpu
def test_a_stopped_and_resumed_run_is_bit_for_bit_the_unbroken_one(job_folder):
    from wadi2d_amd.run import run_job

    cfg = jobmod.load_job(job_folder / "job.json")
    whole = run_job(cfg, job_folder / "whole", backend="hip-exact")
    first = run_job(cfg, job_folder / "split", backend="hip-exact", stop_at_s=240.0)
    assert first.exit_code == 6 and (job_folder / "split" / "checkpoint.npz").exists()
    assert first.t_end == pytest.approx(240.0)
    second = run_job(cfg, job_folder / "split", backend="hip-exact", resume=True)
    print(f"whole {whole.steps} steps, split {first.steps} + resumed to {second.steps}", flush=True)
    assert whole.exit_code == 0 and second.exit_code == 0
    assert second.steps == whole.steps
    for name in ("depth", "mean_depth", "speed", "hazard", "arrival"):
        a = np.load(job_folder / "whole" / "results" / "maxima" / f"{name}.npy")
        b = np.load(job_folder / "split" / "results" / "maxima" / f"{name}.npy")
        assert np.array_equal(a, b), name
    assert second.balance.relative == whole.balance.relative
    for f in ("lines_gap.csv", "points_p1.csv"):
        assert (job_folder / "whole" / "series" / f).read_text() == (job_folder / "split" / "series" / f).read_text()
    # a checkpoint of one job does not continue another
    other = cfg.model_copy(update={"terrain": cfg.terrain.model_copy(update={"manning": 0.05})})
    s = run_job(other, job_folder / "split", backend="hip-exact", resume=True)
    assert s.exit_code == 2


def _same_outputs(a: Path, b: Path) -> None:
    for name in ("depth", "mean_depth", "speed", "hazard", "arrival"):
        x = np.load(a / "results" / "maxima" / f"{name}.npy")
        y = np.load(b / "results" / "maxima" / f"{name}.npy")
        assert np.array_equal(x, y, equal_nan=True), name
    for f in ("lines_gap.csv", "points_p1.csv"):
        assert (a / "series" / f).read_text() == (b / "series" / f).read_text(), f
    maps_a = sorted(p.name for p in (a / "results" / "maps").iterdir())
    assert maps_a == sorted(p.name for p in (b / "results" / "maps").iterdir())
    for m in maps_a:
        for q in ("depth", "mean_depth", "qx"):
            x = np.load(a / "results" / "maps" / m / f"{q}.npy")
            assert np.array_equal(x, np.load(b / "results" / "maps" / m / f"{q}.npy")), (m, q)


@pytest.mark.gpu
@pytest.mark.parametrize("losses", [False, True, "gridded"])
def test_the_device_loop_runs_a_job_bit_for_bit_the_host_loop(job_folder, losses):
    """``--loop device``: the steps on the device in chunks between outputs give the host loop's series, maps and
    maxima to the bit, the same step count, smallest dt and peaks, the same books to rounding (uniform rain, curve
    numbers on a depth grid, gridded rain frames)."""
    from wadi2d_amd.run import run_job

    if losses:
        _loss_job(job_folder, gridded=losses == "gridded")
    cfg = jobmod.load_job(job_folder / "job.json")
    host = run_job(cfg, job_folder / "host", backend="hip-exact")
    dev = run_job(cfg, job_folder / "dev", backend="hip-exact", loop="device")
    print(f"device loop (losses {losses}): exit {dev.exit_code} steps {dev.steps}/{host.steps} loop "
          f"{dev.loop_s:.2f}/{host.loop_s:.2f} s", flush=True)
    assert host.exit_code == 0 and dev.exit_code == 0, dev.message
    assert dev.steps == host.steps and dev.t_end == host.t_end and dev.dt_min == host.dt_min
    _same_outputs(job_folder / "host", job_folder / "dev")
    assert dev.peaks == host.peaks
    bh, bd = host.balance, dev.balance
    # the FP64 books: the device adds each step's own length to the tally, the host loop the difference of its
    # clock's readings, which can differ in the last bit
    assert bd.storage == bh.storage
    assert bd.outflow == pytest.approx(bh.outflow, rel=1e-12) and bd.rain == pytest.approx(bh.rain, rel=1e-12)
    assert bd.relative == pytest.approx(bh.relative, abs=1e-12)


@pytest.mark.gpu
def test_a_run_stopped_and_resumed_on_the_device_loop_is_the_unbroken_host_run(job_folder):
    from wadi2d_amd.run import run_job

    cfg = jobmod.load_job(job_folder / "job.json")
    whole = run_job(cfg, job_folder / "whole", backend="hip-exact")
    first = run_job(cfg, job_folder / "split", backend="hip-exact", stop_at_s=240.0, loop="device")
    assert first.exit_code == 6 and first.t_end == pytest.approx(240.0)
    second = run_job(cfg, job_folder / "split", backend="hip-exact", resume=True, loop="device")
    print(f"whole {whole.steps} steps, device split {first.steps} + resumed to {second.steps}", flush=True)
    assert second.exit_code == 0 and second.steps == whole.steps
    _same_outputs(job_folder / "whole", job_folder / "split")


@pytest.mark.gpu
def test_every_boundary_kind_and_storm_kind_runs_and_a_collapsing_step_ends_in_exit_5(job_folder):
    """A stage and a normal boundary beside the free one, a hyetograph storm, then the same job with a minimum step
    no step can meet: the run stops with exit 5, an error event and a done event with its partial books
    (run.py:1749-1757, stepper.py:117-124)."""
    from wadi2d_amd.run import run_job

    xe, yn = X0 + NX * DX, Y0 + NY * DX
    d = job_dict(duration=300.0)
    d["storm"] = {"kind": "hyetograph", "hyetograph_min": [2.0, 4.0], "hyetograph_mm": [10.0, 20.0]}
    d["boundaries"] = [
        {"kind": "free", "coordinates": [[xe - 1.0, yn], [xe - 1.0, Y0 + 320.0]], "name": "east top"},
        {"kind": "normal", "slope": 0.01, "coordinates": [[xe - 1.0, Y0 + 310.0], [xe - 1.0, Y0]], "name": "east"},
        {"kind": "stage", "stage": 100.5, "coordinates": [[X0 + 1.0, Y0 + 100.0], [X0 + 1.0, Y0]], "name": "west"},
    ]
    (job_folder / "job.json").write_text(json.dumps(d), encoding="utf-8")
    cfg = jobmod.load_job(job_folder / "job.json")
    s = run_job(cfg, job_folder / "kinds", backend="hip")
    print(f"kinds: exit {s.exit_code} steps {s.steps} mass {s.balance.relative:.2e}", flush=True)
    assert s.exit_code == 0, s.message
    assert abs(s.balance.relative) <= 1e-5  # the stage boundary's inflow is not booked as inflow here either
    d["numerics"] = {"dt_min": 1.0}
    (job_folder / "job.json").write_text(json.dumps(d), encoding="utf-8")
    cfg = jobmod.load_job(job_folder / "job.json")
    s = run_job(cfg, job_folder / "fail", backend="hip")
    assert s.exit_code == 5 and "below the minimum" in s.message
    events = [json.loads(x) for x in (job_folder / "fail" / "progress.jsonl").read_text().splitlines()]
    kinds = [e["ev"] for e in events]
    assert "error" in kinds and kinds[-1] == "done" and events[-1]["code"] == 5 and events[-1]["partial"]


@pytest.mark.gpu
def test_a_collapsing_step_on_the_device_loop_ends_as_the_host_loop_does(job_folder):
    """Exit 5 on ``--loop device`` (stepper.py:117-124, run.py:1749-1757): with the minimum step set just above
    the smallest step a finished run took, the run collapses at that step, somewhere inside a chunk of device
    steps. Both loops stop at the same step with the same message; the failing step is not booked (no tally, no
    smallest dt), the failed member stays listed so the final maxima sample covers its tiles, and the series,
    maxima and books equal the host loop's. Stage, normal and free boundaries under a hyetograph storm."""
    from wadi2d_amd.run import run_job

    xe, yn = X0 + NX * DX, Y0 + NY * DX
    d = job_dict(duration=300.0)
    d["storm"] = {"kind": "hyetograph", "hyetograph_min": [2.0, 4.0], "hyetograph_mm": [10.0, 20.0]}
    d["boundaries"] = [
        {"kind": "free", "coordinates": [[xe - 1.0, yn], [xe - 1.0, Y0 + 320.0]], "name": "east top"},
        {"kind": "normal", "slope": 0.01, "coordinates": [[xe - 1.0, Y0 + 310.0], [xe - 1.0, Y0]], "name": "east"},
        {"kind": "stage", "stage": 100.5, "coordinates": [[X0 + 1.0, Y0 + 100.0], [X0 + 1.0, Y0]], "name": "west"},
    ]
    (job_folder / "job.json").write_text(json.dumps(d), encoding="utf-8")
    ok = run_job(jobmod.load_job(job_folder / "job.json"), job_folder / "ok", backend="hip-exact")
    assert ok.exit_code == 0, ok.message
    d["numerics"] = {"dt_min": ok.dt_min * (1.0 + 1e-6)}
    (job_folder / "job.json").write_text(json.dumps(d), encoding="utf-8")
    cfg = jobmod.load_job(job_folder / "job.json")
    host = run_job(cfg, job_folder / "host", backend="hip-exact")
    dev = run_job(cfg, job_folder / "dev", backend="hip-exact", loop="device")
    print(f"collapse: finished run {ok.steps} steps, dt_min {ok.dt_min:.6e}; host exit {host.exit_code} at step "
          f"{host.steps}, device exit {dev.exit_code} at step {dev.steps}", flush=True)
    assert host.exit_code == 5 and "below the minimum" in host.message
    assert dev.exit_code == 5 and dev.message == host.message
    assert dev.steps == host.steps and dev.t_end == host.t_end and dev.dt_min == host.dt_min
    assert 0 < host.steps < ok.steps
    for name in ("depth", "mean_depth", "speed", "hazard", "arrival"):
        x = np.load(job_folder / "host" / "results" / "maxima" / f"{name}.npy")
        y = np.load(job_folder / "dev" / "results" / "maxima" / f"{name}.npy")
        assert np.array_equal(x, y, equal_nan=True), name
    for f in ("lines_gap.csv", "points_p1.csv"):
        assert (job_folder / "host" / "series" / f).read_text() == (job_folder / "dev" / "series" / f).read_text(), f
    bh, bd = host.balance, dev.balance
    assert bd.storage == bh.storage
    assert bd.outflow == pytest.approx(bh.outflow, rel=1e-12) and bd.rain == pytest.approx(bh.rain, rel=1e-12)


# ---- the run folder's files


def test_progress_results_and_checkpoint_files_round_trip(tmp_path):
    from wadi2d_amd.io import Progress, Results, load_checkpoint, save_checkpoint, write_series_csv

    with Progress(tmp_path / "progress.jsonl") as p:
        p.emit("start", cells=3, dx=10.0)
        p.emit("done", code=0, balance={"rain_m3": 1.5})
    lines = [json.loads(x) for x in (tmp_path / "progress.jsonl").read_text().splitlines()]
    assert lines == [{"ev": "start", "cells": 3, "dx": 10.0}, {"ev": "done", "code": 0, "balance": {"rain_m3": 1.5}}]
    g = _grid(8, 8)
    r = Results(tmp_path / "results", g, "{}")
    r.snapshot(900.4, {"depth": np.ones(g.shape)})
    r.maxima({"depth": np.full(g.shape, 2.0)})
    r.series("lines", "a b", np.array([0.0, 60.0]), {"discharge_m3s": [1.0, 2.0]})
    r.attributes({"status": "finished"})
    assert np.load(tmp_path / "results" / "maps" / "t000000900" / "depth.npy").dtype == np.float32
    assert json.loads((tmp_path / "results" / "attributes.json").read_text())["status"] == "finished"
    with np.load(tmp_path / "results" / "series" / "lines" / "a_b.npz") as f:
        assert list(f["discharge_m3s"]) == [1.0, 2.0]
    again = Results(tmp_path / "results", g, "{}", append=True)
    assert again.attrs["status"] == "finished"
    write_series_csv(tmp_path / "s.csv", [0.0, 60.0], {"q": [1.0, 2.5]})
    assert (tmp_path / "s.csv").read_text().splitlines() == ["t_s,q", "0.000,1", "60.000,2.5"]
    save_checkpoint(tmp_path / "checkpoint.npz", {"job_hash": "abc", "t": 30.0, "h": np.arange(4.0)})
    back = load_checkpoint(tmp_path / "checkpoint.npz")
    assert back["job_hash"] == "abc" and back["t"] == 30.0 and np.array_equal(back["h"], np.arange(4.0))
    (tmp_path / "bad.npz").write_bytes(b"not a checkpoint")
    assert load_checkpoint(tmp_path / "bad.npz") is None


@pytest.mark.gpu
def test_the_command_line_runs_a_job_on_the_device(job_folder):
    """``python -m wadi2d_amd run job.json --out DIR --backend hip-f64``: exit 0, the events on stdout as in the file,
    the maxima and the series written; a second run with --resume and no stop finds the finished run's checkpoint
    missing (none is written on finishing) and starts again with a warning."""
    d = job_dict(duration=120.0)
    (job_folder / "job.json").write_text(json.dumps(d), encoding="utf-8")
    cmd = [sys.executable, "-m", "wadi2d_amd", "run", str(job_folder / "job.json"), "--out", str(job_folder / "o"),
           "--backend", "hip-f64"]  # fmt: skip
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=ROOT, timeout=240)
    print(r.stdout[-600:], r.stderr[-600:], flush=True)
    assert r.returncode == 0
    out = [json.loads(x) for x in r.stdout.splitlines() if x.startswith("{")]
    file = [json.loads(x) for x in (job_folder / "o" / "progress.jsonl").read_text().splitlines()]
    assert out == file and file[-1]["ev"] == "done" and file[-1]["code"] == 0
    assert np.load(job_folder / "o" / "results" / "maxima" / "depth.npy").max() > 0.1
    r = subprocess.run([*cmd, "--resume"], capture_output=True, text=True, cwd=ROOT, timeout=240)
    assert r.returncode == 0 and "no checkpoint to resume from" in r.stdout
