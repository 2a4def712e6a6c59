"""A storm/ARF ensemble of uneven members advanced together in one device-resident run, and the same members run
one after another the plain way (the serial loop of wadi2d/arf.py:347-371, one run_job per storm).

Members are stacked into one mosaic grid: member m's padded grid sits in its own band of tile rows, the bands
separated by one empty tile row. Tiles are numbered row-major, so each member's tiles form one contiguous range in
the member's own order, no tile has a neighbour in another member (the void tile stands in), and the member's
tables are its own, built on its own grid and laid end to end. Every member keeps its own clock, step length,
source windows, active tiles and books on the device (engine/device_run.py).
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from wadi2d_amd import synth
from wadi2d_amd.engine.device_run import DeviceRun, MemberPlan
from wadi2d_amd.engine.driver import Driver
from wadi2d_amd.engine.forcing import Boundary, Rain, boundary_set
from wadi2d_amd.engine.storms import ScsStorm
from wadi2d_amd.grid import TileGrid
from wadi2d_amd.layout import STRIDE2, TILE
from wadi2d_amd.solver.model import Model
from wadi2d_amd.solver.state import from_state, state_index
from wadi2d_amd.tables import GridTables

SIDES = (128, 256, 512, 768, 1024)
TILE_FIELDS = ("hc", "dtab", "flx", "frx", "glx", "grx", "efx", "cwx", "ftabx", "fly", "fry", "gly", "gry", "efy",
               "cwy", "ftaby", "zcx", "zcy")  # fmt: skip


@dataclass
class MemberSpec:
    """One member: its ground (seed, side) and its storm (point depth, duration, area-reduction factor)."""

    index: int
    seed: int
    side: int
    depth_mm: float
    duration_h: float
    arf: float

    @property
    def storm(self) -> ScsStorm:
        return ScsStorm(self.depth_mm, self.duration_h, self.arf)


def draw_members(count: int = 64, seed: int = 2024, sides=SIDES) -> list[MemberSpec]:
    """Sides uniform over ``sides``, depth uniform on 20-120 mm, duration on 1-6 h, factor on 0.6-1.0."""
    rng = np.random.default_rng(seed)
    out = []
    for m in range(count):
        out.append(MemberSpec(m, int(rng.integers(1, 2**31)), int(rng.choice(sides)), float(rng.uniform(20.0, 120.0)),
                              float(rng.uniform(1.0, 6.0)), float(rng.uniform(0.6, 1.0))))  # fmt: skip
    return out


def member_scenario(spec: MemberSpec, dx: float = 20.0, k: int = 4, duration_s: float = 3 * 3600.0):
    """The member's ground (the 256 x 256 plain case's recipe at its own side and seed) and its storm."""
    return synth.c0_style_case(spec.side, dx, k, spec.seed, spec.storm, duration_s,
                               f"member {spec.index} {spec.side}x{spec.side}")  # fmt: skip


@dataclass
class Slot:
    """Where one member sits in the mosaic."""

    row0: int  # first padded row of the member in the mosaic
    tile0: int
    tile1: int
    grid: TileGrid


def stack_tables(members: list[GridTables]) -> tuple[GridTables, list[Slot]]:
    """The mosaic of the members' tables (see the module docstring)."""
    if not members:
        raise ValueError("an ensemble needs at least one member")
    dx = members[0].grid.dx
    if any(t.grid.dx != dx for t in members) or len({t.rows for t in members}) != 1:
        raise ValueError("ensemble members share the cell size and the number of table rows")
    width = max(t.grid.shape[1] for t in members)
    height = sum(t.grid.shape[0] + TILE for t in members) - TILE
    mask = np.zeros((height, width), bool)
    zmin = np.zeros((height, width))
    zbar = np.zeros((height, width))
    slots, row, tile = [], 0, 1
    for t in members:
        ny, nx = t.grid.shape
        mask[row : row + ny, :nx] = t.grid.mask
        zmin[row : row + ny, :nx] = t.zmin
        zbar[row : row + ny, :nx] = t.zbar
        n = t.grid.ntiles - 1
        slots.append(Slot(row, tile, tile + n, t.grid))
        row += ny + TILE
        tile += n
    grid = TileGrid(0.0, 0.0, dx, mask)
    for s in slots:  # the mosaic's numbering is each member's own, shifted
        own = s.grid.tiles[1:] + np.array([s.row0 // TILE, 0], np.int32)
        if not np.array_equal(grid.tiles[s.tile0 : s.tile1], own):
            raise AssertionError("mosaic tiles out of member order")
    arrays = {}
    for name in TILE_FIELDS:
        parts = [getattr(members[0], name)[:1]] + [getattr(t, name)[1:] for t in members]
        arrays[name] = np.concatenate(parts)
    return GridTables(grid=grid, zmin=zmin, zbar=zbar, **arrays), slots


class MemberRain:
    """A member's storm (cumulative depth in mm against time in s) over its interior cells (state indices of the
    mosaic): the window scale the device forms rates from, as ``Rain.scale`` (the weight form); with curve numbers
    (``cn`` per interior cell, NaN: none) the window's storm depths, as ``Rain.terms`` (the excess form)."""

    pixels = None
    depths = None

    def __init__(self, storm, state: np.ndarray, cn: np.ndarray | None = None) -> None:
        self.storm = storm
        self.state = state
        self.cn = cn

    @property
    def excess_form(self) -> bool:
        return self.cn is not None

    def scale(self, t0: float, t1: float) -> float:
        return (float(self.storm(t1)) - float(self.storm(t0))) / (1000.0 * (t1 - t0))

    def terms(self, t0: float, t1: float) -> tuple[float, float, float]:
        return float(self.storm(t0)), float(self.storm(t1)), 1000.0 * (t1 - t0)


@dataclass
class Ensemble:
    specs: list[MemberSpec]
    scenarios: list
    tables: GridTables
    slots: list[Slot]
    boundaries: list[list[Boundary]]
    interior: list[np.ndarray] = field(default_factory=list)  # per member, mosaic state indices
    interior_ji: list[tuple[np.ndarray, np.ndarray]] = field(default_factory=list)  # the same on the member's grid

    @classmethod
    def build(cls, specs: list[MemberSpec], dx: float = 20.0, k: int = 4, duration_s: float = 3 * 3600.0,
              workers: int | None = None) -> Ensemble:  # fmt: skip
        return cls.from_scenarios([member_scenario(s, dx, k, duration_s) for s in specs], specs, workers)

    @classmethod
    def from_scenarios(cls, scens: list, specs: list | None = None, workers: int | None = None,
                       keep_beds: bool = False) -> Ensemble:  # fmt: skip
        """Members from built-or-not scenarios (their storms are their ``cumulative_mm``)."""
        built = []
        for sc in scens:
            built.append(sc.build(workers))
            if not keep_beds:
                sc.bed = None  # the pixels are not needed again
        tables, slots = stack_tables(built)
        bounds = []
        for sc, sl in zip(scens, slots, strict=True):
            bounds.append([Boundary(b.kind, np.asarray(b.jj) + sl.row0, np.asarray(b.ii), b.stage, b.slope)
                           for b in sc.boundaries])  # fmt: skip
        ens = cls(specs, scens, tables, slots, bounds)
        bset = boundary_set([b for bs in bounds for b in bs], tables)
        interior = tables.grid.mask & ~bset.mask
        for sl in slots:
            part = np.zeros_like(interior)
            ny, nx = sl.grid.shape
            part[sl.row0 : sl.row0 + ny, :nx] = interior[sl.row0 : sl.row0 + ny, :nx]
            jj, ii = np.nonzero(part)
            ens.interior.append(np.asarray(state_index(tables.grid, jj, ii), np.int64))
            ens.interior_ji.append((jj - sl.row0, ii))
        ens.bset = bset
        return ens

    def storm(self, m: int):
        return self.scenarios[m].cumulative_mm

    @property
    def losses(self) -> bool:
        """Whether any member has curve numbers: then every member's rain is in the excess form."""
        return any(getattr(sc, "curve_number", None) is not None for sc in self.scenarios)

    def member_cn(self, m: int) -> np.ndarray:
        """Member m's curve numbers at its interior cells, in the order of ``interior[m]`` (NaN without)."""
        jj, ii = self.interior_ji[m]
        cn = getattr(self.scenarios[m], "curve_number", None)
        return np.full(jj.size, np.nan) if cn is None else np.asarray(cn, np.float64)[jj, ii]

    @property
    def cells(self) -> int:
        return int(sum(int(s.grid.mask.sum()) for s in self.slots))

    def member_field(self, m: int, mosaic: np.ndarray) -> np.ndarray:
        """Member m's block of a (ny, nx) mosaic field, on its own padded grid."""
        sl = self.slots[m]
        ny, nx = sl.grid.shape
        return mosaic[sl.row0 : sl.row0 + ny, :nx]

    def tile_member(self) -> np.ndarray:
        out = np.full(self.tables.grid.ntiles, -1, np.int64)
        for m, sl in enumerate(self.slots):
            out[sl.tile0 : sl.tile1] = m
        return out


@dataclass
class BatchedResult:
    wall_s: float
    steps: list[int]
    t_end: list[float]
    max_depth: list[np.ndarray]
    mass_error: list[float]
    launch: tuple[int, int] | None = None
    tuning: list | None = None
    peak_q: list[float] = field(default_factory=list)  # per member: the largest outflow after a step, m3/s
    t_peak: list[float] = field(default_factory=list)


@dataclass
class Batch:
    """An ensemble's mosaic model with its device run set up: the members' tally faces (in member order) and the
    member of every tile."""

    model: Model
    run: DeviceRun
    faces: np.ndarray
    member_of: np.ndarray


def setup_batched(ens: Ensemble, backend: str = "hip", maxima_every: int = 1, prepare=None,
                  pipeline: bool | None = None, **options) -> Batch:  # fmt: skip
    """The mosaic model of every member, its boundary cells, source, tally and maxima, and one device run over all
    members from t = 0 (``pipeline``: the device run's pipelined step; ``options`` go to the backend)."""
    num = ens.scenarios[0].numerics
    model = Model(ens.tables, num, backend, **options)
    model.set_depth(np.zeros(ens.tables.grid.shape))
    b = model.backend
    model.set_boundary(ens.bset.cells)
    state_all = np.concatenate(ens.interior)
    losses = ens.losses
    if losses:
        b.source_losses(state_all, np.ones(state_all.size), np.concatenate([ens.member_cn(m) for m in
                                                                            range(len(ens.slots))]))  # fmt: skip
    else:
        b.source_weights(state_all, np.ones(state_all.size))
    member_of = ens.tile_member()
    faces = np.concatenate((ens.bset.xfaces, ens.bset.yfaces)).astype(np.int64)
    axis = np.concatenate((np.zeros(ens.bset.xfaces.size), np.ones(ens.bset.yfaces.size))).astype(np.int32)
    sign = np.concatenate((ens.bset.xsign, ens.bset.ysign)).astype(np.float32)
    order = np.argsort(member_of[faces // STRIDE2], kind="stable")
    faces, axis, sign = faces[order], axis[order], sign[order]
    b.tally_set(faces, axis, sign)
    b.maxima_set(0.1, 0.05)
    st = model.stepper
    plans = []
    for m, sl in enumerate(ens.slots):
        rain = MemberRain(ens.storm(m), ens.interior[m], ens.member_cn(m) if losses else None)
        plans.append(MemberPlan(sl.tile0, sl.tile1, rain, None, 0.0, ens.scenarios[m].duration_s, st.dt, 0, True,
                                ens.interior[m]))  # fmt: skip
    if prepare is not None:
        prepare(model)
    run = DeviceRun(b, plans, num, maxima=maxima_every, idle=st.idle, pipeline=pipeline)
    return Batch(model, run, faces, member_of)


def run_batched(ens: Ensemble, backend: str = "hip", graph_steps: int = 64, chunk: int = 512, tune: bool = True,
                launch: tuple[int, int] | None = None, maxima_every: int = 1, log=None, on_chunk=None,
                prepare=None, pipeline: bool | None = None) -> BatchedResult:  # fmt: skip
    """Every member in one device-resident run to its end; per-member max depth (sampled every ``maxima_every``
    steps), mass error."""
    from wadi2d_amd.engine.autotune import apply_launch, tune_launch

    bt = setup_batched(ens, backend, maxima_every, prepare, pipeline)
    model, run, faces, member_of = bt.model, bt.run, bt.faces, bt.member_of
    b = model.backend
    plans = run.plans
    b.synchronize()
    t0 = time.perf_counter()
    tuning = None
    if launch is not None:
        apply_launch(b, launch)
    elif tune:
        launch, tuning = tune_launch(run, graph_steps=graph_steps)
    state = run.state
    while not all(s.done for s in state):
        state = run.steps(chunk, graph_steps)
        if on_chunk is not None:
            on_chunk(model, state)
        if log is not None:
            live = sum(not s.done for s in state)
            log(f"[ensemble] {backend}: {max(s.nsteps for s in state)} steps, {live} members live, "
                f"{b.nact} tiles active")  # fmt: skip
    wall = time.perf_counter() - t0
    if any(s.failed for s in state):
        raise RuntimeError(f"members failed: {[m for m, s in enumerate(state) if s.failed]}")
    dmax = from_state(ens.tables.grid, b.maxima_get(0)).astype(np.float64)
    total = b.tally_totals()
    member_face = member_of[faces // STRIDE2]
    masses = []
    for m in range(len(plans)):
        bal = run.balance(m, float(total[member_face == m].sum()))
        masses.append(bal.relative)
    return BatchedResult(wall, [int(s.nsteps) for s in state], [float(s.t) for s in state],
                         [ens.member_field(m, dmax) for m in range(len(plans))], masses, launch, tuning,
                         [float(s.q_peak) for s in state], [float(s.t_peak) for s in state])  # fmt: skip


@dataclass
class SerialResult:
    wall_s: float
    steps: list[int]
    max_depth: list[np.ndarray]
    mass_error: list[float]


def run_serial(ens: Ensemble, backend: str = "hip", members: list[int] | None = None, maxima_every: int = 1,
               log=None) -> SerialResult:  # fmt: skip
    """The plain way: each member its own model, driven from the host one step at a time, one member after another
    (a storm study's serial loop, arf.py:347-371)."""
    which = range(len(ens.scenarios)) if members is None else members
    wall, steps, dmaxes, masses = 0.0, [], [], []
    for m in which:
        sc = ens.scenarios[m]
        tables = sc.tables
        bset = boundary_set(sc.boundaries, tables)
        rain = Rain(tables, bset, ens.storm(m), curve_number=getattr(sc, "curve_number", None))
        t0 = time.perf_counter()
        model = Model(tables, sc.numerics, backend)
        model.set_depth(np.zeros(tables.grid.shape))
        drv = Driver(model, bset, rain, maxima_every=maxima_every)
        drv.run(sc.duration_s)
        model.backend.synchronize()
        wall += time.perf_counter() - t0
        steps.append(model.stepper.nsteps)
        dmaxes.append(from_state(tables.grid, model.backend.maxima_get(0)).astype(np.float64))
        masses.append(drv.accounts().relative)
        if log is not None:
            log(f"[serial] member {m} ({sc.grid.shape[0] - 2} side): {model.stepper.nsteps} steps, {wall:.1f} s so far")
        del model, drv
    return SerialResult(wall, steps, dmaxes, masses)


def max_depth_error(a: list[np.ndarray], b: list[np.ndarray], ens: Ensemble, members=None) -> list[float]:
    """Per member, the largest |difference| of max depth over the member's cells."""
    which = range(len(ens.scenarios)) if members is None else members
    return [float(np.abs(x - y)[ens.slots[m].grid.mask].max()) for m, x, y in zip(which, a, b, strict=True)]
