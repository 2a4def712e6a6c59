"""The oracle: a CPU restatement of Wadi2D's uniform-grid step and its time loop. TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may use this module, and only as
the checker (or the timed CPU baseline), never as a path of the product.

What it restates, file by file:
  kernels and step order ...... oracle/step_oracle.cpp (wadi2d/solver/cpu/kernels.py, cpu/backend.py:95-126)
  CpuBackend protocol ......... OracleBackend   (wadi2d/solver/cpu/backend.py:14-141)
  Stepper ..................... OracleStepper   (wadi2d/solver/stepper.py:14-129)
  local time stepping ......... oracle/levels.py (wadi2d/solver/level_stepper.py, solver/levels.py,
                                solver/cpu/levels.py)
  Model ....................... OracleModel     (wadi2d/solver/model.py:32-185)
  host state conversions ...... to_state/from_state/level_of/plan_of (wadi2d/solver/state.py:198-301)
  water ledger ................ Ledger          (wadi2d/engine/monitor.py:35-79,197-248; engine/maxima.py:28-192)

``prec=8`` runs the arithmetic in double: the FP64 oracle. ``prec=4`` runs it in float with the reference's
float32 statement order: the CPU baseline.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "lib" / "liboracle.so"

TILE, HALO = 32, 2
STRIDE = TILE + 2 * HALO
STRIDE2 = STRIDE * STRIDE
TILE2 = TILE * TILE
FIELDS = ("h", "d", "u", "v", "mx", "my", "qx", "qy", "u1", "v1", "mx1", "my1", "qx1", "qy1", "r", "src", "res", "om")
NEIGHBOURS = ((-1, -1), (-1, 0), (-1, 1), (0, -1), (0, 1), (1, -1), (1, 0), (1, 1))
G = 9.80665
SLEEP_CHECKS = 2  # stepper.py:14


class NumericalFailure(RuntimeError):
    """The time step collapsed or a NaN appeared (stepper.py:19-20)."""


class _Tables(ctypes.Structure):
    _fields_ = [("ntiles", ctypes.c_int32), ("rows", ctypes.c_int32), ("nbr", ctypes.c_void_p)] + [
        (n, ctypes.c_void_p)
        for n in ("hc", "dtab", "flx", "frx", "glx", "grx", "efx", "cwx", "ftabx",
                  "fly", "fry", "gly", "gry", "efy", "cwy", "ftaby", "zcx", "zcy")  # fmt: skip
    ]


class _Params(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("g", "dx", "e_wet", "h_cfl", "u_max", "lead", "a_min")] + [
        (n, ctypes.c_int32) for n in ("rows", "second", "sheet", "mirror", "carry")
    ]


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise ImportError(f"{LIB} is not built: make -C {HERE}")
        L = ctypes.CDLL(str(LIB))
        VP, I32, I64, D = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
        L.oracle_create.restype = VP
        L.oracle_create.argtypes = [ctypes.POINTER(_Tables), ctypes.POINTER(_Params), I32, I32]
        L.oracle_destroy.argtypes = [VP]
        L.oracle_set_field.argtypes = [VP, I32, VP, I64]
        L.oracle_get_field.argtypes = [VP, I32, VP, I64]
        L.oracle_set_active.argtypes = [VP, VP, I32]
        L.oracle_set_boundary.argtypes = [VP, VP, VP, VP, VP, VP, I32]
        L.oracle_step.argtypes = [VP, D, I32, I32]
        L.oracle_max_wave.restype = D
        L.oracle_max_wave.argtypes = [VP]
        L.oracle_wave.argtypes = [VP, VP]
        L.oracle_wet_flags.argtypes = [VP, VP]
        L.oracle_set_wet.argtypes = [VP, VP]
        L.oracle_refresh_plan.argtypes = [VP, VP, I32]
        L.oracle_refresh_halos.argtypes = [VP]
        L.oracle_set_threads.argtypes = [VP, I32]
        L.oracle_lts_init.argtypes = [VP]
        L.oracle_lts_set_level.argtypes = [VP, VP]
        L.oracle_lts_tiles.argtypes = [VP, I32, VP, I32, D, I32, I32]
        L.oracle_lts_edges.argtypes = [VP, I32, VP, I32, D, I32]
        _lib = L
    return _lib


@dataclass(frozen=True)
class Numerics:
    """The fields of the reference's ``Numerics`` this restatement reads (config/schema.py:457-516)."""

    cfl: float = 0.45
    limiter: str = "minmod"
    sheet_flow: bool = True
    mirror_dry_faces: bool = True
    face_depth_carry: bool = True
    e_wet: float = 1e-4
    h_cfl: float = 0.01
    u_max: float = 25.0
    dt_max: float = 5.0
    dt_min: float = 1e-3
    dt_every: int = 8
    wave_floor: float = 3.0
    plan_min: float = 0.02
    active_every: int = 16
    local_time_stepping: bool = False
    lts_levels: int = 3
    lts_wave_floor: float = 1.0
    lts_buffer: float = 1.0

    @classmethod
    def like(cls, other) -> Numerics:
        """The same numbers from any object with these attribute names (a job's Numerics)."""
        return cls(**{k: getattr(other, k) for k in cls.__dataclass_fields__})


# ---------------------------------------------------------------------------------------------------------------
# host conversions, state.py


def to_state(grid, a: np.ndarray) -> np.ndarray:
    """A (ny, nx) grid into own cells and halos, halos by clipped grid reads (state.py:198-209)."""
    out = np.zeros(grid.ntiles * STRIDE2, np.float64)
    ny, nx = grid.shape
    for t in range(1, grid.ntiles):
        tj, ti = grid.tiles[t]
        r = np.clip(np.arange(STRIDE) - HALO + tj * TILE, 0, ny - 1)
        c = np.clip(np.arange(STRIDE) - HALO + ti * TILE, 0, nx - 1)
        out[t * STRIDE2 : (t + 1) * STRIDE2] = a[np.ix_(r, c)].ravel()
    return out


def from_state(grid, a: np.ndarray) -> np.ndarray:
    """Own cells of a tile-major array as a (ny, nx) grid (state.py:241-252)."""
    ny, nx = grid.shape
    out = np.zeros((ny, nx), a.dtype)
    for t in range(1, grid.ntiles):
        tj, ti = grid.tiles[t]
        blk = a[t * STRIDE2 : (t + 1) * STRIDE2].reshape(STRIDE, STRIDE)[HALO : HALO + TILE, HALO : HALO + TILE]
        out[tj * TILE : (tj + 1) * TILE, ti * TILE : (ti + 1) * TILE] = blk
    return out


def state_index(grid, j, i):
    """State-array index of grid cells (state.py:255-260)."""
    tile = grid.tile_of[np.asarray(j) // TILE, np.asarray(i) // TILE]
    return tile * STRIDE2 + (np.asarray(j) % TILE + HALO) * STRIDE + np.asarray(i) % TILE + HALO


def level_of(h, hc, dtab) -> np.ndarray:
    """Level above zmin for mean depths, float64 (state.py:269-282)."""
    last = dtab.shape[-1] - 1
    out = np.array(h, dtype=np.float64)
    wet = hc > 0.0
    above = wet & (h >= hc)
    out[above] = dtab[above, last] + (h[above] - hc[above])
    mid = wet & ~above
    s = np.sqrt(h[mid] / hc[mid]) * last
    r = np.minimum(s.astype(int), last - 1)
    idx = np.arange(r.size)
    t = dtab[mid]
    out[mid] = t[idx, r] + (s - r) * (t[idx, r + 1] - t[idx, r])
    return out


def plan_of(h, hc, dtab) -> np.ndarray:
    """Wet share of the cell plan, float64 (state.py:291-301)."""
    last = dtab.shape[-1] - 1
    out = np.ones(h.shape, np.float64)
    mid = (hc > 0.0) & (h < hc)
    s = np.sqrt(h[mid] / hc[mid]) * last
    r = np.minimum(s.astype(int), last - 1)
    idx = np.arange(r.size)
    t = dtab[mid]
    out[mid] = hc[mid] * (2 * r + 1) / last**2 / np.maximum(t[idx, r + 1] - t[idx, r], 1e-12)
    return out


# ---------------------------------------------------------------------------------------------------------------
# backend, cpu/backend.py


class OracleBackend:
    """The CpuBackend protocol over the C++ restatement; fields cross as float64 whatever the arithmetic."""

    name = "oracle"

    def __init__(self, tables, num: Numerics, prec: int = 8, threads: int = 1) -> None:
        self.grid = tables.grid
        self.num = num
        self.prec = prec
        self.rows = tables.rows
        self.n = self.grid.ntiles * STRIDE2
        keep = {}

        def f32(a):
            arr = np.ascontiguousarray(a, dtype=np.float32)
            keep[id(arr)] = arr
            return arr.ctypes.data

        nbr = np.ascontiguousarray(self.grid.nbr, dtype=np.int32)
        t = _Tables(self.grid.ntiles, tables.rows, nbr.ctypes.data,
                    *[f32(getattr(tables, n)) for n in ("hc", "dtab", "flx", "frx", "glx", "grx", "efx", "cwx",
                      "ftabx", "fly", "fry", "gly", "gry", "efy", "cwy", "ftaby", "zcx", "zcy")])  # fmt: skip
        lead = float(max(num.dt_every, 1 << num.lts_levels) if num.local_time_stepping else num.dt_every)  # backend.py:32
        p = _Params(G, self.grid.dx, num.e_wet, num.h_cfl, num.u_max, lead, num.plan_min, tables.rows,
                    int(num.limiter != "none"), int(num.sheet_flow), int(num.mirror_dry_faces),
                    int(num.face_depth_carry))  # fmt: skip
        self.h = lib().oracle_create(ctypes.byref(t), ctypes.byref(p), prec, threads)
        del keep
        self.wet_host = np.zeros(self.grid.ntiles, np.int32)
        self.bc_n = 0
        self.set_active(np.arange(1, self.grid.ntiles, dtype=np.int32))

    def __del__(self) -> None:
        h = getattr(self, "h", None)
        if h:
            lib().oracle_destroy(h)
            self.h = None

    def set_threads(self, threads: int) -> None:
        lib().oracle_set_threads(self.h, threads)

    def set_active(self, active: np.ndarray) -> None:
        self.active_tiles = np.ascontiguousarray(active, dtype=np.int32)
        self.nact = int(self.active_tiles.size)
        lib().oracle_set_active(self.h, self.active_tiles.ctypes.data, self.nact)

    def set_field(self, name: str, a: np.ndarray) -> None:
        v = np.ascontiguousarray(a, dtype=np.float64)
        if v.size != self.n:
            raise ValueError(f"{name}: {v.size} entries for {self.n}")
        lib().oracle_set_field(self.h, FIELDS.index(name), v.ctypes.data, self.n)

    def get_field(self, name: str) -> np.ndarray:
        out = np.zeros(self.n, np.float64)
        lib().oracle_get_field(self.h, FIELDS.index(name), out.ctypes.data, self.n)
        return out

    def set_boundary(self, bc) -> None:
        arrs = [np.ascontiguousarray(bc.cell, np.int32), np.ascontiguousarray(bc.ref, np.int32),
                np.ascontiguousarray(bc.kind, np.int32), np.ascontiguousarray(bc.d0, np.float32),
                np.ascontiguousarray(bc.h0, np.float32)]  # fmt: skip
        self.bc_n = int(arrs[0].size)
        lib().oracle_set_boundary(self.h, *[a.ctypes.data for a in arrs], self.bc_n)
        self._bc = arrs

    def set_wet(self, flags: np.ndarray) -> None:
        f = np.ascontiguousarray(flags, dtype=np.int32)
        lib().oracle_set_wet(self.h, f.ctypes.data)

    def refresh_plan(self, tiles: np.ndarray) -> None:
        if tiles.size:
            t = np.ascontiguousarray(tiles, dtype=np.int32)
            lib().oracle_refresh_plan(self.h, t.ctypes.data, int(t.size))

    def refresh_halos(self) -> None:
        lib().oracle_refresh_halos(self.h)

    def step(self, dt: float, want_wave: bool, want_wet: bool) -> None:
        # the reference passes np.float32(dt) to its float32 kernels (cpu/backend.py:421); double keeps dt
        lib().oracle_step(self.h, float(np.float32(dt)) if self.prec == 4 else float(dt), int(want_wave),
                          int(want_wet))  # fmt: skip

    def max_wave(self) -> float:
        return float(lib().oracle_max_wave(self.h))

    def wave(self) -> np.ndarray:
        out = np.zeros(self.grid.ntiles, np.float64)
        lib().oracle_wave(self.h, out.ctypes.data)
        return out

    def wet_flags(self) -> np.ndarray:
        out = np.zeros(self.grid.ntiles, np.int32)
        lib().oracle_wet_flags(self.h, out.ctypes.data)
        return out

    def synchronize(self) -> None:
        return None


# ---------------------------------------------------------------------------------------------------------------
# stepper.py


class OracleStepper:
    def __init__(self, backend: OracleBackend, grid, num: Numerics) -> None:
        self.backend = backend
        self.grid = grid
        self.num = num
        self.t = 0.0
        self.nsteps = 0
        self.adaptive_active = True
        self.keep_active = np.zeros(grid.ntiles, bool)
        self.refresh_due = True
        self.dt = num.cfl * grid.dx / num.wave_floor
        self.idle = np.zeros(grid.ntiles, np.int32)

    def rebuild_active(self, retain: bool = False, check: bool = False) -> None:  # stepper.py:56-92
        ntiles = self.grid.ntiles
        b = self.backend
        if not self.adaptive_active:
            if int(b.nact) != ntiles - 1:
                b.set_active(np.arange(1, ntiles, dtype=np.int32))
            return
        wet = b.wet_flags().astype(bool) | self.keep_active
        wet[0] = False
        act = wet.copy()
        act[self.grid.nbr[wet].ravel()] = True
        current = np.zeros(ntiles, bool)
        current[b.active_tiles] = True
        if retain:
            act |= current
        elif check:
            self.idle[act] = 0
            self.idle[~act] += 1
            act |= current & (self.idle < SLEEP_CHECKS)
        else:
            self.idle[:] = 0
        act[0] = False
        woken = act & ~current
        b.set_active(np.nonzero(act)[0].astype(np.int32))
        if (retain or check) and woken.any():
            b.refresh_plan(np.nonzero(woken)[0].astype(np.int32))
            b.refresh_halos()

    def advance(self, t_end: float, on_step=None) -> float:  # stepper.py:101-129
        b = self.backend
        num = self.num
        dx = self.grid.dx
        while self.t < t_end - 1e-9:
            refresh = self.refresh_due or self.nsteps % num.dt_every == 0
            self.refresh_due = False
            activate = self.nsteps % num.active_every == 0
            dt = min(self.dt, t_end - self.t, num.dt_max)
            b.step(dt, refresh, activate)
            self.t += dt
            self.nsteps += 1
            if refresh:
                w = b.max_wave()
                if not np.isfinite(w):
                    raise NumericalFailure(f"non-finite wave speed at t = {self.t:.3f} s")
                self.dt = min(num.cfl * dx / max(w, num.wave_floor), num.dt_max)
                if self.dt < num.dt_min:
                    raise NumericalFailure(f"time step {self.dt:.2e} s below the minimum at t = {self.t:.3f} s")
            if activate and self.adaptive_active:
                self.rebuild_active(check=True)
            if on_step is not None:
                on_step(self)
        return self.t


# ---------------------------------------------------------------------------------------------------------------
# model.py


class OracleModel:
    def __init__(self, tables, num=None, prec: int = 8, threads: int = 1) -> None:
        self.tables = tables
        self.grid = tables.grid
        self.num = Numerics.like(num) if num is not None else Numerics()
        self.backend = OracleBackend(tables, self.num, prec, threads)
        self.stepper = OracleStepper(self.backend, self.grid, self.num)
        self.boundary_tiles = np.zeros(0, np.int64)
        self._hc = self.grid.from_tiles(tables.hc).astype(np.float64)
        self._dtab = self.grid.from_tiles(tables.dtab).astype(np.float64)

    def _tile_wet(self, h: np.ndarray) -> np.ndarray:
        out = np.zeros(self.grid.ntiles, np.int32)
        for t in range(1, self.grid.ntiles):
            tj, ti = self.grid.tiles[t]
            out[t] = int(np.any(h[tj * TILE : (tj + 1) * TILE, ti * TILE : (ti + 1) * TILE] > 0.0))
        return out

    def set_depth(self, hbar: np.ndarray) -> None:  # model.py:43-54
        h = np.where(self.grid.mask, hbar, 0.0)
        self.backend.set_field("h", to_state(self.grid, h))
        self.backend.set_field("d", to_state(self.grid, level_of(h, self._hc, self._dtab)))
        self.backend.set_field("om", to_state(self.grid, plan_of(h, self._hc, self._dtab)))
        self.backend.set_wet(self._tile_wet(h))
        self.stepper.rebuild_active()

    def set_source(self, rate: np.ndarray) -> None:  # model.py:99-110
        self.backend.set_field("src", to_state(self.grid, np.where(self.grid.mask, rate, 0.0)))
        keep = self.stepper.keep_active
        keep[:] = False
        for t in range(1, self.grid.ntiles):
            tj, ti = self.grid.tiles[t]
            keep[t] = bool(np.any(rate[tj * TILE : (tj + 1) * TILE, ti * TILE : (ti + 1) * TILE] != 0.0))
        keep[self.boundary_tiles] = True
        self.stepper.rebuild_active(retain=True)
        self.stepper.refresh_due = True

    def set_source_cells(self, cells: np.ndarray, rate: np.ndarray) -> None:  # model.py:112-122
        src = np.zeros(self.grid.ntiles * STRIDE2, np.float32)
        src[cells] = rate
        self.backend.set_field("src", src)
        keep = self.stepper.keep_active
        keep[:] = False
        keep[cells[rate != 0.0] // STRIDE2] = True
        keep[self.boundary_tiles] = True
        self.stepper.rebuild_active(retain=True)
        self.stepper.refresh_due = True

    def set_boundary(self, bc) -> None:  # model.py:145-150
        self.backend.set_boundary(bc)
        self.boundary_tiles = np.unique(np.asarray(bc.cell) // STRIDE2)
        self.stepper.keep_active[self.boundary_tiles] = True
        self.stepper.rebuild_active(retain=True)

    def advance(self, t_end: float, on_step=None) -> float:
        return self.stepper.advance(t_end, on_step)

    def field(self, name: str) -> np.ndarray:
        return from_state(self.grid, self.backend.get_field(name))

    def volume(self) -> float:
        return float(self.field("h").sum(dtype=np.float64) * self.grid.dx**2)

    def cell_state_index(self, j: int, i: int) -> int:
        t = int(self.grid.tile_of[j // TILE, i // TILE])
        return t * STRIDE2 + (j % TILE + HALO) * STRIDE + i % TILE + HALO


# ---------------------------------------------------------------------------------------------------------------
# the water ledger and the maxima, engine/monitor.py and engine/maxima.py


@dataclass
class Balance:
    """Water accounts of the interior cells, m3 (monitor.py:35-79, the fields this path fills)."""

    rain: float = 0.0
    inflow: float = 0.0
    outflow: float = 0.0
    storage0: float = 0.0
    storage: float = 0.0
    transmission: float = 0.0

    @property
    def error(self) -> float:
        return self.rain + self.inflow - self.transmission - self.outflow - (self.storage - self.storage0)

    @property
    def relative(self) -> float:
        scale = max(self.rain + self.inflow + max(-self.outflow, 0.0) + self.storage0, 1e-9)
        return self.error / scale


class Maxima:
    """The running maxima of the active tiles' own cells, restating engine/maxima.py:29-51 (``_update_cpu``) statement
    for statement: level above the bed (depth), mean depth, speed at the cell centre from its west/east and
    north/south face velocities where the level exceeds ``deep``, speed times level (hazard), and the first time the
    level reaches ``arrive`` (-1 before). In the backend's precision: float32 arrays and arithmetic for ``prec=4``."""

    NAMES = ("depth", "mean_depth", "speed", "hazard", "arrival")

    def __init__(self, model: OracleModel, arrive: float = 0.1, deep: float = 0.05) -> None:
        self.model = model
        self.dtype = np.float32 if model.backend.prec == 4 else np.float64
        n = model.grid.ntiles * STRIDE2
        self.arrays = [np.zeros(n, self.dtype) for _ in self.NAMES]
        self.arrays[4][:] = -1.0
        self.arrive, self.deep = self.dtype(arrive), self.dtype(deep)

    def update(self, t: float) -> None:
        b = self.model.backend
        d, h, u, v = (b.get_field(f).astype(self.dtype) for f in ("d", "h", "u", "v"))
        dmax, hmax, smax, zmax, tarr = self.arrays
        active = np.asarray(b.active_tiles, np.int64)
        jj, ii = np.mgrid[0:TILE, 0:TILE]
        own = ((jj + HALO) * STRIDE + ii + HALO).ravel()
        k = (active[:, None] * STRIDE2 + own[None, :]).ravel()
        dk = d[k]
        dmax[k] = np.where(dk > dmax[k], dk, dmax[k])
        hmax[k] = np.where(h[k] > hmax[k], h[k], hmax[k])
        half = self.dtype(0.5)
        uc = half * (u[k] + u[k + 1])  # the east face is the next state entry, a halo entry at the tile's edge
        vc = half * (v[k] + v[k + STRIDE])
        sp = np.sqrt(uc * uc + vc * vc)
        wet = dk > self.deep
        smax[k] = np.where(wet & (sp > smax[k]), sp, smax[k])
        hz = sp * dk
        zmax[k] = np.where(wet & (hz > zmax[k]), hz, zmax[k])
        tarr[k] = np.where((tarr[k] < 0) & (dk >= self.arrive), self.dtype(t), tarr[k])


class Ledger:
    """What a run accumulates beside the step: the FP64 boundary tally of every step (maxima.py:149-192), the storage
    of the interior cells with their residuals (monitor.py:214-218), the rain credit of each window
    (monitor.py:197-212), the per-cell maximum depth sampled every step, the peak outfall discharge, and with
    ``maxima_every`` the five running maxima sampled after every step whose count is a multiple of it
    (maxima.py:103-146 as the run loop calls it)."""

    def __init__(self, model: OracleModel, interior: np.ndarray, xfaces, xsign, yfaces, ysign,
                 maxima_every: int = 0, arrive: float = 0.1, deep: float = 0.05) -> None:  # fmt: skip
        self.model = model
        grid = model.grid
        self.storage_index = np.asarray(state_index(grid, *np.nonzero(interior)), np.int64)
        self.xfaces, self.xsign = np.asarray(xfaces, np.int64), np.asarray(xsign, np.float64)
        self.yfaces, self.ysign = np.asarray(yfaces, np.int64), np.asarray(ysign, np.float64)
        self.total = np.zeros(self.xfaces.size + self.yfaces.size)
        self.balance = Balance()
        self.window = (0.0, 1.0, 0.0, 0.0)
        self.settled = 0.0
        self.last_t = model.stepper.t
        self.dmax = np.zeros(grid.ntiles * STRIDE2)
        self.peak_q = 0.0
        self.maxima_every = int(maxima_every)
        self.maxima = Maxima(model, arrive, deep) if self.maxima_every else None
        self.balance.storage0 = self.storage()

    def storage(self) -> float:
        b = self.model.backend
        held = b.get_field("h")[self.storage_index] + b.get_field("res")[self.storage_index]
        return float(held.sum()) * self.model.grid.dx**2

    def open_window(self, t0: float, t1: float, rain: float, inflow: float) -> None:
        self.window = (t0, t1, rain, inflow)
        self.settled = t0

    def settle(self, t: float) -> None:
        t0, t1, rain, inflow = self.window
        share = (min(t, t1) - self.settled) / (t1 - t0)
        self.balance.rain += rain * share
        self.balance.inflow += inflow * share
        self.settled = min(t, t1)

    def on_step(self, st: OracleStepper) -> None:
        b = self.model.backend
        dt = st.t - self.last_t
        self.last_t = st.t
        qx, qy = b.get_field("qx"), b.get_field("qy")
        q = np.concatenate((self.xsign * qx[self.xfaces], self.ysign * qy[self.yfaces]))
        self.total += dt * q
        self.peak_q = max(self.peak_q, float(q.sum()))
        np.maximum(self.dmax, b.get_field("d"), out=self.dmax)
        if self.maxima is not None and st.nsteps % self.maxima_every == 0:
            self.maxima.update(st.t)

    def accounts(self, t: float) -> Balance:
        self.settle(t)
        self.balance.outflow = float(self.total.sum())
        self.balance.storage = self.storage()
        return self.balance
