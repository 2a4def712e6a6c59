"""HipBackend: the reference's backend protocol over the HIP step library (include/wadi2d_hip.h).

It answers what ``CpuBackend`` and ``GpuBackend`` answer (wadi2d/solver/cpu/backend.py:14-141,
wadi2d/solver/gpu/backend.py:53-203): ``set_active``, ``set_field``/``get_field`` in the state layout,
``set_boundary``, ``step``, ``max_wave``, ``wet_flags``, ``refresh_plan``, ``refresh_halos``, ``synchronize``; the
device holds every array. Errors from the library raise: an invalid argument ``ValueError``, out of memory
``MemoryError``, a device error ``RuntimeError``.
"""

from __future__ import annotations

import ctypes

import numpy as np

from wadi2d_amd import _native
from wadi2d_amd.layout import STATE_ARRAYS, STRIDE2, WORK_ARRAYS

FIELDS = STATE_ARRAYS + WORK_ARRAYS
F32, F64 = 4, 8
EXACT, FAST = 0, 1
VP, I32, I64, D = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double


class _Tables(ctypes.Structure):
    _fields_ = [("ntiles", I32), ("rows", I32), ("nbr", VP)] + [
        (n, VP)
        for n in ("hc", "dtab", "flx", "frx", "glx", "grx", "efx", "cwx", "ftabx",
                  "fly", "fry", "gly", "gry", "efy", "cwy", "ftaby", "zcx", "zcy")  # fmt: skip
    ]


class _Params(ctypes.Structure):
    _fields_ = [(n, D) for n in ("g", "dx", "e_wet", "h_cfl", "u_max", "lead", "a_min")] + [
        (n, I32) for n in ("second", "sheet", "mirror", "carry")
    ]


class Member(ctypes.Structure):
    _fields_ = [("tile0", I32), ("tile1", I32), ("t0", D), ("t_end", D), ("dt0", D), ("window", D),
                ("nwindows", I32), ("scale_offset", I64), ("nsteps0", I64), ("refresh_due", I32),
                ("dry", I32)]  # fmt: skip


class RunNumerics(ctypes.Structure):
    _fields_ = [("cfl", D), ("dt_max", D), ("dt_min", D), ("wave_floor", D), ("dt_every", I32),
                ("active_every", I32), ("maxima", I32)]  # fmt: skip


class MemberState(ctypes.Structure):
    _fields_ = [("t", D), ("dt", D), ("rain", D), ("nsteps", I64), ("done", I32), ("failed", I32),
                ("refresh_due", I32), ("dry", I32), ("dt_min", D), ("q_peak", D), ("t_peak", D),
                ("pipelined", I32), ("pipelined_updates", I64)]  # fmt: skip


_SIGS = {
    "w2d_create": [ctypes.POINTER(_Tables), ctypes.POINTER(_Params), I32, I32, ctypes.POINTER(VP)],
    "w2d_destroy": [VP],
    "w2d_set_field": [VP, I32, VP, I64],
    "w2d_get_field": [VP, I32, VP, I64],
    "w2d_read_cells": [VP, I32, VP, I64, VP],
    "w2d_read_cells_noalloc": [VP, I32, VP, I64, VP],
    "w2d_write_cells": [VP, I32, VP, I64, VP],
    "w2d_set_active": [VP, VP, I32],
    "w2d_set_boundary": [VP, VP, VP, VP, VP, VP, I32],
    "w2d_step": [VP, D, I32, I32],
    "w2d_max_wave": [VP, ctypes.POINTER(D)],
    "w2d_wave": [VP, VP],
    "w2d_wet_flags": [VP, VP],
    "w2d_set_wet": [VP, VP],
    "w2d_refresh_plan": [VP, VP, I32],
    "w2d_refresh_halos": [VP],
    "w2d_tally_set": [VP, VP, VP, VP, I32],
    "w2d_tally_add": [VP, D],
    "w2d_tally_volume": [VP, ctypes.POINTER(D)],
    "w2d_tally_flow": [VP, ctypes.POINTER(D)],
    "w2d_maxima_set": [VP, D, D],
    "w2d_maxima_update": [VP, D],
    "w2d_maxima_get": [VP, I32, VP, I64],
    "w2d_maxima_put": [VP, I32, VP, I64],
    "w2d_tally_restore": [VP, VP, I32],
    "w2d_storage_set": [VP, VP, I64],
    "w2d_storage": [VP, ctypes.POINTER(D)],
    "w2d_synchronize": [VP],
    "w2d_set_launch": [VP, I32, I32],
    "w2d_closure_set": [VP, I32, VP, VP, VP, VP, VP, VP, I32],
    "w2d_mlp_sizes": [ctypes.POINTER(I32), ctypes.POINTER(I32)],
    "w2d_dev_alloc": [I64, ctypes.POINTER(VP)],
    "w2d_dev_free": [VP],
    "w2d_dev_copy": [VP, VP, I64, I32],
    "w2d_dev_synchronize": [],
    "w2d_mlp_take": [VP, VP, VP, VP, I64, VP, VP, VP, VP],
    "w2d_mlp_pack": [VP, VP, VP],
    "w2d_mlp_forward": [VP, VP, VP, I64, VP, VP],
    "w2d_mlp_train_blocks": [I64, ctypes.POINTER(I32)],
    "w2d_mlp_gradient": [VP, VP, VP, VP, VP, I64, ctypes.c_float, VP, I32, VP, VP],
    "w2d_mlp_adam": [VP, VP, VP, VP, I32, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, VP, VP],
    "w2d_source_weights": [VP, VP, VP, I64],
    "w2d_source_scale": [VP, D],
    "w2d_source_sum": [VP, ctypes.POINTER(D)],
    "w2d_source_losses": [VP, VP, VP, VP, I64],
    "w2d_source_excess": [VP, D, D, D, ctypes.POINTER(D), VP],
    "w2d_run_terms": [VP, VP, I64],
    "w2d_source_pixels": [VP, VP, VP, I64, I32],
    "w2d_source_excess_pixels": [VP, VP, VP, I32, D, ctypes.POINTER(D), VP],
    "w2d_run_pixel_terms": [VP, VP, I64, I32],
    "w2d_device_count": [ctypes.POINTER(I32)],
    "w2d_build_tables_device": [VP, I64, I64, I64, I64, D, I32, I32, D, VP, I32, VP, VP, I32,
                                VP, VP, VP, VP, VP, VP, VP, VP],  # fmt: skip
    "w2d_timing": [VP, I32],
    "w2d_run_setup": [VP, ctypes.POINTER(Member), I32, VP, I64, ctypes.POINTER(RunNumerics), VP],
    "w2d_run_steps": [VP, I32, I32],
    "w2d_run_read": [VP, ctypes.POINTER(MemberState), I32],
    "w2d_run_idle": [VP, VP],
    "w2d_run_pause": [VP, VP, I32],
    "w2d_run_pipeline": [VP, I32],
    "w2d_active_tiles": [VP, VP, ctypes.POINTER(I32)],
    "w2d_tally_totals": [VP, VP],
    "w2d_copy_probe": [I64, I32, ctypes.POINTER(D)],
    "w2d_timing_read": [VP, ctypes.POINTER(D), ctypes.POINTER(D), ctypes.POINTER(I64)],
    "w2d_level_setup": [VP],
    "w2d_level_set": [VP, VP],
    "w2d_level_list": [VP, VP, I32, I32, ctypes.POINTER(I32)],
    "w2d_level_lists_clear": [VP],
    "w2d_level_halo_state": [VP, I32],
    "w2d_level_faces": [VP, I32, D, I32],
    "w2d_level_strips": [VP, I32, D, I32],
    "w2d_level_limit": [VP, I32, D],
    "w2d_level_save": [VP, I32],
    "w2d_level_estimate": [VP, I32],
    "w2d_level_flux": [VP, I32, D],
    "w2d_level_update": [VP, I32, D, I32],
    "w2d_level_commit": [VP, I32],
    "w2d_level_commit_strips": [VP, I32],
    "w2d_level_clear": [VP, I32],
    "w2d_level_boundary": [VP],
}


def api() -> ctypes.CDLL:
    lib = _native.hip()
    if not getattr(lib, "_w2d_typed", False):
        for name, args in _SIGS.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        lib.w2d_last_error.restype = ctypes.c_char_p
        lib.w2d_last_error.argtypes = []
        lib._w2d_typed = True
    return lib


def check(status: int) -> None:
    if status == 0:
        return
    msg = api().w2d_last_error().decode()
    if status == 1:
        raise ValueError(msg)
    if status == 2:
        raise MemoryError(msg)
    raise RuntimeError(msg)


def device_count() -> int:
    n = I32(0)
    check(api().w2d_device_count(ctypes.byref(n)))
    return int(n.value)


def copy_probe(nbytes: int, repeats: int = 10) -> float:
    """Milliseconds of one plain device copy of ``nbytes`` (mean of ``repeats``)."""
    ms = D(0.0)
    check(api().w2d_copy_probe(int(nbytes), int(repeats), ctypes.byref(ms)))
    return float(ms.value)


def _c(a: np.ndarray, dtype) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=dtype)


class HipBackend:
    """Owns the device copies; ``precision`` ``"f32"`` or ``"f64"``, ``variant`` ``"fast"`` or ``"exact"``."""

    name = "hip"

    def __init__(self, tables, num, precision: str = "f32", variant: str = "fast", face_cache: bool = True) -> None:
        self.grid = tables.grid
        self.num = num
        self.rows = tables.rows
        self.precision = precision
        self.dtype = np.float64 if precision == "f64" else np.float32
        self.n = self.grid.ntiles * STRIDE2
        nbr = _c(self.grid.nbr, np.int32)
        keep = [nbr]

        def f32(a):
            arr = _c(a, np.float32)
            keep.append(arr)
            return arr.ctypes.data

        t = _Tables(self.grid.ntiles, tables.rows, nbr.ctypes.data,
                    *[f32(getattr(tables, n)) for n in ("hc", "dtab", "flx", "frx", "glx", "grx", "efx", "cwx",
                      "ftabx", "fly", "fry", "gly", "gry", "efy", "cwy", "ftaby", "zcx", "zcy")])  # fmt: skip
        p = _Params(num.g, num.dx, num.e_wet, num.h_cfl, num.u_max, num.lead, num.a_min, int(num.second),
                    int(num.sheet), int(num.mirror), int(num.carry))  # fmt: skip
        handle = VP()
        check(api().w2d_create(ctypes.byref(t), ctypes.byref(p), F64 if precision == "f64" else F32,
                               (FAST if variant == "fast" else EXACT) | (0 if face_cache else 4),
                               ctypes.byref(handle)))  # fmt: skip
        self.h = handle
        self.extra_wave = 0.0
        self.hooks = None
        self.active_tiles = np.arange(1, self.grid.ntiles, dtype=np.int32)
        self.nact = int(self.active_tiles.size)
        self._tally_n = 0

    def __del__(self) -> None:
        h = getattr(self, "h", None)
        if h is not None and h.value:
            api().w2d_destroy(h)
            self.h = VP()

    # ---- the protocol
    def set_active(self, active: np.ndarray) -> None:
        a = _c(active, np.int32)
        check(api().w2d_set_active(self.h, a.ctypes.data, int(a.size)))
        self.active_tiles = a
        self.nact = int(a.size)

    @property
    def active(self) -> np.ndarray:
        return self.active_tiles

    def set_field(self, name: str, a: np.ndarray) -> None:
        v = _c(a, self.dtype)
        if v.size != self.n:
            raise ValueError(f"{name}: {v.size} entries for {self.n}")
        check(api().w2d_set_field(self.h, FIELDS.index(name), v.ctypes.data, self.n))

    def get_field(self, name: str) -> np.ndarray:
        out = np.zeros(self.n, self.dtype)
        check(api().w2d_get_field(self.h, FIELDS.index(name), out.ctypes.data, self.n))
        return out

    def write_cells(self, name: str, index: np.ndarray, values) -> None:
        """Values into own cells at state indices (the writes callers make through the reference's arrays)."""
        idx = _c(index, np.int64)
        val = _c(np.broadcast_to(np.asarray(values, self.dtype), idx.shape), self.dtype)
        check(api().w2d_write_cells(self.h, FIELDS.index(name), idx.ctypes.data, int(idx.size), val.ctypes.data))

    def read_cells(self, name: str, index: np.ndarray) -> np.ndarray:
        idx = _c(index, np.int64)
        out = np.zeros(idx.size, self.dtype)
        check(api().w2d_read_cells(self.h, FIELDS.index(name), idx.ctypes.data, int(idx.size), out.ctypes.data))
        return out

    def read_cells_noalloc(self, name: str, index: np.ndarray) -> np.ndarray:
        """``read_cells`` through the library's fixed device scratch array: no device allocation, at most 262,144
        cells per call."""
        idx = _c(index, np.int64)
        out = np.zeros(idx.size, self.dtype)
        check(api().w2d_read_cells_noalloc(self.h, FIELDS.index(name), idx.ctypes.data, int(idx.size),
                                           out.ctypes.data))  # fmt: skip
        return out

    def set_boundary(self, bc) -> None:
        arrs = [_c(bc.cell, np.int32), _c(bc.ref, np.int32), _c(bc.kind, np.int32), _c(bc.d0, np.float32),
                _c(bc.h0, np.float32)]  # fmt: skip
        self.bc = bc
        check(api().w2d_set_boundary(self.h, *[a.ctypes.data for a in arrs], int(arrs[0].size)))

    def set_wet(self, flags: np.ndarray) -> None:
        f = _c(flags, np.int32)
        check(api().w2d_set_wet(self.h, f.ctypes.data))

    def refresh_plan(self, tiles: np.ndarray) -> None:
        t = _c(tiles, np.int32)
        check(api().w2d_refresh_plan(self.h, t.ctypes.data, int(t.size)))

    def refresh_halos(self) -> None:
        check(api().w2d_refresh_halos(self.h))

    def step(self, dt: float, want_wave: bool, want_wet: bool) -> None:
        check(api().w2d_step(self.h, float(dt), int(want_wave), int(want_wet)))

    def max_wave(self) -> float:
        out = D(0.0)
        check(api().w2d_max_wave(self.h, ctypes.byref(out)))
        return max(float(out.value), float(self.extra_wave))

    def wave(self) -> np.ndarray:
        out = np.zeros(self.grid.ntiles, np.float64)
        check(api().w2d_wave(self.h, out.ctypes.data))
        return out

    def wet_flags(self) -> np.ndarray:
        out = np.zeros(self.grid.ntiles, np.int32)
        check(api().w2d_wet_flags(self.h, out.ctypes.data))
        return out

    def synchronize(self) -> None:
        check(api().w2d_synchronize(self.h))

    def closure_set(self, cell_r, cell_l, axis, stat, norm, params, conveyance: bool = True) -> None:
        """The learned closure on the listed faces (own-layout cell indices); an empty list turns it off."""
        arrs = [_c(cell_r, np.int32), _c(cell_l, np.int32), _c(axis, np.int32), _c(stat, np.float32),
                _c(norm, np.float32), _c(params, np.float32)]  # fmt: skip
        n = int(arrs[0].size)
        check(api().w2d_closure_set(self.h, n, *[a.ctypes.data for a in arrs], int(conveyance)))

    def level_ops(self) -> HipLevels:
        """The level stepper's operations on this backend (gpu/backend.py:190-191)."""
        return HipLevels(self)

    def set_launch(self, faces_threads: int = 512, cells_threads: int = 256) -> None:
        """Workgroup sizes of the face and cell kernels (256 or 512 each); the fields do not depend on them."""
        check(api().w2d_set_launch(self.h, int(faces_threads), int(cells_threads)))

    # ---- the run's accumulators on the device (engine/maxima.py, engine/monitor.py)
    def tally_set(self, faces: np.ndarray, axis: np.ndarray, sign: np.ndarray) -> None:
        f, a, s = _c(faces, np.int32), _c(axis, np.int32), _c(sign, np.float32)
        check(api().w2d_tally_set(self.h, f.ctypes.data, a.ctypes.data, s.ctypes.data, int(f.size)))
        self._tally_n = int(f.size)

    def tally_add(self, dt: float) -> None:
        check(api().w2d_tally_add(self.h, float(dt)))

    def tally_volume(self) -> float:
        out = D(0.0)
        check(api().w2d_tally_volume(self.h, ctypes.byref(out)))
        return float(out.value)

    def tally_totals(self) -> np.ndarray:
        out = np.zeros(self._tally_n, np.float64)
        if self._tally_n:
            check(api().w2d_tally_totals(self.h, out.ctypes.data))
        return out

    def tally_restore(self, totals: np.ndarray) -> None:
        """Saved per-face totals back, as ``Tally.restore`` (engine/maxima.py:189-192) for a checkpoint's resume."""
        t = _c(totals, np.float64)
        check(api().w2d_tally_restore(self.h, t.ctypes.data if t.size else None, int(t.size)))

    def tally_flow(self) -> float:
        out = D(0.0)
        check(api().w2d_tally_flow(self.h, ctypes.byref(out)))
        return float(out.value)

    def maxima_set(self, arrival_depth: float, speed_depth: float) -> None:
        check(api().w2d_maxima_set(self.h, float(arrival_depth), float(speed_depth)))

    def maxima_update(self, t: float) -> None:
        check(api().w2d_maxima_update(self.h, float(t)))

    def maxima_get(self, which: int) -> np.ndarray:
        out = np.zeros(self.n, self.dtype)
        check(api().w2d_maxima_get(self.h, int(which), out.ctypes.data, self.n))
        return out

    def maxima_put(self, which: int, a: np.ndarray) -> None:
        """Saved maxima back (state layout), as ``Maxima.restore`` (engine/maxima.py:136-138)."""
        src = _c(a, self.dtype)
        if src.size != self.n:
            raise ValueError(f"maxima_put: {src.size} values for {self.n} state entries")
        check(api().w2d_maxima_put(self.h, int(which), src.ctypes.data, self.n))

    def source_weights(self, index: np.ndarray, weight: np.ndarray) -> None:
        idx, w = _c(index, np.int64), _c(weight, np.float64)
        check(api().w2d_source_weights(self.h, idx.ctypes.data, w.ctypes.data, int(idx.size)))

    def source_scale(self, scale: float) -> None:
        check(api().w2d_source_scale(self.h, float(scale)))

    def source_sum(self) -> float:
        out = D(0.0)
        check(api().w2d_source_sum(self.h, ctypes.byref(out)))
        return float(out.value)

    def source_losses(self, index: np.ndarray, depth_factor: np.ndarray, curve_number: np.ndarray) -> None:
        """The excess form of the source: a depth factor and a curve number (NaN: none) per own cell."""
        idx, w, cn = _c(index, np.int64), _c(depth_factor, np.float64), _c(curve_number, np.float64)
        if not (idx.size == w.size == cn.size):
            raise ValueError("source_losses: index, depth factor and curve number differ in length")
        check(api().w2d_source_losses(self.h, idx.ctypes.data, w.ctypes.data, cn.ctypes.data, int(idx.size)))

    def source_excess(self, d0: float, d1: float, denom: float) -> tuple[float, np.ndarray]:
        """One window's excess rates on the device: their FP64 sum and the per-tile non-zero flags."""
        total = D(0.0)
        tiles = np.zeros(self.grid.ntiles, np.int32)
        check(api().w2d_source_excess(self.h, float(d0), float(d1), float(denom), ctypes.byref(total),
                                      tiles.ctypes.data))  # fmt: skip
        return float(total.value), tiles.astype(bool)

    def source_pixels(self, index: np.ndarray, pixel: np.ndarray, npix: int) -> None:
        """Gridded rain: each listed own cell's pixel, a column of the per-window pixel depths."""
        idx, px = _c(index, np.int64), _c(pixel, np.int32)
        if idx.size != px.size:
            raise ValueError("source_pixels: index and pixel differ in length")
        check(api().w2d_source_pixels(self.h, idx.ctypes.data, px.ctypes.data, int(idx.size), int(npix)))

    def source_excess_pixels(self, d0: np.ndarray, d1: np.ndarray, denom: float) -> tuple[float, np.ndarray]:
        """One window of gridded rain: the pixels' depths at its ends; the rates' FP64 sum and the kept tiles."""
        a, b2 = _c(d0, np.float64), _c(d1, np.float64)
        total = D(0.0)
        tiles = np.zeros(self.grid.ntiles, np.int32)
        check(api().w2d_source_excess_pixels(self.h, a.ctypes.data, b2.ctypes.data, int(a.size), float(denom),
                                             ctypes.byref(total), tiles.ctypes.data))  # fmt: skip
        return float(total.value), tiles.astype(bool)

    def run_pixel_terms(self, depths: np.ndarray) -> None:
        """Gridded rain in a device run: (nwindows, 2, npix) pixel depths at each window's start and end."""
        d = _c(np.asarray(depths), np.float64)
        check(api().w2d_run_pixel_terms(self.h, d.ctypes.data if d.size else None, int(d.shape[0]), int(d.shape[2])))

    def run_terms(self, terms: np.ndarray) -> None:
        """(d0, d1, 1000 (t1 - t0)) of every window of a device run, as (nwindows, 3)."""
        t = _c(np.asarray(terms).reshape(-1, 3), np.float64)
        check(api().w2d_run_terms(self.h, t.ctypes.data if t.size else None, int(t.shape[0])))

    def timing(self, on: bool) -> None:
        check(api().w2d_timing(self.h, int(on)))

    def timing_read(self) -> tuple[float, float, int]:
        """Summed face-kernel and cell-kernel milliseconds since the last read, and the steps they cover."""
        f, c, n = D(0.0), D(0.0), I64(0)
        check(api().w2d_timing_read(self.h, ctypes.byref(f), ctypes.byref(c), ctypes.byref(n)))
        return float(f.value), float(c.value), int(n.value)

    # ---- device-resident run
    def run_setup(self, members: list[Member], scales: np.ndarray, num: RunNumerics, idle: np.ndarray | None) -> None:
        arr = (Member * len(members))(*members)
        sc = _c(scales, np.float64)
        idl = None if idle is None else _c(idle, np.int32)
        check(api().w2d_run_setup(self.h, arr, len(members), sc.ctypes.data if sc.size else None, int(sc.size),
                                  ctypes.byref(num), None if idl is None else idl.ctypes.data))  # fmt: skip
        self._run_m = len(members)

    def run_steps(self, nsteps: int, graph_steps: int = 0) -> None:
        check(api().w2d_run_steps(self.h, int(nsteps), int(graph_steps)))

    def run_read(self) -> list[MemberState]:
        out = (MemberState * self._run_m)()
        check(api().w2d_run_read(self.h, out, self._run_m))
        n = I32(0)
        check(api().w2d_active_tiles(self.h, None, ctypes.byref(n)))
        tiles = np.zeros(max(n.value, 1), np.int32)
        check(api().w2d_active_tiles(self.h, tiles.ctypes.data, ctypes.byref(n)))
        self.active_tiles = tiles[: n.value].copy()
        self.nact = int(n.value)
        return list(out)

    def run_pause(self, t_stop) -> None:
        """Hold each member of the run once its clock reaches t_stop[m] (inf: run on)."""
        ts = _c(np.asarray(t_stop, np.float64).ravel(), np.float64)
        check(api().w2d_run_pause(self.h, ts.ctypes.data, int(ts.size)))

    def run_pipeline(self, on: bool) -> None:
        """Whether the next run set up takes the pipelined step (the default) or the two-kernel one."""
        check(api().w2d_run_pipeline(self.h, int(bool(on))))

    def run_idle(self) -> np.ndarray:
        out = np.zeros(self.grid.ntiles, np.int32)
        check(api().w2d_run_idle(self.h, out.ctypes.data))
        return out

    def storage_set(self, index: np.ndarray) -> None:
        idx = _c(index, np.int64)
        check(api().w2d_storage_set(self.h, idx.ctypes.data, int(idx.size)))

    def storage_sum(self) -> float:
        out = D(0.0)
        check(api().w2d_storage(self.h, ctypes.byref(out)))
        return float(out.value)


class HipLevels:
    """The level stepper's operations on the device (the methods of GpuLevels, wadi2d/solver/gpu/levels.py:11-203)
    over the ``w2d_level_*`` calls. ``lists`` turns a host array into a device list handle; the tiles carry no halos,
    so ``faces_done`` and ``limit_done`` do nothing, and ``halo_state`` only sets the smallest step the staged
    neighbours are read at."""

    def __init__(self, backend: HipBackend) -> None:
        self.b = backend
        self.h = backend.h
        check(api().w2d_level_setup(self.h))

    def lists(self, a: np.ndarray) -> int:
        arr = np.ascontiguousarray(a, dtype=np.int32)
        width = 4 if arr.ndim == 2 else 1
        out = I32(0)
        n = arr.shape[0] if arr.ndim == 2 else arr.size
        check(api().w2d_level_list(self.h, arr.ctypes.data, int(n), width, ctypes.byref(out)))
        return int(out.value)

    def clear_lists(self) -> None:
        check(api().w2d_level_lists_clear(self.h))

    def set_level(self, level: np.ndarray) -> None:
        lv = np.ascontiguousarray(level, dtype=np.int32)
        check(api().w2d_level_set(self.h, lv.ctypes.data))

    def halo_state(self, tiles: int, n: int, want_plan: bool) -> None:
        check(api().w2d_level_halo_state(self.h, int(n)))

    def clear_wave(self) -> None:
        check(api().w2d_level_clear(self.h, 0))

    def clear_wet(self) -> None:
        check(api().w2d_level_clear(self.h, 1))

    def faces(self, tiles: int, dt: float, want_wave: bool) -> None:
        check(api().w2d_level_faces(self.h, tiles, float(dt), int(want_wave)))

    def strips(self, rows: int, dt: float, want_wave: bool) -> None:
        check(api().w2d_level_strips(self.h, rows, float(dt), int(want_wave)))

    def faces_done(self, tiles: int) -> None:
        return None

    def limit(self, tiles: int, dt: float) -> None:
        check(api().w2d_level_limit(self.h, tiles, float(dt)))

    def limit_done(self, tiles: int) -> None:
        return None

    def save(self, tiles: int) -> None:
        check(api().w2d_level_save(self.h, tiles))

    def estimate(self, rows: int) -> None:
        check(api().w2d_level_estimate(self.h, rows))

    def flux(self, rows: int, dt: float) -> None:
        check(api().w2d_level_flux(self.h, rows, float(dt)))

    def update(self, tiles: int, dt: float, want_plan: bool) -> None:
        check(api().w2d_level_update(self.h, tiles, float(dt), int(want_plan)))

    def boundary(self) -> None:
        check(api().w2d_level_boundary(self.h))

    def commit(self, tiles: int) -> None:
        check(api().w2d_level_commit(self.h, tiles))

    def commit_strips(self, rows: int) -> None:
        check(api().w2d_level_commit_strips(self.h, rows))

    def wave(self) -> np.ndarray:
        return self.b.wave()
