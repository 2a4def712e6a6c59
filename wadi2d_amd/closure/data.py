"""Training and judging data for the closure: fine-grid runs of this solver, coarse-grained.

A member is one seed of ``synth.crest_case`` sampled twice: a fine grid of 2 m cells (one 2 m pixel each) and a
coarse grid of 20 m cells (10 x 10 of the same pixels in its tables). The fine members run together in one
device-resident run (engine/ensemble.py); every ``chunk`` steps their fields are coarse-grained onto the coarse grid:
mean depth is the fine cells' mean (volume over area), the coarse level follows from the coarse cell table, and the
discharge through a coarse face is the sum of the fine face discharges along it.

A sample is one crest face of the coarse grid (crest height over the invert at least ``CREST_MIN``) with water on one
side, at one snapshot: the 16 features (closure/features.py) and two targets with their masks:

* the sill over the invert, geometric (always valid);
* log of the conveyance factor c that makes Manning's relation Q = (c T(e))^(5/3) sqrt(S) hold with the fine
  discharge Q, the coarse surface slope S across the face and the coarse table's T at the upwind level over the
  sill; valid where the face carries flow down the surface slope (S > 1e-3, |Q| > 0.02 m3/s, e > 5 cm).
"""

from __future__ import annotations

import os

from dataclasses import dataclass

import numpy as np

from wadi2d_amd import synth
from wadi2d_amd.closure.features import NSTATIC, dynamic, face_statics
from wadi2d_amd.engine.storms import ScsStorm
from wadi2d_amd.solver.state import from_state, level_of

RATIO = 10  # fine cells per coarse cell side
COARSE_DX = 20.0
CREST_MIN = 0.3  # metres of crest over the invert that make a face a closure face
LOGC_CLIP = 1.5
SILL_CLIP = 4.0


def member_pair(seed: int, n: int, storm, duration_s: float):
    """(fine, coarse) scenarios of one seed on an n x n coarse grid."""
    extent = n * COARSE_DX
    coarse = synth.crest_case(n, COARSE_DX, RATIO, seed, storm, duration_s, extent, f"coarse {seed}")
    fine = synth.crest_case(n * RATIO, COARSE_DX / RATIO, 1, seed, storm, duration_s, extent, f"fine {seed}")
    return fine, coarse


def draw_storm(rng) -> ScsStorm:
    return ScsStorm(float(rng.uniform(40.0, 120.0)), float(rng.uniform(1.0, 3.0)), 1.0)


def held_out_member(m: int, n: int = 1024, duration_s: float = 3 * 3600.0):
    """Held-out member m of the closure's evaluation (seed 9000 + m, never trained on): the coarse n x n crested
    scenario at 20 m from 2 m pixels with its drawn storm, and its face statistics; the pixels are dropped after."""
    storm = draw_storm(np.random.default_rng(9000 + m))
    c = synth.crest_case(n, COARSE_DX, RATIO, 9000 + m, storm, duration_s, n * COARSE_DX, f"coarse {9000 + m}")
    cf = CoarseFaces.of(c)
    c.bed = None
    return c, cf


def _one_thread() -> None:
    os.environ["OMP_NUM_THREADS"] = "1"


def held_out_members(count: int, workers: int, log=None) -> list:
    """Members 0 .. count-1 built in fresh processes (spawned, one thread each: a forked child of a process whose
    OpenMP pool is running can deadlock), in member order."""
    import multiprocessing
    import time
    from concurrent.futures import ProcessPoolExecutor

    t0 = time.time()
    out = []
    ctx = multiprocessing.get_context("spawn")
    with ProcessPoolExecutor(max_workers=max(1, workers), mp_context=ctx, initializer=_one_thread) as pool:
        for m, pair in enumerate(pool.map(held_out_member, range(count))):
            out.append(pair)
            if log is not None:
                log(f"[closure] held-out member {m} built, {time.time() - t0:.0f} s")
    return out


@dataclass
class CoarseFaces:
    """Static description of one coarse member's faces for both axes (x: axis 1, y: axis 0)."""

    tables: object
    stat: dict  # axis -> face_statics(...)
    hc: np.ndarray
    dtab: np.ndarray

    @classmethod
    def of(cls, scen) -> CoarseFaces:
        t = scen.build()
        st = {ax: face_statics(t, scen.bed.z, RATIO, ax) for ax in (1, 0)}
        g = t.grid
        return cls(t, st, g.from_tiles(t.hc).astype(np.float64), g.from_tiles(t.dtab).astype(np.float64))

    def closure_faces(self, axis: int) -> np.ndarray:
        s = self.stat[axis]
        return s["valid"] & (s["static"][..., 5] >= CREST_MIN)


def face_T(tables, axis: int, jj: np.ndarray, ii: np.ndarray, e: np.ndarray) -> np.ndarray:
    """The face table's T = K^(3/5) at depth e over the invert (common.cuh:64-90, in float64)."""
    g = tables.grid
    name = "x" if axis == 1 else "y"
    ftab = getattr(tables, "ftab" + name)
    E = g.from_tiles(getattr(tables, "ef" + name))[jj, ii].astype(np.float64)
    t, j, i = g.tile_of[jj // 32, ii // 32], jj % 32, ii % 32
    rows = ftab[t, j, i].astype(np.float64)  # (n, rows, 2)
    m = rows.shape[1] - 1
    out = np.empty(e.size)
    top = e >= E
    a1, a0 = rows[:, m, 0], rows[:, m - 1, 0]
    c = (m - 1) / m
    W = (a1 - a0) / (E * (1.0 - c * c))
    A = a1 + W * (e - E)
    out[top] = (rows[:, m, 1] * A / np.maximum(a1, 1e-12))[top]
    s = np.sqrt(np.clip(e / E, 0.0, None)) * m
    r = np.clip(np.floor(s).astype(np.int64), 0, m - 1)
    c0, c1 = r / m, (r + 1) / m
    f = (e - c0 * c0 * E) / ((c1 * c1 - c0 * c0) * E)
    k = np.arange(e.size)
    T = rows[k, r, 1] + f * (rows[k, r + 1, 1] - rows[k, r, 1])
    out[~top] = T[~top]
    return out


def coarse_grain(h: np.ndarray, qx: np.ndarray, qy: np.ndarray, n: int, r: int = RATIO):
    """Fine domain fields (n r square) -> coarse domain mean depth (n x n) and face discharges: Qx at the west face
    of (j, i), Qy at its north face."""
    nf = n * r
    hb = h.reshape(n, r, n, r).mean(axis=(1, 3))
    Qx = np.zeros((n, n))
    Qx[:, 1:] = qx[:, r : nf : r].reshape(n, r, n - 1).sum(axis=1)
    Qy = np.zeros((n, n))
    Qy[1:, :] = qy[r : nf : r, :].reshape(n - 1, n, r).sum(axis=2)
    return hb, Qx, Qy


class Collector:
    """Coarse-grains snapshots of fine members (``ratio`` fine cells per coarse cell side) into samples, and keeps
    each member's coarse-grained max level."""

    def __init__(self, ens, pairs: list, n: int, ratio: int = RATIO, samples: bool = True, coarse=None) -> None:
        self.ens = ens
        self.n = n
        self.ratio = ratio
        self.keep_samples = samples
        self.coarse = coarse if coarse is not None else [CoarseFaces.of(c) for _, c in pairs]
        self.samples: list[tuple[np.ndarray, np.ndarray, np.ndarray]] = []
        self.dmax = [np.zeros(cf.hc.shape) for cf in self.coarse]
        self.snapshots = 0

    def __call__(self, model, state) -> None:
        b = model.backend
        g = self.ens.tables.grid
        fields = {f: from_state(g, b.get_field(f)).astype(np.float64) for f in ("h", "qx", "qy")}
        n, nf = self.n, self.n * self.ratio
        for m, cf in enumerate(self.coarse):
            block = {f: self.ens.member_field(m, a)[1 : nf + 1, 1 : nf + 1] for f, a in fields.items()}
            hb, Qx, Qy = coarse_grain(block["h"], block["qx"], block["qy"], n, self.ratio)
            shape = cf.hc.shape
            hbar = np.zeros(shape)
            hbar[1 : n + 1, 1 : n + 1] = hb
            d = level_of(hbar, cf.hc, cf.dtab)
            self.dmax[m] = np.maximum(self.dmax[m], d)
            if not self.keep_samples:
                continue
            Q = {1: np.zeros(shape), 0: np.zeros(shape)}
            Q[1][1 : n + 1, 1 : n + 1] = Qx
            Q[0][1 : n + 1, 1 : n + 1] = Qy
            for axis in (1, 0):
                self._faces(cf, axis, hbar, d, Q[axis])
        self.snapshots += 1

    def _faces(self, cf: CoarseFaces, axis: int, hbar, d, Q) -> None:
        t = cf.tables
        g = t.grid
        st = cf.stat[axis]
        jj, ii = np.nonzero(cf.closure_faces(axis))
        lj, li = (jj, ii - 1) if axis == 1 else (jj - 1, ii)
        name = "x" if axis == 1 else "y"
        fl = g.from_tiles(getattr(t, "fl" + name))[jj, ii].astype(np.float64)
        fr = g.from_tiles(getattr(t, "fr" + name))[jj, ii].astype(np.float64)
        eL, eR = d[lj, li] - fl, d[jj, ii] - fr
        wet = np.maximum(eL, eR) > 0.01
        if not wet.any():
            return
        jj, ii, lj, li, eL, eR = jj[wet], ii[wet], lj[wet], li[wet], eL[wet], eR[wet]
        stat = st["static"][jj, ii]
        x = np.concatenate([stat, dynamic(eL, eR, hbar[lj, li], hbar[jj, ii], stat[:, 5])], axis=1)
        sill = st["sill"][jj, ii]
        etaL, etaR = t.zmin[lj, li] + d[lj, li], t.zmin[jj, ii] + d[jj, ii]
        S = np.abs(etaL - etaR) / g.dx
        q = Q[jj, ii]
        e_face = np.maximum(np.maximum(eL, eR) - sill, 0.0)
        ok = (S > 1e-3) & (np.abs(q) > 0.02) & (e_face > 0.05) & (np.sign(q) == np.sign(etaL - etaR))
        logc = np.zeros(jj.size)
        if ok.any():
            T = face_T(t, axis, jj[ok], ii[ok], e_face[ok])
            logc[ok] = 0.6 * np.log(np.abs(q[ok]) / np.sqrt(S[ok])) - np.log(np.maximum(T, 1e-9))
        target = np.stack([np.clip(logc, -LOGC_CLIP, LOGC_CLIP), np.clip(sill, 0.0, SILL_CLIP)], axis=1)
        mask = np.stack([ok & (np.abs(logc) < 3.0 * LOGC_CLIP), np.ones(jj.size, bool)], axis=1).astype(np.float32)
        self.samples.append((x.astype(np.float32), target.astype(np.float32), mask))

    def arrays(self):
        if not self.samples:
            return np.zeros((0, NSTATIC + 5), np.float32), np.zeros((0, 2), np.float32), np.zeros((0, 2), np.float32)
        return tuple(np.concatenate(a) for a in zip(*self.samples, strict=True))


def middle_runs(pairs: list, coarse: list, n: int, ratio: int, duration_s: float, chunk: int = 64, log=None):
    """The members of ``pairs`` run the plain way on a grid ``ratio`` times finer than the coarse one (same
    pixels, 20 m / ratio cells), coarse-grained onto the coarse grid as the fine runs are. Returns the collector
    and the run (its wall time)."""
    from wadi2d_amd.engine.ensemble import Ensemble, run_batched

    scens = []
    for (fine, c), cf in zip(pairs, coarse, strict=True):
        seed = int(c.name.split()[-1])
        scens.append(synth.crest_case(n * ratio, COARSE_DX / ratio, RATIO // ratio, seed, c.cumulative_mm,
                                      duration_s, n * COARSE_DX, f"middle {seed}"))  # fmt: skip
    ens = Ensemble.from_scenarios(scens)
    col = Collector(ens, None, n, ratio, samples=False, coarse=coarse)
    run = run_batched(ens, "hip", graph_steps=64, chunk=chunk, tune=False, launch=(512, 256), maxima_every=0, log=log,
                      on_chunk=col)  # fmt: skip
    return col, run


def fine_runs(seeds: list[int], n: int, duration_s: float, chunk: int = 256, log=None, storm_seed: int = 0):
    """Run the fine members of ``seeds`` together and coarse-grain them every ``chunk`` steps. Returns the
    collector (samples, coarse-grained max levels) and the (fine, coarse) scenario pairs."""
    from wadi2d_amd.engine.ensemble import Ensemble, run_batched

    rng = np.random.default_rng(storm_seed)
    pairs = [member_pair(s, n, draw_storm(rng), duration_s) for s in seeds]
    ens = Ensemble.from_scenarios([f for f, _ in pairs])
    col = Collector(ens, pairs, n)
    run_batched(ens, "hip", graph_steps=64, chunk=chunk, tune=False, launch=(512, 256), maxima_every=0, log=log,
                on_chunk=col)  # fmt: skip
    return col, pairs
