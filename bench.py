"""Benchmark of the uniform-grid step on one MI355X: cell-updates/s and the fraction of HBM peak.

Workload: the full-extent plain FP32 step, 4096 x 4096 cells at 30 m (16,777,216 cells), seeded synthetic ground,
12-row sub-grid tables, a Gaussian storm cell on a triangular 3 h hyetograph, four free outflow segments
(``wadi2d_amd.synth.full_extent_case``). The run is taken hot to t = 1.5 h (the storm's peak) through the same
window loop a run uses, then W steps of warm-up and exactly K timed steps. A step is what the stepper does per
step: the face step and the cell update, the boundary cells, the FP64 boundary tally, a wave-speed reduction every
8 steps, the active-tile check every 16 steps and a source refresh every 30 s of simulated time. By default the loop
runs on the device (engine/device_run.py: captured graphs of 64 steps, no host in the loop) with the pipelined step
(a step's cell update made by the next step's kernel with its faces; ``--pipeline off`` for the two-kernel step);
``--loop host`` drives the same steps from Python through the backend protocol.

Printed: one JSON line. ``value`` counts domain cells (padding excluded) times steps over the wall time of the
timed region (graph-replayed steps, no events), with the inputs resident in HBM. ``roofline`` is the dominant
kernel's (the pipe kernel, or the face kernel with the pipeline off): its algorithmic bytes per launch over its mean
launch time, from HIP events on the library's stream in a pass of its own after the timed steps. ``config.maxima_every_4`` is a further pass with the running
maxima sampled every 4 steps, as a run samples them by default. ``cpu_baseline`` is the oracle's float
instantiation (the reference's float32 statement order, OpenMP over tiles) on the host's cores, at least 100 steps
from the same hot state, with the CPU's model and its nanoseconds per cell-update.

Other workloads (``--workload``): ``ensemble``, member-hours per wall-hour of 64 uneven storm/ARF members run
together on the device against the same members run one after another the plain way, both held against the
members run together in FP64; ``precision``, the steep 2048 x 2048 reservoir break in FP32 against FP64;
``learned``, the learned closure inside the batched launch against the plain table path.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, B/s
# algorithmic bytes per own cell (float32), each array counted once per cell per step
# The face step's algorithmic bytes per own cell, its share of STEP_BYTES: reads h,d,u,v,mx,my,qx,qy 32 + zcx,zcy 8
# + face statics 48 + two table row pairs 32, writes u,v,mx,my,qx,qy 24. The kernel moves 124 of them: it recomputes
# mx,my from u and the previous step's depths (4 B read instead of 8 read and 8 written) and carries the crest flag in
# the sign of the table top (the crest coefficient read only for a crest face).
FACE_BYTES = 144
CELL_BYTES = 52   # h,qx1,qy1,res,src 20 + Hc 4 + level pair 8 + writes qx,qy,h,d,res 20
STEP_BYTES = 176  # the whole step as one pass: 140 read + 36 written
# The pipelined device loop's pipe kernel makes, per launch, the last step's cell update and this step's faces: one
# step's work in one pass (STEP_BYTES per cell); after a step whose update the cell kernel made it has no update to
# make and does the face step alone (FACE_BYTES).


def domain_cells_of(tables) -> int:
    return int(tables.grid.mask.sum())


def log(*a) -> None:
    print(*a, file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--hot", type=float, default=5400.0, help="simulated seconds before the timed steps")
    ap.add_argument("--variant", choices=("fast", "exact"), default="fast")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="CPU-baseline sample: at least this long")
    ap.add_argument("--cpu-steps", type=int, default=100, help="CPU-baseline sample: at least this many steps")
    ap.add_argument("--events", choices=("on", "off"), default="on",
                    help="a pass with HIP events around the kernels, after the timed steps (device loop); around the "
                         "timed steps themselves with --loop host")
    ap.add_argument("--event-steps", type=int, default=200, help="steps of the event-timed pass (device loop)")
    ap.add_argument("--maxima-steps", type=int, default=500,
                    help="steps of the pass with the running maxima sampled every 4 steps (device loop); 0 skips it")
    ap.add_argument("--copy-probe", type=float, default=0.0, help="MiB of a plain device copy to time first")
    ap.add_argument("--loop", choices=("device", "host"), default="device")
    ap.add_argument("--graph", type=int, default=64, help="steps per captured graph (device loop)")
    ap.add_argument("--pipeline", choices=("on", "off"), default="on",
                    help="device loop: the pipelined step (a step's cell update made with the next step's faces)")
    ap.add_argument("--face-cache", choices=("on", "off"), default="on", help="per-face row cache of the face tables")
    ap.add_argument("--faces-threads", type=int, choices=(256, 512), default=512, help="face kernel workgroup size")
    ap.add_argument("--cells-threads", type=int, choices=(256, 512), default=256, help="cell kernel workgroup size")
    ap.add_argument("--workload", choices=("step", "ensemble", "precision", "learned"), default="step",
                    help="step: the 4096x4096 plain step (default); ensemble: the 64-member storm/ARF ensemble; "
                         "precision: FP32 against FP64 on the steep 2048x2048 reservoir break")  # fmt: skip
    ap.add_argument("--tables", choices=("host", "device"), default="host",
                    help="step: build the sub-grid tables on the host (OpenMP) or on the GPU")
    ap.add_argument("--losses", choices=("off", "on"), default="off",
                    help="step: the storm as a depth grid with SCS curve-number losses (CN 60-95), rates formed on "
                         "the device each window; ensemble: every member with its own curve numbers")  # fmt: skip
    ap.add_argument("--members", type=int, default=64, help="ensemble members")
    ap.add_argument("--serial-members", type=int, default=-1, help="members the serial baseline runs (-1: all)")
    ap.add_argument("--reference", choices=("on", "off"), default="on", help="the FP64 batched run (ensemble)")
    ap.add_argument("--check-members", type=int, default=8, help="learned: held-out members with 2 m runs")
    ap.add_argument("--held-out", type=int, default=32, help="learned: held-out 1024x1024 members in the launch")
    ap.add_argument("--workers", type=int, default=8, help="learned: processes building the members' tables")
    ap.add_argument("--maxima-every", type=int, default=4,
                    help="ensemble: sample the running maxima every N steps (the run loop's default, 4)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for {args.gpus})", file=sys.stderr)
        return 2
    if args.workload == "ensemble":
        return ensemble(args)
    if args.workload == "precision":
        return precision(args)
    if args.workload == "learned":
        return learned(args)

    from wadi2d_amd import synth
    from wadi2d_amd.engine.driver import Driver
    from wadi2d_amd.solver.hip_backend import copy_probe

    if args.copy_probe > 0:
        nbytes = int(args.copy_probe * 2**20)
        ms = copy_probe(nbytes, 10)
        log(f"[bench] copy probe: {nbytes} B read + {nbytes} B written in {ms:.4f} ms = {2 * nbytes / ms / 1e6:.0f} GB/s")
    from wadi2d_amd.engine.forcing import Rain, boundary_set
    from wadi2d_amd.solver.model import Model

    t0 = time.time()
    scen = synth.full_extent_case(args.size)
    log(f"[bench] ground {scen.name}: {time.time() - t0:.1f} s")
    t0 = time.time()
    if args.tables == "device":
        from wadi2d_amd.tables import build_tables

        scen.tables = build_tables(scen.grid, scen.bed, scen.tables_cfg, scen.manning, device=True)
    tables = scen.build()
    log(f"[bench] tables ({args.tables}): {time.time() - t0:.1f} s, {tables.grid.ntiles - 1} tiles")
    bset = boundary_set(synth.edge_outlets(scen), tables)
    if args.losses == "on":
        cn = 60.0 + 35.0 * np.random.default_rng(17).random(tables.grid.shape)
        rain = Rain(tables, bset, scen.cumulative_mm, curve_number=cn, depth_grid=scen.weight)
    else:
        rain = Rain(tables, bset, scen.cumulative_mm, scen.weight)
    del scen.bed
    backend = "hip" if args.variant == "fast" else "hip-exact"
    t0 = time.time()
    model = Model(tables, scen.numerics, backend, face_cache=args.face_cache == "on")
    model.set_depth(np.zeros(tables.grid.shape))
    model.backend.set_launch(args.faces_threads, args.cells_threads)
    drv = Driver(model, bset, rain)
    log(f"[bench] device upload: {time.time() - t0:.1f} s")
    b = model.backend
    st = model.stepper
    t0 = time.time()
    hot = args.hot
    made = 0
    if args.loop == "host":
        while st.t < hot - 1e-9:
            drv.run(min(st.t + 600.0, hot))
            log(f"[bench] hot start t={st.t:.0f} s steps={st.nsteps} dt={st.dt:.3f} active={b.nact}")
        log(f"[bench] hot start: {time.time() - t0:.1f} s")
        drv.run(1e12, max_steps=st.nsteps + args.warmup)
        b.synchronize()
        b.timing_read()
        b.timing(args.events == "on")
        n0, tsim0 = st.nsteps, st.t
        t0 = time.perf_counter()
        drv.run(1e12, max_steps=st.nsteps + args.steps)
        b.synchronize()
        wall = time.perf_counter() - t0
        b.timing(False)
        faces_ms, cells_ms, launches = b.timing_read()
        steps, tsim = st.nsteps - n0, st.t - tsim0
        bal = drv.accounts()
    else:
        from wadi2d_amd.engine.device_run import DeviceRun, MemberPlan

        plan = MemberPlan(1, tables.grid.ntiles, rain, bset, st.t, scen.duration_s, st.dt, st.nsteps,
                          st.refresh_due, rain.state)  # fmt: skip
        run = DeviceRun(b, [plan], scen.numerics, idle=st.idle, pipeline=args.pipeline == "on")
        state = run.state[0]
        while state.t < hot - 1e-9:
            state = run.steps(256, args.graph)[0]
            if state.nsteps % 1024 == 0:
                log(f"[bench] hot start t={state.t:.0f} s steps={state.nsteps} dt={state.dt:.3f} active={b.nact}")
        log(f"[bench] hot start: {time.time() - t0:.1f} s, t={state.t:.0f} s, {state.nsteps} steps")
        run.steps(args.warmup, args.graph)
        # the timed steps: captured graphs, no events; wall time only
        n0, tsim0 = run.state[0].nsteps, run.state[0].t
        t0 = time.perf_counter()
        state = run.steps(args.steps, args.graph)[0]  # waits for the last step
        wall = time.perf_counter() - t0
        steps, tsim = state.nsteps - n0, state.t - tsim0
        assert not state.done, "the run ended inside the timed steps"
        bal = run.balance(0, b.tally_volume())
        # per-kernel times in a pass of their own: HIP events around each kernel (the loop then runs uncaptured)
        faces_ms, cells_ms, launches = 0.0, 0.0, 0
        made = 0  # of those launches, the pipe kernel's that made a cell update
        if args.events == "on" and args.event_steps > 0:
            b.timing_read()
            b.timing(True)
            made0 = state.pipelined_updates
            state = run.steps(args.event_steps, args.graph)[0]
            b.timing(False)
            faces_ms, cells_ms, launches = b.timing_read()
            made = state.pipelined_updates - made0
            assert launches == args.event_steps and not state.done, (launches, state.done)
        # the run loop's own sampling: the running maxima every 4 steps, a run of its own from this state
        maxima4 = None
        if args.maxima_steps > 0:
            b.maxima_set(0.1, 0.05)
            plan4 = MemberPlan(1, tables.grid.ntiles, rain, bset, state.t, scen.duration_s, state.dt, state.nsteps,
                               True, rain.state)  # fmt: skip
            run4 = DeviceRun(b, [plan4], scen.numerics, maxima=4, idle=b.run_idle(), pipeline=args.pipeline == "on")
            run4.steps(args.warmup, args.graph)
            t1 = time.perf_counter()
            s4 = run4.steps(args.maxima_steps, args.graph)[0]
            wall4 = time.perf_counter() - t1
            assert not s4.done, "the run ended inside the maxima steps"
            maxima4 = {"maxima_every": 4, "steps": args.maxima_steps, "ms_per_step": wall4 / args.maxima_steps * 1e3,
                       "value": domain_cells_of(tables) * args.maxima_steps / wall4}
            log(f"[bench] maxima every 4: {args.maxima_steps} steps in {wall4:.3f} s "
                f"({maxima4['ms_per_step']:.3f} ms/step)")
    if args.loop == "host":
        maxima4 = None
    assert steps == args.steps and launches in (0, steps, args.event_steps), (steps, launches)
    launches = max(launches, 1)
    domain_cells = int(tables.grid.mask.sum())
    tiles = int(b.nact)
    value = domain_cells * steps / wall
    face_s = faces_ms / 1e3 / launches
    cell_s = cells_ms / 1e3 / launches
    # the dominant kernel: the pipe kernel in the pipelined device loop, else the face kernel
    pipelined = args.loop == "device" and args.pipeline == "on"
    kernel = "pipe_kernel" if pipelined else "faces_kernel"
    alg_bytes = (made * STEP_BYTES + (launches - made) * FACE_BYTES) / launches if pipelined else FACE_BYTES
    face_achieved = alg_bytes * tiles * 1024 / face_s if face_s > 0 else 0.0
    log(f"[bench] timed {steps} steps in {wall:.3f} s ({wall / steps * 1e3:.3f} ms/step), simulated "
        f"{tsim:.0f} s; faces {face_s * 1e3:.3f} ms, cells {cell_s * 1e3:.3f} ms per launch; "
        f"active tiles {tiles}; mass error {bal.relative:.2e}")

    cpu = None
    if args.cpu_seconds > 0 or args.cpu_steps > 0:
        cpu = cpu_baseline(model, tables, scen.numerics, args.cpu_seconds, args.cpu_steps,
                           st.dt if args.loop == "host" else state.dt)

    out = {
        "metric": "cell-updates/s",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": wall / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded ground, tables and storm built in-process)",
        "config": {
            "workload": f"plain FP32 step, {args.size}x{args.size} cells at 30 m, 12-row sub-grid tables, "
                        + (f"hot start at t={hot:.0f} s" if hot > 0 else "cold start (t=0, dry ground)")
                        + " of a Gaussian 80 mm/h storm, four free outflows"
                        + (", the storm as a depth grid with SCS losses (CN 60-95) formed on the device"
                           if args.losses == "on" else ""),
            "cells": domain_cells,
            "active_tiles": tiles,
            "variant": args.variant,
            "loop": args.loop,
            "hbm_frac_step": STEP_BYTES * value / HBM_PEAK,
            "pipeline": pipelined,
            "faces_ms": face_s * 1e3,  # the pipe kernel's when pipelined
            "cells_ms": cell_s * 1e3,
            "pipe_updates_share": made / launches if pipelined else None,
            "pipe_bytes_per_cell": alg_bytes,
            "kernel_times": (f"HIP events, a pass of {launches} steps after the timed ones" if args.loop == "device"
                             else "HIP events inside the timed steps"),
            "timed_path": ("captured graphs of %d steps, wall time" % args.graph if args.loop == "device"
                           else "host loop, wall time"),
            "mass_error": bal.relative,
            "maxima_every_4": maxima4,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": kernel,
            "achieved": face_achieved / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": face_achieved / HBM_PEAK,
            "traffic": measured_traffic(args, tiles, kernel),
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


def ensemble(args) -> int:
    """Member-hours per wall-hour of the storm/ARF ensemble: all members in one device-resident run (launch shape
    from the tuner) against the plain way, each member alone with its step loop on the host, one after another.
    Both in fast FP32; both held by max depth against the same members run together in FP64."""
    from wadi2d_amd.engine.ensemble import Ensemble, draw_members, max_depth_error, run_batched, run_serial

    t0 = time.time()
    specs = draw_members(args.members)
    ens = Ensemble.build(specs)
    if args.losses == "on":  # each member its own seeded curve numbers, 60-95
        for spec, sc in zip(specs, ens.scenarios, strict=True):
            sc.curve_number = 60.0 + 35.0 * np.random.default_rng(spec.seed + 1).random(sc.tables.grid.shape)
    log(f"[bench] {len(specs)} members, {ens.cells} cells, {ens.tables.grid.ntiles - 1} tiles: built in "
        f"{time.time() - t0:.1f} s")
    ref = None
    if args.reference == "on":
        ref = run_batched(ens, "hip-f64", graph_steps=args.graph, tune=False, launch=(512, 256),
                          maxima_every=args.maxima_every, log=log)  # fmt: skip
        log(f"[bench] FP64 batched: {ref.wall_s:.1f} s, steps {min(ref.steps)}..{max(ref.steps)}")
    got = run_batched(ens, "hip", graph_steps=args.graph, maxima_every=args.maxima_every, log=log)
    log(f"[bench] FP32 batched: {got.wall_s:.1f} s, launch {got.launch}, tuning {got.tuning}")
    n_serial = len(specs) if args.serial_members < 0 else min(args.serial_members, len(specs))
    members = list(range(n_serial))
    plain = run_serial(ens, "hip", members, args.maxima_every, log=log)
    hours = [sc.duration_s / 3600.0 for sc in ens.scenarios]
    batched_rate = sum(hours) / (got.wall_s / 3600.0)
    plain_rate = sum(hours[m] for m in members) / (plain.wall_s / 3600.0) if members else None
    err_b = max_depth_error(got.max_depth, ref.max_depth, ens) if ref else [None]
    err_p = max_depth_error(plain.max_depth, [ref.max_depth[m] for m in members], ens, members) if ref and members \
        else [None]
    out = {
        "metric": "ensemble member-hours per wall-hour",
        "value": batched_rate,
        "unit": "member-hours/hour",
        "n_gpus": 1,
        "steps": int(max(got.steps)),
        "warmup": 0,
        "ms_per_step": got.wall_s / max(got.steps) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded members, tables and storms built in-process)",
        "config": {
            "workload": f"{len(specs)} members, sides from 128..1024 at 20 m ({ens.cells} cells), own seeded ground "
                        "each, NRCS Type II storms of 20-120 mm over 1-6 h times an ARF of 0.6-1.0, 3 h each"
                        + (", SCS losses (CN 60-95 per cell) formed on the device" if args.losses == "on" else ""),
            "launch": list(got.launch) if got.launch else None,
            "maxima_every": args.maxima_every,
            "plain_member_hours_per_hour": plain_rate,
            "plain_members_timed": n_serial,
            "speedup_vs_plain": batched_rate / plain_rate if plain_rate else None,
            "max_depth_error_batched_m": max(err_b),
            "max_depth_error_plain_m": max(err_p),
            "mass_error_max": max(abs(e) for e in got.mass_error),
            "batched_wall_s": got.wall_s,
            "plain_wall_s": plain.wall_s,
            "fp64_wall_s": ref.wall_s if ref else None,
        },
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    return 0


def precision(args) -> int:
    """The steep 2048x2048 reservoir break, 1 h, in fast FP32 (compensated residual) and in FP64: max-depth, peak
    outfall discharge and mass errors of the FP32 run against the FP64 one."""
    from wadi2d_amd import synth
    from wadi2d_amd.engine.precision import run_study

    size = args.size if args.size != 4096 else 2048
    t0 = time.time()
    scen = synth.precision_case(size)
    h0 = synth.reservoir_depth(scen)
    scen.bed = None
    log(f"[bench] {scen.name}: built in {time.time() - t0:.1f} s; reservoir {float(h0.sum()) * 100.0:.3e} m3")
    ref = run_study(scen, h0, "hip-f64", log=log)
    got = run_study(scen, h0, "hip", log=log)
    mask = scen.tables.grid.mask
    diff = np.abs(got.max_depth - ref.max_depth)[mask]
    err = float(diff.max())
    q_err = abs(got.peak_q - ref.peak_q) / max(abs(ref.peak_q), 1e-30)
    cells = int(mask.sum())
    wet = int((ref.max_depth[mask] > 0.01).sum())
    out = {
        "metric": "max-depth error FP32 against FP64",
        "value": err,
        "unit": "m",
        "n_gpus": 1,
        "steps": got.steps,
        "warmup": 0,
        "ms_per_step": got.wall_s / got.steps * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded steep ground, crests and reservoir built in-process)",
        "config": {
            "workload": f"{size}x{size} cells at 10 m, ridged ground (median slope 20 %), 40 crest breaklines, a 1 km2 "
                        "reservoir 5 m above a crest that breaks at t=0, four free edges, 1 h",
            "max_depth_error_p99_m": float(np.percentile(diff, 99)),
            "max_depth_error_p999_m": float(np.percentile(diff, 99.9)),
            "cells_over_2cm": int((diff > 0.02).sum()),
            "cells_wet": wet,
            "peak_q_fp64": ref.peak_q,
            "peak_q_fp32": got.peak_q,
            "peak_q_rel_error": q_err,
            "mass_error_fp32": got.mass_error,
            "mass_error_fp64": ref.mass_error,
            "steps_fp64": ref.steps,
            "cells": cells,
            "wall_s_fp32": got.wall_s,
            "wall_s_fp64": ref.wall_s,
        },
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    return 0


def learned(args) -> int:
    """The learned closure inside the batched launch: held-out 1024 x 1024 members at 20 m (crested ground, 2 m
    pixels), run together with and without the closure, and one after another the plain way; and its gain on
    smaller held-out members against their own 2 m runs."""
    from wadi2d_amd.closure.data import fine_runs
    from wadi2d_amd.closure.evaluate import coarse_run, errors
    from wadi2d_amd.closure.model import ClosureModel
    from wadi2d_amd.engine.ensemble import Ensemble, run_serial

    from wadi2d_amd.closure.data import held_out_members

    model = ClosureModel.load(os.path.join(ROOT, "wadi2d_amd", "closure", "trained.npz"))
    # the members' ground, tables and face statistics on the host, in worker processes
    pairs = held_out_members(args.held_out, args.workers, log=log)
    coarse, faces = [c for c, _ in pairs], [f for _, f in pairs]
    plain, _ = coarse_run(coarse, faces, None, graph_steps=args.graph)
    log(f"[bench] batched plain: {plain.wall_s:.1f} s")
    got, _ = coarse_run(coarse, faces, model, graph_steps=args.graph)
    log(f"[bench] batched learned: {got.wall_s:.1f} s")
    ens = Ensemble.from_scenarios(coarse)
    n_serial = len(coarse) if args.serial_members < 0 else min(args.serial_members, len(coarse))
    serial = run_serial(ens, "hip", list(range(n_serial)), 1, log=log)
    serial_wall = serial.wall_s
    diff = max(float(np.abs(a - b)[f.tables.grid.mask].max()) for a, b, f in zip(got.max_depth, plain.max_depth, faces))
    hours = 3.0 * len(coarse)
    rate = hours / (got.wall_s / 3600.0)
    rate_plain_batched = hours / (plain.wall_s / 3600.0)
    rate_serial = 3.0 * n_serial / (serial.wall_s / 3600.0) if n_serial else None
    check = {}
    if args.check_members > 0:
        from wadi2d_amd.closure.data import middle_runs

        col, pairs = fine_runs(list(range(5000, 5000 + args.check_members)), 64, 5400.0, chunk=64, storm_seed=1)
        cs = [c for _, c in pairs]
        jp, _ = coarse_run(cs, col.coarse, None)
        jl, _ = coarse_run(cs, col.coarse, model)
        mid, mrun = middle_runs(pairs, col.coarse, 64, 2, 5400.0)
        e_plain, e_learned = errors(jp.max_depth, col.dmax, col.coarse), errors(jl.max_depth, col.dmax, col.coarse)
        e_mid = errors(mid.dmax, col.dmax, col.coarse)
        # the plain way at a finer grid, member after member: what the reference would run for the accuracy the
        # closure buys on the coarse grid
        ser10 = run_serial(mid.ens, "hip", None, 1, log=log)
        ser20 = run_serial(Ensemble.from_scenarios(cs), "hip", None, 1, log=log)
        check = {"plain_20m": e_plain.__dict__, "learned_20m": e_learned.__dict__, "plain_10m": e_mid.__dict__,
                 "wall_s": {"plain_20m": jp.wall_s, "learned_20m": jl.wall_s, "plain_10m": mrun.wall_s,
                            "plain_10m_serial": ser10.wall_s, "plain_20m_serial": ser20.wall_s},
                 "matched_accuracy": {
                     "learned_20m_batched_rmse_m": e_learned.rmse, "plain_10m_serial_rmse_m": e_mid.rmse,
                     "plain_20m_serial_rmse_m": e_plain.rmse,
                     "speedup_learned_batched_vs_plain_10m_serial": ser10.wall_s / jl.wall_s,
                     "speedup_learned_batched_vs_plain_20m_serial": ser20.wall_s / jl.wall_s,
                     "target_rmse_m": 0.02, "target_met": bool(e_learned.rmse <= 0.02)}}  # fmt: skip
    out = {
        "metric": "ensemble member-hours per wall-hour, batched with the learned closure",
        "value": rate,
        "unit": "member-hours/hour",
        "n_gpus": 1,
        "steps": int(max(got.steps)),
        "warmup": 0,
        "ms_per_step": got.wall_s / max(got.steps) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 state, bf16 network on matrix cores",
        "data": "synthetic (seeded crested ground, tables and storms built in-process)",
        "config": {
            "workload": f"{len(coarse)} held-out members of 1024x1024 cells at 20 m (crests 10-80 m wide, 0.3-1.5 m "
                        "high, 10 per km2), NRCS Type II storms of 40-120 mm over 1-3 h, 3 h each",
            "closure": {"conveyance": model.conveyance, "rise_scale": model.rise_scale},
            "batched_plain_member_hours_per_hour": rate_plain_batched,
            "serial_plain_member_hours_per_hour": rate_serial,
            "serial_members_timed": n_serial,
            "speedup_vs_serial_plain": rate / rate_serial if rate_serial else None,
            "closure_cost": got.wall_s / plain.wall_s - 1.0,
            "max_depth_change_learned_vs_plain_m": diff,
            "against_2m_runs": check,
            "wall_s": {"plain_batched": plain.wall_s, "learned_batched": got.wall_s, "serial": serial_wall},
        },
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    return 0


def measured_traffic(args, tiles: int, kernel: str):
    """HBM bytes per launch of ``kernel`` from the committed PMC summary of this workload (FETCH_SIZE and WRITE_SIZE,
    each scaled by what the counter reads for a copy of known size), or None for another workload or launch."""
    path = os.path.join(ROOT, "profiles", "traffic_full_extent_4096.json")
    if args.size != 4096 or not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f)
    if tiles != 16641 or rec["launch"]["faces_threads"] != args.faces_threads:
        return None
    for per in (rec["per_launch"], rec.get("two_kernel_step", {}).get("per_launch", {})):
        if kernel in per:
            return per[kernel]["bytes"]
    return None


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(model, tables, num, seconds: float, min_steps: int, dt: float):
    """The oracle's float32 step on the host's cores, from the GPU's hot state: at least ``min_steps`` steps and at
    least ``seconds`` of stepping."""
    from oracle.oracle import Numerics as ONumerics
    from oracle.oracle import OracleBackend

    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    t0 = time.time()
    ob = OracleBackend(tables, ONumerics.like(num), prec=4, threads=cores)
    for f in ("h", "d", "u", "v", "mx", "my", "qx", "qy", "res", "om", "src"):
        ob.set_field(f, model.backend.get_field(f))
    ob.set_boundary(model.backend.bc)
    cpu = cpu_model()
    log(f"[bench] cpu baseline set up in {time.time() - t0:.1f} s on {cores} threads of {cpu}")
    cells = domain_cells_of(tables)
    done, spent = 0, 0.0
    while spent < seconds or done < min_steps:
        t = time.perf_counter()
        ob.step(dt, done % 8 == 0, done % 16 == 0)
        spent += time.perf_counter() - t
        done += 1
        if done % 10 == 0:
            log(f"[bench] cpu step {done}: {spent:.1f} s")
    value = cells * done / spent
    return {
        "value": value,
        "unit": "cell-updates/s",
        "cores": cores,
        "kind": "port",
        "cpu": cpu,
        "ns_per_cell_step": 1e9 / value,
        "core_ns_per_cell_step": 1e9 * cores / value,
        "sample": f"{done} steps ({spent:.1f} s) of the full {cells}-cell grid from the same hot state (float32, "
                  f"the reference's statement order, OpenMP over tiles on {cores} threads)",
    }


if __name__ == "__main__":
    sys.exit(main())
