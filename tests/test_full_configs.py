"""The three extension workloads at their stated sizes, on the GPU:

* the precision study: 2048 x 2048 cells at 10 m, steep ridged ground with 40 crest breaklines, a 1 km2 reservoir let
  go at t = 0, 1 h; fast FP32 (compensated residual) against FP64, both on the device;
* the batched ensemble: all 64 drawn storm/ARF members (sides 128-1024 at 20 m, 3 h each) in one device-resident
  launch in fast FP32, against the same launch in FP64;
* the learned closure inside that launch: all 32 held-out 1024 x 1024 crested members (3 h each) with the trained
  network, with a network answering zero, and plain.

The parity bars are the target's: max depth within 2 cm of FP64 for the ensemble, peak outfall discharge within
0.1 %, mass within 1e-7. The precision study's violent sloshing moves a handful of cells by centimetres in FP32
(the oracle's own float32 run does so at the sizes it finishes, ``test_full_run_parity.py``): there the bar is the
99.9th percentile within 2 cm and at most 0.1 % of the wet cells over it, with the worst cell reported.
"""

from __future__ import annotations

import os
import time

import numpy as np
import pytest

from wadi2d_amd import synth

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]


def test_precision_study_at_2048_for_one_hour():
    from wadi2d_amd.engine.precision import run_study

    t0 = time.time()
    scen = synth.precision_case(2048)
    h0 = synth.reservoir_depth(scen)
    scen.bed = None
    ref = run_study(scen, h0, "hip-f64")
    got = run_study(scen, h0, "hip")
    mask = scen.tables.grid.mask
    diff = np.abs(got.max_depth - ref.max_depth)[mask]
    wet = int((ref.max_depth[mask] > 0.01).sum())
    over = int((diff > 0.02).sum())
    q_err = abs(got.peak_q - ref.peak_q) / max(abs(ref.peak_q), 1e-30)
    print(f"C2 2048: steps {got.steps}/{ref.steps}, wall {got.wall_s:.1f}/{ref.wall_s:.1f} s; max depth worst "
          f"{diff.max():.3e} m, p99 {np.percentile(diff, 99):.2e}, p99.9 {np.percentile(diff, 99.9):.2e}, {over} of "
          f"{wet} wet cells over 2 cm; peak Q {got.peak_q:.1f}/{ref.peak_q:.1f} ({q_err:.1e}); mass "
          f"{got.mass_error:.1e}/{ref.mass_error:.1e}; {time.time() - t0:.0f} s", flush=True)
    assert wet > 20000
    assert np.percentile(diff, 99.9) <= 0.02 and over <= 1e-3 * wet
    assert q_err <= 1e-3
    assert abs(got.mass_error) <= 1e-7 and abs(ref.mass_error) <= 1e-12


def test_batched_ensemble_of_64_members_against_fp64():
    from wadi2d_amd.engine.ensemble import Ensemble, draw_members, max_depth_error, run_batched

    t0 = time.time()
    specs = draw_members(64)
    ens = Ensemble.build(specs, workers=min(16, os.cpu_count() or 1))
    built = time.time() - t0
    ref = run_batched(ens, "hip-f64", tune=False, launch=(512, 256), maxima_every=4)
    got = run_batched(ens, "hip", maxima_every=4)
    err = max_depth_error(got.max_depth, ref.max_depth, ens)
    print(f"C3 64 members ({ens.cells} cells, built in {built:.0f} s): FP32 {got.wall_s:.1f} s (launch {got.launch}), "
          f"FP64 {ref.wall_s:.1f} s; worst member max-depth error {max(err):.2e} m; mass FP32 "
          f"{max(abs(m) for m in got.mass_error):.1e}, FP64 {max(abs(m) for m in ref.mass_error):.1e}", flush=True)
    assert len(specs) == 64 and got.steps and all(s > 0 for s in got.steps)
    assert {s.side for s in specs} == {128, 256, 512, 768, 1024}
    assert max(err) <= 0.02
    assert max(abs(m) for m in got.mass_error) <= 1e-7
    assert max(abs(m) for m in ref.mass_error) <= 1e-12


def test_learned_closure_on_32_held_out_members_in_the_batched_launch(capsys):
    """The closure inside the batched launch at its stated size: a network answering zero leaves every member's
    maxima bit for bit the plain launch's; the trained network moves them (on crest faces) and keeps the books. The
    members' build (minutes of CPU work) reports each member as it comes, past the output capture."""

    def say(line: str) -> None:
        with capsys.disabled():
            print(line, flush=True)

    from wadi2d_amd.closure.evaluate import coarse_run
    from wadi2d_amd.closure.model import ClosureModel

    from wadi2d_amd.closure.data import held_out_members

    t0 = time.time()
    pairs = held_out_members(32, min(16, os.cpu_count() or 1), log=say)
    coarse, faces = [c for c, _ in pairs], [f for _, f in pairs]
    built = time.time() - t0
    path = os.path.join(os.path.dirname(os.path.dirname(__file__)), "wadi2d_amd", "closure", "trained.npz")
    trained = ClosureModel.load(path)
    zero = trained.with_(params=np.zeros_like(trained.params))  # every output 0: conveyance factor 1, no invert rise
    plain, ens = coarse_run(coarse, faces, None)
    say(f"[closure] plain launch {plain.wall_s:.1f} s")
    nothing, _ = coarse_run(coarse, faces, zero)
    say(f"[closure] zero network {nothing.wall_s:.1f} s")
    learned, _ = coarse_run(coarse, faces, trained)
    moved = max(float(np.abs(a - b)[f.tables.grid.mask].max()) for a, b, f in zip(learned.max_depth, plain.max_depth,
                                                                                  faces))  # fmt: skip
    print(f"C4 32 x 1024^2 ({ens.cells} cells, built in {built:.0f} s): plain {plain.wall_s:.1f} s, zero network "
          f"{nothing.wall_s:.1f} s, trained {learned.wall_s:.1f} s; largest max-depth change {moved:.2f} m; mass "
          f"{max(abs(m) for m in learned.mass_error):.1e}", flush=True)
    assert len(coarse) == 32 and all(c.tables.grid.shape == (1056, 1056) for c in coarse)
    for a, b in zip(nothing.max_depth, plain.max_depth, strict=True):
        assert np.array_equal(a, b)
    assert nothing.steps == plain.steps
    assert moved > 0.05
    assert max(abs(m) for m in learned.mass_error) <= 1e-7
    assert max(abs(m) for m in plain.mass_error) <= 1e-7
