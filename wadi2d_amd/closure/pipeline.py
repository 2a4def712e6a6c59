"""Build the closure end to end on one GPU: fine-grid runs of training members, coarse-grained samples, training on
the GPU, and the check on held-out members against their own fine-grid runs, with and without the closure.

    python -m wadi2d_amd.closure.pipeline --out wadi2d_amd/closure/trained.npz
"""

from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from wadi2d_amd.closure.data import fine_runs
from wadi2d_amd.closure.evaluate import coarse_run, errors
from wadi2d_amd.closure.model import ClosureModel, train


def log(*a) -> None:
    print(*a, file=sys.stderr, flush=True)


def score(model: ClosureModel | None, col, pairs) -> dict:
    """Errors of a coarse run (plain when ``model`` is None) against the coarse-grained fine runs in ``col``."""
    run, _ = coarse_run([c for _, c in pairs], col.coarse, model, log=None)
    return errors(run.max_depth, col.dmax, col.coarse).__dict__


CALIBRATION = [(conv, scale) for conv in (False, True) for scale in (0.1, 0.2, 0.3, 0.4, 0.5, 0.75, 1.0)]


def calibrate(model: ClosureModel, col, pairs) -> tuple[ClosureModel, dict]:
    """The conveyance switch and the rise scale that give the least max-depth RMSE on the training members."""
    scores = {"plain": score(None, col, pairs)}
    best, best_rmse = None, float("inf")
    for conv, scale in CALIBRATION:
        m = model.with_(conveyance=conv, rise_scale=scale)
        e = score(m, col, pairs)
        scores[f"conveyance={conv},rise={scale}"] = e
        log(f"[calibrate] conveyance {conv} rise x{scale}: rmse {e['rmse']:.4f} (plain {scores['plain']['rmse']:.4f})")
        if e["rmse"] < best_rmse:
            best, best_rmse = m, e["rmse"]
    return best, scores


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--train-members", type=int, default=32)
    ap.add_argument("--held-out-members", type=int, default=8)
    ap.add_argument("--n", type=int, default=64, help="coarse cells per side of a member (20 m)")
    ap.add_argument("--duration", type=float, default=5400.0)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--max-samples", type=int, default=2_000_000)
    ap.add_argument("--out", default="")
    args = ap.parse_args()
    t0 = time.time()
    seeds = list(range(1000, 1000 + args.train_members))
    col, pairs = fine_runs(seeds, args.n, args.duration, chunk=64, log=log)
    x, t, m = col.arrays()
    log(f"[pipeline] {x.shape[0]} samples from {col.snapshots} snapshots, {m[:, 0].mean():.2f} with a conveyance "
        f"target, in {time.time() - t0:.0f} s")  # fmt: skip
    if x.shape[0] > args.max_samples:
        keep = np.random.default_rng(0).choice(x.shape[0], args.max_samples, replace=False)
        x, t, m = x[keep], t[keep], m[keep]
    model, history = train(x, t, m, steps=args.steps, log=log)
    model, scores = calibrate(model, col, pairs)
    log(f"[pipeline] chosen: conveyance {model.conveyance}, rise x{model.rise_scale}")
    if args.out:
        model.save(args.out)
    held = list(range(5000, 5000 + args.held_out_members))
    hcol, hpairs = fine_runs(held, args.n, args.duration, chunk=64, log=log, storm_seed=1)
    held_scores = {"plain": score(None, hcol, hpairs), "learned": score(model, hcol, hpairs)}
    print(json.dumps({"samples": int(x.shape[0]), "history": history, "calibration": scores, "held_out": held_scores,
                      "chosen": {"conveyance": model.conveyance, "rise_scale": model.rise_scale},
                      "wall_s": time.time() - t0}), flush=True)  # fmt: skip
    return 0


if __name__ == "__main__":
    sys.exit(main())
