"""The closure checked: coarse runs of held-out members with and without it against the same members' fine-grid
runs coarse-grained (closure/data.py), and the plain table path as the reference point. Errors are over the coarse
cells the water reaches in either run (max level over 2 cm), the free outflow column left out."""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from wadi2d_amd.closure.data import CoarseFaces
from wadi2d_amd.closure.model import ClosureModel, install
from wadi2d_amd.engine.ensemble import Ensemble, run_batched


def coarse_run(coarse_scens: list, faces: list[CoarseFaces], closure: ClosureModel | None, backend: str = "hip",
               log=None, graph_steps: int = 64):  # fmt: skip
    """All coarse members in one device-resident run, with the closure on their crest faces or without it."""
    ens = Ensemble.from_scenarios(coarse_scens)

    def prepare(model):
        n = install(model.backend, closure, [(cf, sl.row0) for cf, sl in zip(faces, ens.slots, strict=True)])
        if log is not None:
            log(f"[closure] {n} faces corrected")

    return run_batched(ens, backend, graph_steps=graph_steps, chunk=512, tune=False, launch=(512, 256), maxima_every=1,
                       prepare=prepare if closure is not None else None, log=log), ens  # fmt: skip


@dataclass
class Errors:
    rmse: float
    max_abs: float
    bias: float
    cells: int


def errors(run: list[np.ndarray], truth: list[np.ndarray], faces: list[CoarseFaces]) -> Errors:
    d = []
    for r, t, cf in zip(run, truth, faces, strict=True):
        mask = cf.tables.grid.mask.copy()
        n = int(mask[1].sum())
        mask[:, n] = False  # the free outflow column
        wet = mask & ((r > 0.02) | (t > 0.02))
        d.append((r - t)[wet])
    a = np.concatenate(d) if d else np.zeros(0)
    if a.size == 0:
        return Errors(0.0, 0.0, 0.0, 0)
    return Errors(float(np.sqrt((a * a).mean())), float(np.abs(a).max()), float(a.mean()), int(a.size))
