"""The learned closure inside the step (csrc/step.hip closure_kernel and the learned face variant): a network that
answers (0, 0) leaves every field exactly as the plain table path computes it; a network that raises the inverts
of a line of faces by 4 m holds a 2 m dam break behind that line; and the coarse-graining and face sample formation
of the training data (closure/data.py) on known fields."""

from __future__ import annotations

import numpy as np
import pytest

from tests.cases import rough_dam_break
from wadi2d_amd.closure.data import RATIO, coarse_grain
from wadi2d_amd.closure.features import NSTATIC, tile_index_of
from wadi2d_amd.closure.model import ClosureModel
from wadi2d_amd.config import Numerics

NPARAM = 5378


def _zero_model(rise_bias: float = 0.0) -> ClosureModel:
    p = np.zeros(NPARAM, np.float32)
    p[-1] = rise_bias  # b3[1]: the invert-rise output
    return ClosureModel(p, np.zeros(16, np.float32), np.zeros(16, np.float32), conveyance=True, rise_scale=1.0)


def _line_faces(grid, column: int):
    """The west faces of every domain cell of one padded column."""
    jj = np.arange(1, grid.shape[0] - 1)
    jj = jj[grid.mask[jj, column] & grid.mask[jj, column - 1]]
    ii = np.full(jj.size, column)
    return tile_index_of(grid, jj, ii), tile_index_of(grid, jj, ii - 1), np.ones(jj.size, np.int32)


def test_coarse_graining_conserves_volume_and_sums_face_discharge():
    n = 6
    rng = np.random.default_rng(2)
    h = rng.random((n * RATIO, n * RATIO))
    qx = rng.normal(size=h.shape)
    qy = rng.normal(size=h.shape)
    hb, Qx, Qy = coarse_grain(h, qx, qy, n)
    assert hb.sum() * RATIO**2 == pytest.approx(h.sum())
    assert Qx[2, 3] == pytest.approx(qx[2 * RATIO : 3 * RATIO, 3 * RATIO].sum())
    assert Qy[4, 1] == pytest.approx(qy[4 * RATIO, RATIO : 2 * RATIO].sum())
    assert not Qx[:, 0].any() and not Qy[0, :].any()


@pytest.mark.gpu
def test_a_closure_that_answers_nothing_changes_nothing():
    from wadi2d_amd.solver.model import Model

    tables, h0 = rough_dam_break()
    num = Numerics(dt_every=1, dt_max=0.2)
    plain = Model(tables, num, "hip-exact")
    learned = Model(tables, num, "hip-exact")
    r, l, ax = _line_faces(tables.grid, 40)
    stat = np.random.default_rng(0).random((r.size, NSTATIC)).astype(np.float32)
    m = _zero_model()
    learned.backend.closure_set(r, l, ax, stat, m.norm, m.params, True)
    for mdl in (plain, learned):
        mdl.set_depth(h0)
        mdl.advance(20.0)
    assert plain.stepper.nsteps == learned.stepper.nsteps
    for f in ("h", "d", "u", "v", "qx", "qy"):
        assert np.array_equal(plain.field(f), learned.field(f)), f


def _flat_dam():
    from wadi2d_amd.config import Tables
    from wadi2d_amd.grid import TileGrid
    from wadi2d_amd.tables import ArrayBed, build_tables

    grid = TileGrid.from_mask(np.ones((60, 90), bool), 0.0, 0.0, 10.0)
    tables = build_tables(grid, ArrayBed(np.zeros(grid.shape), 10.0), Tables(), 0.03)
    h0 = np.zeros(grid.shape)
    h0[:, :30] = 2.0
    return tables, h0


@pytest.mark.gpu
def test_a_closure_that_raises_a_line_of_faces_holds_the_water():
    """A 2 m dam break on flat ground: with the inverts of one line of faces 4 m up, no water crosses it."""
    from wadi2d_amd.solver.model import Model

    tables, h0 = _flat_dam()
    num = Numerics(dt_every=1, dt_max=0.2)
    runs = {}
    for rise in (0.0, 4.0):
        mdl = Model(tables, num, "hip")
        r, l, ax = _line_faces(tables.grid, 36)
        m = _zero_model(rise)
        mdl.backend.closure_set(r, l, ax, np.zeros((r.size, NSTATIC), np.float32), m.norm, m.params, True)
        mdl.set_depth(h0)
        mdl.advance(60.0)
        runs[rise] = mdl.field("h")
    # the dam sits on padded columns 1..29; the raised line is the west face of column 36
    east = np.s_[:, 37:]
    print("water east of the line:", float(runs[0.0][east].sum()), float(runs[4.0][east].sum()), flush=True)
    assert runs[0.0][east].sum() > 1.0
    assert runs[4.0][east].sum() == 0.0


@pytest.mark.gpu
def test_the_trained_closure_beats_the_plain_table_path_on_unseen_ground():
    """The committed network on four members it never saw (seeds apart from training and calibration): their 20 m
    runs with and without it against their own 2 m runs coarse-grained. The plain table path is the yardstick: the
    learned run must cut its max-depth RMSE by at least 15 %."""
    from wadi2d_amd.closure.data import fine_runs
    from wadi2d_amd.closure.evaluate import coarse_run, errors
    from wadi2d_amd.closure.model import ClosureModel

    model = ClosureModel.load("wadi2d_amd/closure/trained.npz")
    col, pairs = fine_runs([7001, 7002, 7003, 7004], 64, 5400.0, chunk=64, storm_seed=7)
    coarse = [c for _, c in pairs]
    plain, _ = coarse_run(coarse, col.coarse, None)
    learned, _ = coarse_run(coarse, col.coarse, model)
    ep = errors(plain.max_depth, col.dmax, col.coarse)
    el = errors(learned.max_depth, col.dmax, col.coarse)
    print(f"max depth against the 2 m runs: plain rmse {ep.rmse:.4f} max {ep.max_abs:.3f}; learned rmse "
          f"{el.rmse:.4f} max {el.max_abs:.3f} ({ep.cells} cells)", flush=True)
    assert el.rmse <= 0.85 * ep.rmse
