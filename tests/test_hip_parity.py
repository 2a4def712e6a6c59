"""The HIP step against the oracle on the same seeded inputs, through the C-ABI.

* float32 exact: IEEE division and square root, the cube root rounded once from double, no contraction. This is
  the float32 statement order of the reference's CPU twin (solver/cpu/kernels.py), which the oracle's float
  instantiation follows, so the fields are compared bit for bit.
* float64: the same arithmetic in double; the device's cube root and the host's may differ in the last bit, so a
  tolerance of 1e-9 (relative to the field's scale) stands in for bit equality.
* float32 fast: the hardware reciprocal, square root and a log2/exp2 cube root; held to the reference's own
  CPU/GPU parity limits (tests/solver/test_subgrid.py:95-116: |dh| < 1e-3 m, |dqx| < 0.05 m3/s).
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle.oracle import OracleModel
from oracle.oracle import Numerics as ONumerics
from tests.cases import rough_dam_break, sloped_valley
from wadi2d_amd.config import Numerics
from wadi2d_amd.solver.backend import BoundaryCells
from wadi2d_amd.solver.model import Model

FIELDS = ("h", "d", "u", "v", "mx", "my", "qx", "qy", "res", "om")


def _pair(tables, num, backend, prec, h0=None, bc=None, rain=None):
    hip = Model(tables, num, backend)
    ora = OracleModel(tables, ONumerics.like(num), prec=prec, threads=4)
    for m in (hip, ora):
        if h0 is not None:
            m.set_depth(h0)
        if bc is not None:
            m.set_boundary(bc)
        if rain is not None:
            m.set_source(rain)
    return hip, ora


def _east_free(m, tables):
    """Kind-1 (free) boundary cells on the last interior column, like the reference's harness."""
    ny, nx = tables.grid.shape
    cells = [m.cell_state_index(j, nx - 2) for j in range(1, ny - 1)]
    n = len(cells)
    return BoundaryCells(np.array(cells, np.int32), np.array(cells, np.int32), np.ones(n, np.int32),
                         np.zeros(n, np.float32), np.zeros(n, np.float32))  # fmt: skip


def _compare(hip, ora, atol_rel: float, label: str):
    worst = {}
    depth_scale = max(float(np.abs(ora.field("h")).max()), 1e-30)
    for f in FIELDS:
        a = hip.field(f).astype(np.float64)
        b = ora.field(f).astype(np.float64)
        # the residual is the rounding left over from a depth update: it is measured against the depth
        scale = depth_scale if f == "res" else max(float(np.abs(b).max()), 1e-30)
        err = float(np.abs(a - b).max()) / scale
        worst[f] = err
    print(label, {k: f"{v:.2e}" for k, v in worst.items()}, flush=True)
    bad = {k: v for k, v in worst.items() if v > atol_rel}
    assert not bad, f"{label}: fields differ beyond {atol_rel}: {bad}"


@pytest.mark.gpu
def test_exact_float32_is_bitwise_the_oracle_on_a_rough_dam_break():
    tables, h0 = rough_dam_break()
    num = Numerics(dt_every=1, dt_max=0.2)
    hip, ora = _pair(tables, num, "hip-exact", 4, h0=h0)
    for t_end in (0.2, 5.0, 30.0):
        hip.advance(t_end)
        ora.advance(t_end)
        assert hip.stepper.nsteps == ora.stepper.nsteps
        _compare(hip, ora, 0.0, f"exact f32 t={t_end}")


@pytest.mark.gpu
def test_float64_matches_the_fp64_oracle_on_a_rough_dam_break():
    tables, h0 = rough_dam_break()
    num = Numerics(dt_every=1, dt_max=0.2)
    hip, ora = _pair(tables, num, "hip-f64", 8, h0=h0)
    hip.advance(30.0)
    ora.advance(30.0)
    assert hip.stepper.nsteps == ora.stepper.nsteps
    _compare(hip, ora, 1e-9, "f64 t=30")


@pytest.mark.gpu
def test_fast_float32_within_the_reference_parity_limits():
    tables, h0 = rough_dam_break()
    num = Numerics(dt_every=1, dt_max=0.2)
    hip, ora = _pair(tables, num, "hip", 4, h0=h0)
    hip.advance(60.0)
    ora.advance(60.0)
    assert hip.stepper.nsteps == ora.stepper.nsteps
    assert np.abs(hip.field("h") - ora.field("h")).max() < 1e-3
    assert np.abs(hip.field("qx") - ora.field("qx")).max() < 0.05


@pytest.mark.gpu
@pytest.mark.parametrize("backend, prec, tol", [("hip-exact", 4, 0.0), ("hip-f64", 8, 1e-9)])
def test_rain_crest_and_free_outflow_with_adaptive_stepping(backend, prec, tol):
    """Rain on a crested valley draining east: wet/dry fronts, natural crests, sheet flow, sleeping tiles, the
    adaptive step and a free boundary, over enough steps that tiles wake."""
    tables = sloped_valley()
    num = Numerics()
    rain = np.where(tables.grid.mask, 80.0 / 3.6e6, 0.0)
    hip, ora = _pair(tables, num, backend, prec)
    for m in (hip, ora):
        m.set_boundary(_east_free(m, tables))
    rain[:, tables.grid.shape[1] // 3 :] = 0.0  # rain on the west third only, so the east tiles start asleep
    for m in (hip, ora):
        m.set_source(rain)
    for t_end in (60.0, 300.0, 900.0):
        hip.advance(t_end)
        ora.advance(t_end)
        assert hip.stepper.nsteps == ora.stepper.nsteps
        assert np.array_equal(np.sort(hip.backend.active_tiles), np.sort(ora.backend.active_tiles))
        _compare(hip, ora, tol, f"{backend} valley t={t_end}")
    assert hip.stepper.nsteps > 50


@pytest.mark.gpu
@pytest.mark.parametrize("faces_threads, cells_threads", [(256, 256), (256, 512), (512, 512)])
def test_every_launch_shape_is_bitwise_the_oracle(faces_threads, cells_threads):
    """The workgroup sizes change which thread stages and computes which cell, never the arithmetic."""
    tables, h0 = rough_dam_break()
    num = Numerics(dt_every=1, dt_max=0.2)
    hip, ora = _pair(tables, num, "hip-exact", 4, h0=h0)
    hip.backend.set_launch(faces_threads, cells_threads)
    hip.advance(5.0)
    ora.advance(5.0)
    assert hip.stepper.nsteps == ora.stepper.nsteps
    _compare(hip, ora, 0.0, f"launch {faces_threads}/{cells_threads}")


@pytest.mark.gpu
def test_a_momentum_or_depth_set_between_steps_is_taken_as_given():
    """The device keeps no face momentum between steps (the face kernel recomputes it from the velocity and the
    depths the last step read): reading it, then setting it (a kick, as FaceKick.apply writes b.mx[faces],
    momentum.py:49-54) and setting the depth between steps must leave every field bit for bit the oracle's, which
    stores the momentum as the reference does."""
    tables, h0 = rough_dam_break()
    num = Numerics(dt_every=1, dt_max=0.2)
    hip, ora = _pair(tables, num, "hip-exact", 4, h0=h0)
    for m in (hip, ora):
        m.advance(5.0)
    _compare(hip, ora, 0.0, "exact f32 t=5 (momentum stored on read)")
    from wadi2d_amd.solver.state import to_state

    for m in (hip, ora):
        mx = m.field("mx").astype(np.float32)
        mx[:, ::7] *= np.float32(1.5)  # a kick on every seventh column of faces
        m.backend.set_field("mx", to_state(m.grid, mx, np.float32))
    for m in (hip, ora):
        m.advance(8.0)
    _compare(hip, ora, 0.0, "exact f32 t=8 after a kick")
    for m in (hip, ora):
        h = m.field("h").astype(np.float32)
        h[::11, :] += np.where(m.grid.mask[::11, :], np.float32(0.25), np.float32(0.0))
        m.backend.set_field("h", to_state(m.grid, h, np.float32))
        m.advance(12.0)
    _compare(hip, ora, 0.0, "exact f32 t=12 after a depth change")


@pytest.mark.gpu
def test_cells_written_between_steps_are_what_set_field_would_give():
    """w2d_write_cells is the reference callers' writes through the backend's arrays (FaceKick.apply's
    b.mx[faces], momentum.py:49-54; tests' b.u[idx]): a kick on some faces, a velocity and a depth written cell by
    cell, then steps, must leave every field bit for bit the oracle's after the same whole-array set_field."""
    from wadi2d_amd.solver.state import state_index, to_state

    tables, h0 = rough_dam_break()
    num = Numerics(dt_every=1, dt_max=0.2)
    hip, ora = _pair(tables, num, "hip-exact", 4, h0=h0)
    for m in (hip, ora):
        m.advance(4.0)
    rng = np.random.default_rng(21)
    jj, ii = np.nonzero(tables.grid.mask)
    pick = rng.choice(jj.size, 400, replace=False)
    j, i = jj[pick], ii[pick]
    idx = np.asarray(state_index(tables.grid, j, i), np.int64)
    for name, scale in (("mx", 1.5), ("u", 0.5), ("h", 1.25)):
        whole = ora.field(name).astype(np.float32)
        whole[j, i] *= np.float32(scale)
        ora.backend.set_field(name, to_state(ora.grid, whole, np.float32))
        values = hip.field(name).astype(np.float32)[j, i] * np.float32(scale)
        # the first write of each cell is overwritten by the second: the last one counts
        hip.backend.write_cells(name, np.concatenate((idx, idx)), np.concatenate((values * 0, values)))
        assert np.array_equal(hip.field(name), whole), name
    for m in (hip, ora):
        m.advance(9.0)
    _compare(hip, ora, 0.0, "exact f32 t=9 after cells written")
