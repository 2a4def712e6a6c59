"""The allocation-free gather, w2d_read_cells_noalloc, against w2d_read_cells through the C-ABI.

It reads through one fixed device array per device, shared by every call: two calls in a row (the second with
fewer cells) must each give what read_cells gives, bit for bit, and a refused call must leave the next one right.
"""

from __future__ import annotations

import numpy as np
import pytest

from tests.cases import rough_dam_break
from wadi2d_amd.config import Numerics
from wadi2d_amd.solver.model import Model

MAX_CELLS = 262_144


def _own_indices(m, rng, n):
    ny, nx = m.tables.grid.shape
    j = rng.integers(0, ny, n)
    i = rng.integers(0, nx, n)
    return np.array([m.cell_state_index(int(a), int(b)) for a, b in zip(j, i)], np.int64)


@pytest.mark.gpu
@pytest.mark.parametrize("backend", ["hip-exact", "hip-f64"])
def test_noalloc_gather_equals_read_cells_on_two_calls_in_a_row(backend):
    tables, h0 = rough_dam_break(nx=90, ny=60)
    m = Model(tables, Numerics(dt_every=1, dt_max=0.2), backend)
    m.set_depth(h0)
    m.advance(3.0)
    b = m.backend
    rng = np.random.default_rng(11)
    first = _own_indices(m, rng, 4000)
    second = first[::7][:300].copy()  # fewer cells, some of them the first call's
    for field in ("h", "d", "u", "qx", "mx", "res"):
        for idx in (first, second):
            want = b.read_cells(field, idx)
            got = b.read_cells_noalloc(field, idx)
            assert got.dtype == want.dtype and got.shape == want.shape
            assert np.array_equal(got, want), (backend, field, idx.size)
    # a step between two calls: the scratch keeps the last call's contents, which must not leak into the next
    m.advance(0.5)
    for idx in (first, second):
        assert np.array_equal(b.read_cells_noalloc("h", idx), b.read_cells("h", idx))


@pytest.mark.gpu
def test_noalloc_gather_refuses_what_it_cannot_hold_and_recovers():
    tables, h0 = rough_dam_break(nx=60, ny=40)
    m = Model(tables, Numerics(dt_every=1, dt_max=0.2), "hip-exact")
    m.set_depth(h0)
    b = m.backend
    idx = _own_indices(m, np.random.default_rng(2), 500)
    with pytest.raises(ValueError, match="262,144"):
        b.read_cells_noalloc("h", np.repeat(idx[:1], MAX_CELLS + 1))
    halo = np.array([0, 1, 36], np.int64)  # tile 0's halo corner entries: not own cells
    with pytest.raises(ValueError, match="not an own cell"):
        b.read_cells_noalloc("h", np.concatenate([idx[:5], halo]))
    assert np.array_equal(b.read_cells_noalloc("h", idx), b.read_cells("h", idx))
    full = np.repeat(idx, MAX_CELLS // idx.size + 1)[:MAX_CELLS]
    assert np.array_equal(b.read_cells_noalloc("d", full), b.read_cells("d", full))
    assert b.read_cells_noalloc("h", idx[:0]).size == 0
