"""Model: tables + backend + stepper, what tests and the run driver use (wadi2d/solver/model.py:32-185)."""

from __future__ import annotations

import numpy as np

from wadi2d_amd.config import Numerics
from wadi2d_amd.layout import HALO, STRIDE2, TILE, local_index
from wadi2d_amd.solver.backend import BoundaryCells, NumericsRecord, select_backend
from wadi2d_amd.solver.state import from_state, level_of, plan_of, to_state
from wadi2d_amd.solver.stepper import Stepper


class Model:
    def __init__(self, tables, num: Numerics | None = None, backend: str = "auto", **options) -> None:
        self.tables = tables
        self.grid = tables.grid
        self.num = num or Numerics()
        record = NumericsRecord.from_config(self.num, self.grid.dx)
        self.backend = select_backend(backend, tables, record, **options)
        self.stepper = Stepper(self.backend, self.grid, self.num)
        self.boundary_tiles = np.zeros(0, np.int64)
        self._hc = self.grid.from_tiles(tables.hc).astype(np.float64)
        self._dtab = self.grid.from_tiles(tables.dtab).astype(np.float64)

    def _tile_wet(self, h: np.ndarray) -> np.ndarray:
        g = self.grid
        ny, nx = g.shape
        blocks = (h > 0.0).reshape(ny // TILE, TILE, nx // TILE, TILE).any(axis=(1, 3))
        out = np.zeros(g.ntiles, np.int32)
        out[1:] = blocks[g.tiles[1:, 0], g.tiles[1:, 1]]
        return out

    def _tile_any(self, a: np.ndarray) -> np.ndarray:
        g = self.grid
        ny, nx = g.shape
        blocks = (a != 0.0).reshape(ny // TILE, TILE, nx // TILE, TILE).any(axis=(1, 3))
        out = np.zeros(g.ntiles, bool)
        out[1:] = blocks[g.tiles[1:, 0], g.tiles[1:, 1]]
        return out

    def set_depth(self, hbar: np.ndarray) -> None:
        """Mean depth grid, with levels through the table and zero velocities (model.py:43-54)."""
        h = np.where(self.grid.mask, hbar, 0.0)
        self.backend.set_field("h", to_state(self.grid, h))
        self.backend.set_field("d", to_state(self.grid, level_of(h, self._hc, self._dtab)))
        self.backend.set_field("om", to_state(self.grid, plan_of(h, self._hc, self._dtab)))
        self.backend.set_wet(self._tile_wet(h))
        self.stepper.rebuild_active()

    def set_velocity(self, ucell: np.ndarray, vcell: np.ndarray) -> None:
        """Face velocities from cell-centred ones; momentum and discharge follow (model.py:81-97)."""
        dx = self.grid.dx
        h = self.field("h").astype(np.float64)
        for axis, (uname, mname, qname), cell in ((1, ("u", "mx", "qx"), ucell), (0, ("v", "my", "qy"), vcell)):
            un = np.roll(cell, 1, axis=axis)
            hn = np.roll(h, 1, axis=axis)
            face = 0.5 * (cell + un)
            face = np.where((h > 0.0) & (hn > 0.0), face, 0.0)
            abar = 0.5 * (h + hn) * dx
            up = np.where(face > 0.0, hn, h) * dx
            self.backend.set_field(uname, to_state(self.grid, face))
            self.backend.set_field(mname, to_state(self.grid, abar * face))
            self.backend.set_field(qname, to_state(self.grid, up * face))

    def set_source(self, rate: np.ndarray) -> None:
        """Net source rate per cell, m/s, on a (ny, nx) grid; its tiles stay active (model.py:99-110)."""
        self.backend.set_field("src", to_state(self.grid, np.where(self.grid.mask, rate, 0.0)))
        keep = self.stepper.keep_active
        keep[:] = self._tile_any(rate)
        keep[self.boundary_tiles] = True
        self.stepper.rebuild_active(retain=True)
        self.stepper.refresh_due = True

    def set_source_cells(self, cells: np.ndarray, rate: np.ndarray) -> None:
        """``rate`` per entry of ``cells``, state indices of own cells (model.py:112-122)."""
        src = np.zeros(self.grid.ntiles * STRIDE2, np.float32)
        src[cells] = rate
        self.backend.set_field("src", src)
        keep = self.stepper.keep_active
        keep[:] = False
        keep[cells[rate != 0.0] // STRIDE2] = True
        keep[self.boundary_tiles] = True
        self.stepper.rebuild_active(retain=True)
        self.stepper.refresh_due = True

    def read_cells(self, name: str, index) -> np.ndarray:
        return self.backend.read_cells(name, index)

    def device_index(self, index: np.ndarray):
        return np.asarray(index, np.int64)

    def set_boundary(self, bc: BoundaryCells) -> None:
        self.backend.set_boundary(bc)
        self.boundary_tiles = np.unique(np.asarray(bc.cell) // STRIDE2)
        self.stepper.keep_active[self.boundary_tiles] = True
        self.stepper.rebuild_active(retain=True)

    def advance(self, t_end: float, on_step=None) -> float:
        return self.stepper.advance(t_end, on_step)

    def field(self, name: str) -> np.ndarray:
        self.backend.synchronize()
        return from_state(self.grid, self.backend.get_field(name))

    def cell_speed(self) -> np.ndarray:
        u, v = self.field("u"), self.field("v")
        return np.hypot(0.5 * (u + np.roll(u, -1, axis=1)), 0.5 * (v + np.roll(v, -1, axis=0)))

    def volume(self) -> float:
        return float(self.field("h").sum(dtype=np.float64) * self.grid.dx**2)

    def cell_state_index(self, j: int, i: int) -> int:
        t, oj, oi = self.grid.locate(j, i)
        return int(local_index(t, oj + HALO, oi + HALO))


__all__ = ["BoundaryCells", "Model"]
