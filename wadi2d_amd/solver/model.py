"""Model: one grid's tables, its HIP backend and the stepper that drives it; what the run driver and the tests use.

It answers the calls the reference's callers make on their Model (wadi2d/solver/model.py:32-185): a start from mean
depths, face velocities from cell velocities, the source by grid or by cells, boundary cells, advancing to a time,
fields as (ny, nx) grids, and cell indices into the state layout. Every array it hands the backend is in that layout
(solver/state.py); the tiles it keeps awake for a source or a boundary go to the stepper.
"""

from __future__ import annotations

import numpy as np

from wadi2d_amd.config import Numerics
from wadi2d_amd.layout import HALO, STRIDE2, TILE, local_index
from wadi2d_amd.solver.backend import BoundaryCells, NumericsRecord, select_backend
from wadi2d_amd.solver.state import from_state, level_of, plan_of, to_state
from wadi2d_amd.solver.stepper import Stepper


def tiles_touched(grid, flags: np.ndarray) -> np.ndarray:
    """Per tile number, whether any of its cells is flagged in the (ny, nx) boolean grid; the void tile never is."""
    ny, nx = grid.shape
    per_block = flags.reshape(ny // TILE, TILE, nx // TILE, TILE).any(axis=(1, 3))
    out = np.zeros(grid.ntiles, bool)
    out[1:] = per_block[grid.tiles[1:, 0], grid.tiles[1:, 1]]
    return out


def _from_upstream(a: np.ndarray, axis: int) -> np.ndarray:
    """The value of the cell west (axis 1) or north (axis 0) of each cell; the first column or row reads 0 (it is the
    padding ring, always dry)."""
    out = np.zeros_like(a)
    if axis == 1:
        out[:, 1:] = a[:, :-1]
    else:
        out[1:, :] = a[:-1, :]
    return out


def make_stepper(backend, grid, num: Numerics) -> Stepper:
    """The uniform step, or the level stepper when the numerics ask for local time stepping (model.py:23-29)."""
    if not num.local_time_stepping:
        return Stepper(backend, grid, num)
    from wadi2d_amd.solver.level_stepper import LevelStepper

    return LevelStepper(backend, grid, num)


class Model:
    def __init__(self, tables, num: Numerics | None = None, backend: str = "auto", **options) -> None:
        self.tables = tables
        self.grid = tables.grid
        self.num = num or Numerics()
        self.backend = select_backend(backend, tables, NumericsRecord.from_config(self.num, self.grid.dx), **options)
        self.stepper = make_stepper(self.backend, self.grid, self.num)
        self.boundary_tiles = np.zeros(0, np.int64)
        # the level and wet-share tables as grids, for starts given as mean depths
        self._hc = self.grid.from_tiles(tables.hc).astype(np.float64)
        self._dtab = self.grid.from_tiles(tables.dtab).astype(np.float64)

    def _put(self, name: str, grid_values: np.ndarray) -> None:
        self.backend.set_field(name, to_state(self.grid, grid_values))

    def _keep(self, tiles: np.ndarray) -> None:
        """Tiles that stay awake for a source, with the boundary's; the list is widened, never narrowed, and the
        next step refreshes the wave speed."""
        keep = self.stepper.keep_active
        keep[:] = tiles
        keep[self.boundary_tiles] = True
        self.stepper.rebuild_active(retain=True)
        self.stepper.refresh_due = True

    def set_depth(self, hbar: np.ndarray) -> None:
        """Start from mean depths (off-domain cells dry): levels and wet shares through each cell's table, still water,
        the wet tiles flagged and the active list made afresh (model.py:43-54)."""
        h = np.where(self.grid.mask, hbar, 0.0)
        self._put("h", h)
        self._put("d", level_of(h, self._hc, self._dtab))
        self._put("om", plan_of(h, self._hc, self._dtab))
        self.backend.set_wet(tiles_touched(self.grid, h > 0.0).astype(np.int32))
        self.stepper.rebuild_active()

    def set_velocity(self, ucell: np.ndarray, vcell: np.ndarray) -> None:
        """Face velocities as the mean of the two cells' velocities, on faces wet on both sides; the momentum takes
        the two cells' mean depth, the discharge the upwind cell's (model.py:81-97)."""
        dx = self.grid.dx
        h = self.field("h").astype(np.float64)
        for axis, names, cell in ((1, ("u", "mx", "qx"), ucell), (0, ("v", "my", "qy"), vcell)):
            h_up = _from_upstream(h, axis)
            face = 0.5 * (cell + _from_upstream(cell, axis))
            face = np.where((h > 0.0) & (h_up > 0.0), face, 0.0)
            for name, values in zip(names, (face, 0.5 * (h + h_up) * dx * face,
                                            np.where(face > 0.0, h_up, h) * dx * face)):  # fmt: skip
                self._put(name, values)

    def set_source(self, rate: np.ndarray) -> None:
        """A net source (m/s) on a (ny, nx) grid; tiles with a non-zero rate stay awake (model.py:99-110)."""
        self._put("src", np.where(self.grid.mask, rate, 0.0))
        self._keep(tiles_touched(self.grid, rate != 0.0))

    def set_source_cells(self, cells: np.ndarray, rate: np.ndarray) -> None:
        """A net source given per own cell (state indices), every other cell none (model.py:112-122)."""
        src = np.zeros(self.grid.ntiles * STRIDE2, np.float32)
        src[cells] = rate
        self.backend.set_field("src", src)
        tiles = np.zeros(self.grid.ntiles, bool)
        tiles[np.asarray(cells)[np.asarray(rate) != 0.0] // STRIDE2] = True
        self._keep(tiles)

    def set_boundary(self, bc: BoundaryCells) -> None:
        """Prescribed cells (model.py:145-150); their tiles stay awake."""
        self.backend.set_boundary(bc)
        self.boundary_tiles = np.unique(np.asarray(bc.cell) // STRIDE2)
        self.stepper.keep_active[self.boundary_tiles] = True
        self.stepper.pin(self.boundary_tiles)
        self.stepper.rebuild_active(retain=True)

    def read_cells(self, name: str, index) -> np.ndarray:
        """A field's values at own-cell state indices, gathered on the device."""
        return self.backend.read_cells(name, index)

    def device_index(self, index: np.ndarray):
        """Indices in the form ``read_cells`` takes (host int64 here: the gather copies them over)."""
        return np.asarray(index, np.int64)

    def advance(self, t_end: float, on_step=None) -> float:
        return self.stepper.advance(t_end, on_step)

    def field(self, name: str) -> np.ndarray:
        """A state field as a (ny, nx) grid."""
        self.backend.synchronize()
        return from_state(self.grid, self.backend.get_field(name))

    def cell_speed(self) -> np.ndarray:
        """Speed at cell centres from the mean of each axis's two face velocities (west and east, north and south)."""
        u, v = self.field("u"), self.field("v")
        east, south = np.roll(u, -1, axis=1), np.roll(v, -1, axis=0)
        return np.hypot(0.5 * (u + east), 0.5 * (v + south))

    def volume(self) -> float:
        """Water held, m3: the mean depths' sum times the cell area."""
        return float(self.field("h").sum(dtype=np.float64) * self.grid.dx**2)

    def cell_state_index(self, j: int, i: int) -> int:
        """State index of grid cell (j, i)."""
        t, tj, ti = self.grid.locate(j, i)
        return int(local_index(t, tj + HALO, ti + HALO))


__all__ = ["BoundaryCells", "Model", "make_stepper", "tiles_touched"]
