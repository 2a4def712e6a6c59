"""The tile grid: which 32 x 32 blocks of a raster-aligned grid hold cells, how those blocks are numbered, and which
block borders which.

The numbering is part of the state layout every caller indexes into, so it is the reference's
(wadi2d/grid/tiles.py:19-101): blocks that hold at least one cell are numbered 1, 2, ... in row-major order, number 0
is the void block that stands for anything off the grid or empty, a block's eight neighbours are listed in the order
of ``NEIGHBOURS``, and ``from_mask`` leaves one empty cell around the domain before rounding up to whole blocks.
Rows count south from the north edge ``y0``, columns east from the west edge ``x0``.
"""

from __future__ import annotations

import numpy as np

from wadi2d_amd.layout import TILE

NEIGHBOURS = ((-1, -1), (-1, 0), (-1, 1), (0, -1), (0, 1), (1, -1), (1, 0), (1, 1))
"""Block offsets (rows, columns) of the neighbour columns of ``TileGrid.nbr``: north-west, north, north-east, west,
east, south-west, south, south-east."""


def _held(mask: np.ndarray) -> np.ndarray:
    """Per block of the grid, whether any of its cells is in the domain."""
    nby, nbx = mask.shape[0] // TILE, mask.shape[1] // TILE
    return mask.reshape(nby, TILE, nbx, TILE).any(axis=(1, 3))


class TileGrid:
    """Geometry and block bookkeeping of one grid.

    ``mask`` (ny, nx; both multiples of TILE) is True on the domain's cells. Built from it: ``tile_of`` (block rows,
    block columns) holding each block's number or 0; ``tiles`` (ntiles, 2) holding each number's block row and column,
    (-1, -1) for the void; ``nbr`` (ntiles, 8) holding each block's neighbours' numbers, all 0 for the void.
    """

    def __init__(self, x0: float, y0: float, dx: float, mask: np.ndarray) -> None:
        mask = np.asarray(mask, dtype=bool)
        if mask.ndim != 2 or mask.shape[0] % TILE or mask.shape[1] % TILE:
            raise ValueError("mask shape must be a multiple of the tile size; use TileGrid.from_mask")
        self.x0, self.y0, self.dx, self.mask = float(x0), float(y0), float(dx), mask
        held = _held(mask)
        # a running count over the blocks in row-major order numbers exactly the held ones 1, 2, ...
        count = np.cumsum(held.ravel(), dtype=np.int64).reshape(held.shape)
        self.tile_of = np.where(held, count, 0).astype(np.int32)
        where = np.argwhere(held).astype(np.int32)
        self.tiles = np.concatenate([np.full((1, 2), -1, np.int32), where])
        # a neighbour off the grid reads the frame of zeros around the numbering: the void
        framed = np.pad(self.tile_of, 1)
        self.nbr = np.zeros((self.tiles.shape[0], len(NEIGHBOURS)), np.int32)
        rows, cols = where[:, 0] + 1, where[:, 1] + 1
        for k, (dr, dc) in enumerate(NEIGHBOURS):
            self.nbr[1:, k] = framed[rows + dr, cols + dc]

    @classmethod
    def from_mask(cls, mask: np.ndarray, x0: float, y0: float, dx: float) -> TileGrid:
        """The grid for a domain whose north-west corner is (x0, y0): one empty cell on every side, then whole blocks
        (tiles.py:52-64); the grid's own corner moves one cell out."""
        mask = np.asarray(mask, dtype=bool)
        need = [-(-(n + 2) // TILE) * TILE - (n + 1) for n in mask.shape]  # cells after the domain on each axis
        grown = np.pad(mask, ((1, need[0]), (1, need[1])))
        return cls(x0 - dx, y0 + dx, dx, grown)

    @property
    def ntiles(self) -> int:
        return int(self.tiles.shape[0])

    @property
    def shape(self) -> tuple[int, int]:
        return self.mask.shape

    def cell_of(self, x: float, y: float) -> tuple[int, int]:
        """(row, column) of the cell holding point (x, y)."""
        return int((self.y0 - y) // self.dx), int((x - self.x0) // self.dx)

    def xy_of(self, j, i):
        """Centre of cell (row j, column i); arrays broadcast."""
        return self.x0 + (i + 0.5) * self.dx, self.y0 - (j + 0.5) * self.dx

    def locate(self, j: int, i: int) -> tuple[int, int, int]:
        """(block number, row, column within the block) of cell (j, i)."""
        return int(self.tile_of[j // TILE, i // TILE]), j % TILE, i % TILE

    def _block_view(self, a: np.ndarray) -> np.ndarray:
        """A (ny, nx, ...) grid seen as (block rows, block columns, TILE, TILE, ...)."""
        ny, nx = self.shape
        return a.reshape((ny // TILE, TILE, nx // TILE, TILE) + a.shape[2:]).swapaxes(1, 2)

    def to_tiles(self, a: np.ndarray, dtype=np.float32) -> np.ndarray:
        """The held blocks of a (ny, nx, ...) grid as (ntiles, TILE, TILE, ...) by number; the void block is zero."""
        out = np.zeros((self.ntiles, TILE, TILE) + a.shape[2:], dtype=dtype)
        out[1:] = self._block_view(a)[self.tiles[1:, 0], self.tiles[1:, 1]]
        return out

    def from_tiles(self, a: np.ndarray, fill: float = 0.0) -> np.ndarray:
        """A (ny, nx, ...) grid from (ntiles, TILE, TILE, ...) blocks by number; blocks held by no tile take ``fill``."""
        ny, nx = self.shape
        out = np.full((ny, nx) + a.shape[3:], fill, dtype=a.dtype)
        self._block_view(out)[self.tiles[1:, 0], self.tiles[1:, 1]] = a[1:]
        return out
