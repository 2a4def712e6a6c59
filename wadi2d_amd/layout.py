"""The array contract of the drop-in boundary (wadi2d/solver/layout.py:1-52).

Callers see state arrays in the reference's layout: compact tiles of ``TILE`` x ``TILE`` own cells inside a
``HALO`` ring, ``k = t * STRIDE2 + j * STRIDE + i`` counted from the halo corner, tile 0 the void tile. Static
tables have no halo: ``t * TILE2 + (j - HALO) * TILE + (i - HALO)``.

On the device the HIP library keeps its own layout: the same tiles without halo entries (own cells only,
``t * TILE2 + j * TILE + i``); a kernel reads a neighbour tile's edge cells straight from that tile. The
library converts at the boundary (``set_field``/``get_field``), so nothing above it sees the difference.
"""

from __future__ import annotations

TILE = 32
HALO = 2
STRIDE = TILE + 2 * HALO
STRIDE2 = STRIDE * STRIDE
TILE2 = TILE * TILE

STATE_ARRAYS = ("h", "d", "u", "v", "mx", "my", "qx", "qy")
"""Mean depth, level above zmin, face velocities, face momenta, face discharges."""

WORK_ARRAYS = ("u1", "v1", "mx1", "my1", "qx1", "qy1", "r", "src", "res", "om")
"""Next face state, positivity ratio, net source rate (m/s), depth residual, wet share of the cell plan."""

CLOSED = 1.0e9
"""Invert over zmin of a closed face; no level reaches it, so a wall needs no logic in a kernel."""


def local_index(t, j, i):
    """Index of cell (j, i), halo-based, in tile t of a state array."""
    return t * STRIDE2 + j * STRIDE + i


def table_index(t, j, i):
    """Index of own cell (j, i), halo-based, in tile t of a static array."""
    return t * TILE2 + (j - HALO) * TILE + (i - HALO)
