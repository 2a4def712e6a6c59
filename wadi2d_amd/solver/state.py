"""Host-side conversions between (ny, nx) grids and the state layout, and the FP64 level/plan of a start
(wadi2d/solver/state.py:55-301)."""

from __future__ import annotations

import numpy as np

from wadi2d_amd.layout import HALO, STRIDE, STRIDE2, TILE, local_index


def to_state(grid, a: np.ndarray, dtype=np.float64) -> np.ndarray:
    """A (ny, nx) grid into own cells and halos; halos by clipped grid reads (state.py:198-209)."""
    ny, nx = grid.shape
    nt = grid.ntiles
    out = np.zeros((nt, STRIDE, STRIDE), dtype)
    if nt > 1:
        tj = grid.tiles[1:, 0].astype(np.int64)
        ti = grid.tiles[1:, 1].astype(np.int64)
        off = np.arange(STRIDE) - HALO
        r = np.clip(tj[:, None] * TILE + off[None, :], 0, ny - 1)
        c = np.clip(ti[:, None] * TILE + off[None, :], 0, nx - 1)
        out[1:] = a[r[:, :, None], c[:, None, :]]
    return out.reshape(-1)


def from_state(grid, a: np.ndarray) -> np.ndarray:
    """Own cells of a state array as a (ny, nx) grid (state.py:241-252)."""
    own = np.asarray(a)[: grid.ntiles * STRIDE2].reshape(-1, STRIDE, STRIDE)[:, HALO : HALO + TILE, HALO : HALO + TILE]
    return grid.from_tiles(np.ascontiguousarray(own))


def state_index(grid, j, i):
    """State-array index of grid cells (state.py:255-260)."""
    j = np.asarray(j)
    i = np.asarray(i)
    tile = grid.tile_of[j // TILE, i // TILE]
    return local_index(tile.astype(np.int64), j % TILE + HALO, i % TILE + HALO)


def level_of(h: np.ndarray, hc: np.ndarray, dtab: np.ndarray) -> np.ndarray:
    """Level above zmin for mean depths, float64 (state.py:269-282)."""
    last = dtab.shape[-1] - 1
    out = np.array(h, dtype=np.float64)
    wet = hc > 0.0
    above = wet & (h >= hc)
    out[above] = dtab[above, last] + (h[above] - hc[above])
    mid = wet & ~above
    s = np.sqrt(h[mid] / hc[mid]) * last
    r = np.minimum(s.astype(int), last - 1)
    idx = np.arange(r.size)
    t = dtab[mid]
    out[mid] = t[idx, r] + (s - r) * (t[idx, r + 1] - t[idx, r])
    return out


def plan_of(h: np.ndarray, hc: np.ndarray, dtab: np.ndarray) -> np.ndarray:
    """Wet share of the cell plan, float64 (state.py:291-301)."""
    last = dtab.shape[-1] - 1
    out = np.ones(h.shape, np.float64)
    mid = (hc > 0.0) & (h < hc)
    s = np.sqrt(h[mid] / hc[mid]) * last
    r = np.minimum(s.astype(int), last - 1)
    idx = np.arange(r.size)
    t = dtab[mid]
    out[mid] = hc[mid] * (2 * r + 1) / last**2 / np.maximum(t[idx, r + 1] - t[idx, r], 1e-12)
    return out
