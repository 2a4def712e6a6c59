"""Small seeded cases shared by the tests: beds, tables and starts, at sizes the oracle finishes in seconds."""

from __future__ import annotations

import numpy as np
from scipy.ndimage import gaussian_filter

from wadi2d_amd.config import Tables
from wadi2d_amd.grid import TileGrid
from wadi2d_amd.tables import ArrayBed, build_tables


def rough_dam_break(seed: int = 3, nx: int = 90, ny: int = 60, k: int = 6, dx: float = 10.0):
    """The reference's CPU/GPU parity case (tests/solver/test_subgrid.py:95-116): a rough bed, 2 m of water on the
    west third. Returns (tables, start depth)."""
    rng = np.random.default_rng(seed)
    tg = TileGrid.from_mask(np.ones((ny, nx), bool), 0.0, 0.0, dx)
    z = gaussian_filter(rng.normal(0, 1.0, (tg.shape[0] * k + 2, tg.shape[1] * k + 2)), 5) * 30 + 200.0
    tables = build_tables(tg, ArrayBed(z, dx / k), Tables(), 0.03)
    h0 = np.zeros(tg.shape)
    h0[:, :30] = 2.0
    return tables, h0


def sloped_valley(nx: int = 220, ny: int = 70, k: int = 4, dx: float = 10.0, seed: int = 5, rows: int = 8):
    """A valley falling east with noise and a crest wall with a gap across it: crest faces, sheet flow, wet/dry."""
    rng = np.random.default_rng(seed)
    tg = TileGrid.from_mask(np.ones((ny, nx), bool), 0.0, 0.0, dx)
    nyp, nxp = tg.shape
    pixel = dx / k
    y, x = (np.mgrid[0 : nyp * k, 0 : nxp * k] + 0.5) * pixel
    z = 30.0 - 0.01 * x + 0.02 * np.abs(y - 0.5 * nyp * dx)
    z += gaussian_filter(rng.normal(0, 1.0, z.shape), 3) * 1.5
    wall = (x >= 0.55 * nxp * dx) & (x < 0.55 * nxp * dx + 2 * pixel)
    gap = np.abs(y - 0.5 * nyp * dx) < 10.0
    z[wall & ~gap] += 1.5
    tables = build_tables(tg, ArrayBed(z, pixel), Tables(rows=rows), 0.04)
    return tables


def deep_and_shallow(n: int = 200, k: int = 4, dx: float = 10.0, seed: int = 11, deep: float = 4.0,
                     shallow: float = 0.05):  # fmt: skip
    """A rough plain, deep water (``deep`` m) on its west fifth and a thin sheet (``shallow`` m) everywhere else: wave
    speeds from the floor to about 7 m/s across the grid, so the level stepper holds tiles at several time levels.
    Returns (tables, start depth)."""
    rng = np.random.default_rng(seed)
    tg = TileGrid.from_mask(np.ones((n, n), bool), 0.0, 0.0, dx)
    z = gaussian_filter(rng.normal(0, 1.0, (tg.shape[0] * k + 2, tg.shape[1] * k + 2)), 4) * 2.0 + 100.0
    tables = build_tables(tg, ArrayBed(z, dx / k), Tables(), 0.035)
    h0 = np.full(tg.shape, shallow)
    h0[:, : n // 5] = deep
    return tables, h0
