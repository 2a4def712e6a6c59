"""What the closure sees of a face: static statistics of the sub-grid ground around it, and the water on both sides.

For the x face between cells L = (j, i-1) and R = (j, i) (rows and columns of the padded grid; y faces likewise,
turned), with k pixels per cell side, the face's invert I = zmin[L] + fl (the table's, tables/builder.py:290-308):

* the face line's profile, max of the two pixels either side of the line, row by row (what the face table reads);
* for each of the k pixel rows, the highest pixel on the straight path from L's centre to R's centre (P_r): the
  crest a flow line must top.

Static features (11): percentiles 10/25/50/75/90 of profile - I; crest height median(P_r) - I; crest width, the
mean length of path above the crest's half height (m); the table top E; zbar[R] - zbar[L]; 1 if the face is a
natural crest (cw != 0); 1. Dynamic features (5): e_L, e_R (the water levels over the invert, not below 0), h_L,
h_R, and max(e_L, e_R) - crest height. The sill, min(P_r) - I, is where a flow line can first cross: the
geometric target of the closure's crest-height output.
"""

from __future__ import annotations

import numpy as np

from wadi2d_amd.layout import TILE

NSTATIC = 11
STATIC_NAMES = ("p10", "p25", "p50", "p75", "p90", "crest_h", "crest_w", "E", "dzbar", "natural", "one")
DYNAMIC_NAMES = ("eL", "eR", "hL", "hR", "over")


def _half_max(zb: np.ndarray, k: int):
    """Per pixel row (axis 0) and cell (axis 1): max over the first and the second half of the cell's k pixels."""
    h = k // 2
    return zb[:, :, :h].max(-1), zb[:, :, h:].max(-1)


def face_statics(tables, z: np.ndarray, k: int, axis: int) -> dict:
    """Static features, sill and validity of every x (axis 1) or y (axis 0) face, as (ny, nx) grids indexed by
    the face's R cell (the cell whose west or north face it is). ``z`` is the pixel ground on the padded grid."""
    if k % 2:
        raise ValueError("the closure's features need an even number of pixels per cell")
    grid = tables.grid
    ny, nx = grid.shape
    zz = z if axis == 1 else z.T
    nyy, nxx = (ny, nx) if axis == 1 else (nx, ny)
    zb = zz.reshape(nyy * k, nxx, k)
    first, second = _half_max(zb, k)  # (rows, cells)
    # path max per pixel row and face (face at cell c: second half of c-1, first half of c)
    P = np.full((nyy * k, nxx), np.nan)
    P[:, 1:] = np.maximum(second[:, :-1], first[:, 1:])
    prof = np.full((nyy * k, nxx), np.nan)
    prof[:, 1:] = np.maximum(zb[:, :-1, k - 1], zb[:, 1:, 0])
    P = P.reshape(nyy, k, nxx)
    prof = prof.reshape(nyy, k, nxx)
    zmin = tables.zmin if axis == 1 else tables.zmin.T
    zbar = tables.zbar if axis == 1 else tables.zbar.T
    name = "x" if axis == 1 else "y"
    fl = grid.from_tiles(getattr(tables, "fl" + name)).astype(np.float64)
    ef = grid.from_tiles(getattr(tables, "ef" + name)).astype(np.float64)
    cw = grid.from_tiles(getattr(tables, "cw" + name)).astype(np.float64)
    if axis == 0:
        fl, ef, cw = fl.T, ef.T, cw.T
    zl = np.full_like(zmin, np.nan)
    zl[:, 1:] = zmin[:, :-1]
    inv = zl + fl
    mask = grid.mask if axis == 1 else grid.mask.T
    valid = np.zeros_like(mask)
    valid[:, 1:] = mask[:, 1:] & mask[:, :-1]
    valid &= fl < 1e8  # closed faces carry the CLOSED invert
    # order statistics over the k pixel rows of each face, rows last and sorted (linear interpolation as
    # numpy's default percentile)
    rel = np.sort(np.moveaxis(prof - inv[:, None, :], 1, -1), axis=-1)
    ph_sorted = np.sort(np.moveaxis(P - inv[:, None, :], 1, -1), axis=-1)

    def pct(a, q):
        pos = q / 100.0 * (k - 1)
        lo = int(np.floor(pos))
        hi = min(lo + 1, k - 1)
        return a[..., lo] + (pos - lo) * (a[..., hi] - a[..., lo])

    pcts = [pct(rel, q) for q in (10, 25, 50, 75, 90)]
    crest_h = pct(ph_sorted, 50)
    sill = ph_sorted[..., 0]
    # crest width: path pixels above half the row's crest height, in metres
    pixel = grid.dx / k
    ph = P - inv[:, None, :]
    half = inv[:, None, :] + 0.5 * np.maximum(np.nan_to_num(ph, nan=0.0), 0.0)  # (nyy, k, nxx)
    zr = zz.reshape(nyy, k, nxx, k)
    h2 = k // 2
    count = np.zeros((nyy, k, nxx))
    count[:, :, 1:] = ((zr[:, :, :-1, h2:] > half[:, :, 1:, None]).sum(-1) +
                       (zr[:, :, 1:, :h2] > half[:, :, 1:, None]).sum(-1))  # fmt: skip
    tall = np.nan_to_num(ph, nan=0.0) > 0.05
    width = np.where(tall, count, 0.0).mean(axis=1) * pixel
    dzbar = np.zeros_like(zbar)
    dzbar[:, 1:] = zbar[:, 1:] - zbar[:, :-1]
    feats = np.stack([*pcts, crest_h, width, ef, dzbar, (cw != 0).astype(float), np.ones_like(ef)], axis=-1)
    feats = np.where(valid[..., None], np.nan_to_num(feats, nan=0.0), 0.0)
    sill = np.where(valid, np.maximum(np.nan_to_num(sill, nan=0.0), 0.0), 0.0)
    if axis == 0:
        feats, sill, valid, inv = feats.transpose(1, 0, 2), sill.T, valid.T, inv.T
    return {"static": feats, "sill": sill, "valid": valid, "invert": inv}


def dynamic(eL, eR, hL, hR, crest_h) -> np.ndarray:
    """The five dynamic features from levels over the invert and mean depths either side."""
    eL = np.maximum(eL, 0.0)
    eR = np.maximum(eR, 0.0)
    return np.stack([eL, eR, hL, hR, np.maximum(eL, eR) - crest_h], axis=-1)


def tile_index_of(grid, jj: np.ndarray, ii: np.ndarray) -> np.ndarray:
    """Own-layout (no halo) index t * 1024 + (j % 32) * 32 + (i % 32) of padded cells (j, i)."""
    t = grid.tile_of[jj // TILE, ii // TILE].astype(np.int64)
    return t * TILE * TILE + (jj % TILE) * TILE + (ii % TILE)
