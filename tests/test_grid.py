"""The tile grid's numbering, checked by brute force on small masks: the layout every caller indexes into
(wadi2d/grid/tiles.py:19-101): held blocks numbered from 1 in row-major order, 0 the void, neighbours in the
NEIGHBOURS order, one empty cell around the domain."""

from __future__ import annotations

import numpy as np

from wadi2d_amd.grid import NEIGHBOURS, TileGrid
from wadi2d_amd.layout import TILE


def test_numbering_neighbours_and_padding_by_brute_force():
    rng = np.random.default_rng(1)
    for _ in range(12):
        ny, nx = (int(v) for v in rng.integers(1, 130, 2))
        mask = rng.random((ny, nx)) < rng.uniform(0.005, 1.0)
        g = TileGrid.from_mask(mask, 50.0, 900.0, 5.0)
        assert g.shape[0] % TILE == 0 and g.shape[0] >= ny + 2 and g.shape[0] - TILE < ny + 2
        assert np.array_equal(g.mask[1 : 1 + ny, 1 : 1 + nx], mask) and g.mask.sum() == mask.sum()
        assert (g.x0, g.y0) == (45.0, 905.0)
        number = 0
        for bj in range(g.shape[0] // TILE):
            for bi in range(g.shape[1] // TILE):
                held = g.mask[bj * TILE : (bj + 1) * TILE, bi * TILE : (bi + 1) * TILE].any()
                if held:
                    number += 1
                    assert g.tile_of[bj, bi] == number and tuple(g.tiles[number]) == (bj, bi)
                else:
                    assert g.tile_of[bj, bi] == 0
        assert g.ntiles == number + 1 and tuple(g.tiles[0]) == (-1, -1) and not g.nbr[0].any()
        for t in range(1, g.ntiles):
            bj, bi = g.tiles[t]
            for k, (dr, dc) in enumerate(NEIGHBOURS):
                r, c = bj + dr, bi + dc
                inside = 0 <= r < g.tile_of.shape[0] and 0 <= c < g.tile_of.shape[1]
                assert g.nbr[t, k] == (g.tile_of[r, c] if inside else 0)
        j, i = (int(v) for v in np.argwhere(g.mask)[0])
        t, tj, ti = g.locate(j, i)
        assert g.tiles[t, 0] * TILE + tj == j and g.tiles[t, 1] * TILE + ti == i
        x, y = g.xy_of(j, i)
        assert g.cell_of(x, y) == (j, i)
        z = rng.random(g.shape)
        back = g.from_tiles(g.to_tiles(z, np.float64), fill=-1.0)
        held_cells = np.kron(g.tile_of > 0, np.ones((TILE, TILE), bool))
        assert np.array_equal(back[held_cells], z[held_cells]) and np.all(back[~held_cells] == -1.0)
