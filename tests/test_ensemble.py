"""The storm/ARF ensemble (wadi2d_amd/engine/ensemble.py, storms.py): the design storm against the published NRCS
Type II ordinates and the HMS Hypothetical Storm's rules (storm/hypothetical.py:43-69, storm/patterns.py:33-44);
the mosaic against its members; on the GPU, the batched run against each member run alone the plain way (bit for
bit in the exact variant) and against the FP64 oracle."""

from __future__ import annotations

import numpy as np
import pytest

from tests.runs import run_oracle
from wadi2d_amd.engine.ensemble import (Ensemble, MemberSpec, draw_members, max_depth_error, run_batched,
                                        run_serial, stack_tables)  # fmt: skip
from wadi2d_amd.engine.storms import TYPE_II, ScsStorm, scs_curve
from wadi2d_amd.layout import STRIDE2


def test_type_ii_curve_holds_the_published_ordinates():
    # the NRCS tabular Type II distribution (TR-55), as published
    minutes, frac = scs_curve()
    for hours, want in ((2, 0.022), (6, 0.080), (8, 0.120), (10, 0.181), (11, 0.235), (11.5, 0.283), (12, 0.663),
                        (12.5, 0.735), (13, 0.772), (14, 0.820), (16, 0.880), (20, 0.952), (24, 1.0)):  # fmt: skip
        assert frac[int(hours * 60)] == pytest.approx(want, abs=1e-12)
    assert len(TYPE_II) == 241 and minutes[-1] == 1440.0
    assert np.all(np.diff(frac) >= -1e-4)  # a natural spline through rising knots barely dips between them
    knots = np.asarray(TYPE_II)
    assert np.array_equal(frac[::6], knots)


def test_scs_storm_stretches_the_day_over_its_duration_and_takes_the_factor():
    s = ScsStorm(80.0, 4.0, 0.75)
    assert s(0.0) == 0.0
    assert s(4 * 3600.0) == pytest.approx(60.0)
    assert s(10 * 3600.0) == pytest.approx(60.0)
    # half the duration is the 24-hour curve at 12 h
    assert s(2 * 3600.0) == pytest.approx(60.0 * 0.663)
    # the window mean rate is the depth difference over the window
    assert s.mean_rate(600.0, 630.0) == pytest.approx((s(630.0) - s(600.0)) / 1000.0 / 30.0)
    with pytest.raises(ValueError):
        ScsStorm(10.0, 1.0, 1.5)


def test_draws_cover_the_stated_ranges():
    specs = draw_members(64)
    assert {s.side for s in specs} <= {128, 256, 512, 768, 1024}
    assert all(20.0 <= s.depth_mm <= 120.0 and 1.0 <= s.duration_h <= 6.0 and 0.6 <= s.arf <= 1.0 for s in specs)
    assert len({s.seed for s in specs}) == 64
    assert draw_members(64) == specs  # seeded


def _small(duration_s=1200.0):
    specs = [MemberSpec(0, 11, 96, 90.0, 1.0, 0.9), MemberSpec(1, 12, 64, 60.0, 0.5, 1.0),
             MemberSpec(2, 13, 128, 120.0, 1.0, 0.7)]  # fmt: skip
    return Ensemble.build(specs, dx=20.0, k=4, duration_s=duration_s)


def test_mosaic_keeps_every_member_whole_and_apart():
    ens = _small()
    g = ens.tables.grid
    member = ens.tile_member()
    for m, sl in enumerate(ens.slots):
        t = ens.scenarios[m].tables
        for name in ("hc", "dtab", "flx", "cwy", "ftabx", "zcy"):
            assert np.array_equal(getattr(ens.tables, name)[sl.tile0 : sl.tile1], getattr(t, name)[1:])
        nb = g.nbr[sl.tile0 : sl.tile1]
        # neighbours are the member's own tiles (shifted) or the void tile, never another member's
        assert np.all((nb == 0) | (member[nb] == m))
        own = t.grid.nbr[1:]
        assert np.array_equal(np.where(own == 0, 0, own + sl.tile0 - 1), nb)
        assert np.array_equal(ens.member_field(m, ens.tables.zmin), t.zmin)
        # interior cells: the member's own cells but its free east column, in the member's tiles
        assert ens.interior[m].size == int(t.grid.mask.sum()) - ens.specs[m].side
        assert np.all(member[ens.interior[m] // STRIDE2] == m)
    with pytest.raises(ValueError):
        stack_tables([])


@pytest.mark.gpu
def test_batched_exact_equals_each_member_run_alone():
    """One launch for all members changes nothing: the exact float32 variant gives every member the fields it gets
    alone, driven from the host one step at a time."""
    ens = _small()
    got = run_batched(ens, "hip-exact", graph_steps=16, chunk=128, tune=True)
    alone = run_serial(ens, "hip-exact")
    print("steps batched", got.steps, "alone", alone.steps, "launch", got.launch, flush=True)
    assert got.steps == alone.steps
    for m in range(3):
        assert np.array_equal(got.max_depth[m], alone.max_depth[m]), m
        assert got.mass_error[m] == pytest.approx(alone.mass_error[m], abs=1e-12)
    assert got.t_end == [s.duration_s for s in (ens.scenarios[m] for m in range(3))]


@pytest.mark.gpu
def test_batched_exact_with_curve_number_losses_equals_each_member_run_alone():
    """Members with SCS losses (two of three; the third keeps its rain whole) in one launch: each member's rates
    formed on the device in the excess form, bit for bit what it gets alone the plain way."""
    ens = _small(1800.0)
    for m in (0, 2):
        shape = ens.scenarios[m].tables.grid.shape
        cn = 55.0 + 40.0 * np.random.default_rng(40 + m).random(shape)
        cn[:, : shape[1] // 3] = np.nan
        ens.scenarios[m].curve_number = cn
    assert ens.losses and np.isnan(ens.member_cn(1)).all()
    got = run_batched(ens, "hip-exact", graph_steps=16, chunk=128, tune=False)
    alone = run_serial(ens, "hip-exact")  # member 1 alone takes the weight form: the same rates, bit for bit
    print("losses: steps batched", got.steps, "alone", alone.steps, flush=True)
    assert got.steps == alone.steps
    for m in range(3):
        assert np.array_equal(got.max_depth[m], alone.max_depth[m]), m
        assert got.mass_error[m] == pytest.approx(alone.mass_error[m], abs=1e-12)
        assert (got.max_depth[m] > 0.01).sum() > 100, m


@pytest.mark.gpu
def test_batched_fast_float32_meets_the_fp64_oracle():
    ens = _small(900.0)
    got = run_batched(ens, "hip", graph_steps=16, chunk=128)
    ref = [run_oracle(sc, prec=8, threads=16) for sc in ens.scenarios]
    err = max_depth_error(got.max_depth, [r.max_depth for r in ref], ens)
    print("max depth error", err, "mass", got.mass_error, flush=True)
    assert max(err) <= 0.02
    assert max(abs(e) for e in got.mass_error) <= 1e-6
    assert [r.steps for r in ref] == got.steps
