"""Per-member buffer parities of a device-resident run, through the C-ABI: a member held while another steps keeps
its own parity on the device, so it resumes exactly where it stood, and a read in between sees its state (the
stepper's loop, solver/stepper.py:101-129, run per member; the run's hold at an output, run.py:1712-1739)."""

from __future__ import annotations

import numpy as np
import pytest

from wadi2d_amd.engine.ensemble import Ensemble, MemberSpec, setup_batched
from wadi2d_amd.solver.state import from_state

FIELDS = ("h", "d", "u", "v", "mx", "my", "qx", "qy", "res", "om")


def _two_members():
    specs = [MemberSpec(0, 21, 96, 90.0, 0.5, 0.9), MemberSpec(1, 22, 64, 80.0, 0.5, 1.0)]
    return Ensemble.build(specs, dx=20.0, k=4, duration_s=1500.0)


def _member_fields(bt, ens, m):
    return {f: ens.member_field(m, bt.model.field(f)) for f in FIELDS}


@pytest.mark.gpu
@pytest.mark.parametrize("backend", ["hip-exact", "hip-f64"])
def test_a_member_held_for_an_odd_number_of_steps_resumes_where_it_stood(backend):
    ens = _two_members()
    free = setup_batched(ens, backend, maxima_every=4)
    end_free = free.run.run(chunk=64, graph_steps=16)
    held = setup_batched(ens, backend, maxima_every=4)
    b = held.model.backend
    b.run_pause([600.0, np.inf])
    s = held.run.steps(40, 16)
    while s[0].t < 600.0:
        s = held.run.steps(7, 0)
    t_hold, n_hold = s[0].t, s[0].nsteps
    at_hold = _member_fields(held, ens, 0)
    # member 1 steps on an odd number of times (in a graph and one by one) while member 0 stands
    n1 = s[1].nsteps
    for chunk, graph in ((7, 0), (32, 16), (6, 0)):
        s = held.run.steps(chunk, graph)
    assert s[0].t == t_hold and s[0].nsteps == n_hold
    assert (s[1].nsteps - n1) % 2 == 1
    # a read while it stands sees the state it was held at
    now = _member_fields(held, ens, 0)
    for f in FIELDS:
        assert np.array_equal(now[f], at_hold[f]), f
    b.run_pause([np.inf, np.inf])
    end_held = held.run.run(chunk=64, graph_steps=16)
    for m in range(2):
        assert end_held[m].nsteps == end_free[m].nsteps and end_held[m].t == end_free[m].t, m
        assert end_held[m].dt_min == end_free[m].dt_min, m
        fa, fb = _member_fields(free, ens, m), _member_fields(held, ens, m)
        for f in FIELDS:
            assert np.array_equal(fa[f], fb[f]), (m, f)
    for which in range(5):
        da = from_state(ens.tables.grid, free.model.backend.maxima_get(which))
        db = from_state(ens.tables.grid, held.model.backend.maxima_get(which))
        assert np.array_equal(da, db), which
    assert free.model.backend.tally_volume() == held.model.backend.tally_volume()
