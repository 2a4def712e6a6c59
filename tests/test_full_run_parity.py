"""A whole run on the HIP path against the oracle: the plain 256 x 256 case at 10 m (wadi2d_amd.synth.plain_cpu_case:
1 m sub-grid, 8-row tables, a crest wall with a gap, 50 mm/h for 30 min, 2 h, a free outflow column), FP64 on both
sides. Limits: per-cell maximum depth within 1 mm, peak outfall discharge within 0.1 %, relative mass error of the
FP64 books below 1e-7 on both sides. The float32 path is held to 2 cm of maximum depth and the same mass limit."""

from __future__ import annotations

import numpy as np
import pytest

from tests.runs import run_hip, run_oracle
from wadi2d_amd import synth


@pytest.fixture(scope="module")
def oracle_fp64():
    return run_oracle(synth.plain_cpu_case(), prec=8, threads=16)


@pytest.mark.gpu
def test_plain_case_fp64_run_matches_the_fp64_oracle(oracle_fp64):
    ref = oracle_fp64
    got = run_hip(synth.plain_cpu_case(), "hip-f64")
    print(f"steps {got.steps}/{ref.steps} peak q {got.peak_q:.6f}/{ref.peak_q:.6f} mass {got.mass_error:.2e}/"
          f"{ref.mass_error:.2e} max |d max depth| {np.abs(got.max_depth - ref.max_depth).max():.3e}", flush=True)
    assert got.steps == ref.steps
    assert np.abs(got.max_depth - ref.max_depth).max() <= 1e-3
    assert abs(got.peak_q - ref.peak_q) <= 1e-3 * abs(ref.peak_q)
    assert abs(got.mass_error) <= 1e-7 and abs(ref.mass_error) <= 1e-7
    assert (ref.max_depth > 0.01).sum() > 10000  # the storm wets the domain


@pytest.mark.gpu
@pytest.mark.parametrize("backend", ["hip", "hip-exact"])
def test_plain_case_fp32_run_within_two_centimetres(oracle_fp64, backend):
    ref = oracle_fp64
    got = run_hip(synth.plain_cpu_case(), backend)
    err = np.abs(got.max_depth - ref.max_depth).max()
    print(f"{backend}: steps {got.steps}/{ref.steps} peak q {got.peak_q:.6f}/{ref.peak_q:.6f} mass "
          f"{got.mass_error:.2e} max |d max depth| {err:.3e}", flush=True)
    assert err <= 0.02
    assert abs(got.peak_q - ref.peak_q) <= 1e-2 * abs(ref.peak_q)
    assert abs(got.mass_error) <= 1e-7
