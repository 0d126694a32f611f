"""CPU pins of the oracle's post-solve quality (finalizeBundleAdjustResult, BundleAdjustQuality.cpp:302-502, and
the image checks of resultFailsQualityGate, solver.cpp:123-171) on a problem with closed-form track RMS values."""
import math

import numpy as np
import pytest

from plabundle_amd import options as PO

from quality_fixture import ACTIVE, TRACKS, expected_rms, options, problem


@pytest.fixture(scope="module")
def solved(oracle_mod):
    orc = oracle_mod.Oracle(problem(), threads=1)
    rep = orc.solve(PO.to_c(options()))
    assert rep["solve_status"] == 1  # nothing to refine: converged
    return orc


def _q(oracle_mod, **kw):
    q = oracle_mod.default_quality_options()
    for k, v in kw.items():
        setattr(q, k, v)
    return q


def test_strict_track_rms_closed_forms(solved, oracle_mod):
    r = solved.quality(_q(oracle_mod))
    exp = expected_rms()
    np.testing.assert_allclose(r["track_rms_before"], exp, rtol=1e-12)
    # refined state = input state; tracks the solve never activated have refined RMS inf
    exp_after = np.where(ACTIVE, exp, math.inf)
    np.testing.assert_allclose(r["track_rms_after"], exp_after, rtol=1e-12)


def test_point_filter_median_threshold_and_statistics(solved, oracle_mod):
    r = solved.quality(_q(oracle_mod, min_accepted_valid_track_ratio=0.3))
    cand = sorted(v for v, a in zip(expected_rms(), ACTIVE) if a and math.isfinite(v))  # A, B, E, G
    median = 0.5 * (cand[1] + cand[2])
    assert r["median_rms_after"] == pytest.approx(median, rel=1e-12)
    assert r["filter_threshold"] == pytest.approx(max(2.5, 3.0 * median), rel=1e-12)
    kept = [0, 1, 4]  # G (7.07 px) is above the 2.5 px threshold
    np.testing.assert_array_equal(r["track_valid"], [1 if t in kept else 0 for t in range(len(TRACKS))])
    assert r["optimized_tracks"] == 3 and r["total_tracks"] == 7
    mean = np.mean([TRACKS[t][2] for t in kept])
    assert r["mean_rms_before"] == pytest.approx(mean, rel=1e-12)
    assert r["mean_rms_after"] == pytest.approx(mean, rel=1e-12)
    assert r["valid_track_ratio"] == pytest.approx(3 / 7)
    assert r["quality_gate_rejected"] == 0 and r["solution_usable"] == 1


def test_quality_gate_decisions(solved, oracle_mod):
    r = solved.quality(_q(oracle_mod))  # 3/7 < 0.6
    assert r["quality_gate_rejected"] == 1 and "valid-track ratio" in r["message"]
    r = solved.quality(_q(oracle_mod, min_accepted_valid_track_ratio=0.3, max_accepted_rms_growth=0.5))
    assert r["quality_gate_rejected"] == 1 and "growth" in r["message"]
    r = solved.quality(_q(oracle_mod, min_accepted_valid_track_ratio=0.3, enable_quality_gate=0))
    assert r["quality_gate_rejected"] == 0
    r = solved.quality(_q(oracle_mod, enable_point_filter=0, min_accepted_valid_track_ratio=0.3))
    assert math.isinf(r["filter_threshold"]) and r["optimized_tracks"] == 4


def test_no_kept_track_downgrades_the_solve(solved, oracle_mod):
    r = solved.quality(_q(oracle_mod, filter_max_reproj_error=0.0, filter_sigma_factor=0.1))
    assert r["optimized_tracks"] == 0
    assert r["solve_status"] == 7 and r["solution_usable"] == 0  # NumericalFailure
    assert math.isinf(r["mean_rms_after"])
    assert r["quality_gate_rejected"] == 1
