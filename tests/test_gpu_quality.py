"""GPU parity of the on-device post-solve quality (pb_hip_quality) against the oracle's restatement of
finalizeBundleAdjustResult + the image quality gate, through the C ABI."""
import math
import sys

import numpy as np
import pytest

from plabundle_amd import _abi
from plabundle_amd import options as PO
from plabundle_amd import synthetic as S
from plabundle_amd.solver import Solver, SolverWorkspace, SolveStatus

from quality_fixture import ACTIVE, expected_rms, options, problem

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ws():
    lib = _abi.load_library()
    assert lib.pb_hip_available(0), "no gfx950 device"
    w = SolverWorkspace(0)
    yield w
    w.close()


def _same_report(a, b, rel):
    for k in ("total_tracks", "optimized_tracks", "solve_status", "solution_usable", "quality_gate_rejected"):
        assert a[k] == b[k], (k, a[k], b[k])
    for k in ("valid_track_ratio", "mean_rms_before", "mean_rms_after", "median_rms_after", "filter_threshold"):
        if math.isinf(b[k]):
            assert math.isinf(a[k]), k
        else:
            assert a[k] == pytest.approx(b[k], rel=rel), k


def test_quality_closed_forms(ws, oracle_mod):
    _say("case quality closed forms")
    p, o = problem(), options()
    r = Solver().solve(p, o, ws, precision=PO.Precision.Fp64)
    assert r.status == SolveStatus.Success
    for kw in ({}, {"min_accepted_valid_track_ratio": 0.3},
               {"filter_max_reproj_error": 0.0, "filter_sigma_factor": 0.1}):
        q = PO.quality_to_c(o)
        for k, v in kw.items():
            setattr(q, k, v)
        g = ws.quality(q)
        orc = oracle_mod.Oracle(p, threads=1)
        orc.solve(PO.to_c(o))
        ref = orc.quality(q)
        np.testing.assert_allclose(g["track_rms_before"], expected_rms(), rtol=1e-12)
        np.testing.assert_allclose(g["track_rms_after"], np.where(ACTIVE, expected_rms(), math.inf), rtol=1e-12)
        np.testing.assert_array_equal(g["track_valid"], ref["track_valid"])
        _same_report(g, ref, 1e-12)


@pytest.mark.parametrize("precision", [PO.Precision.Fp64, PO.Precision.Mixed])
def test_quality_after_solves_matches_oracle(ws, oracle_mod, precision):
    _say(f"case quality ring {precision.name}")
    p = S.ring_problem(seed=9, num_cameras=16, num_points=600, max_views=6)
    o = S.ring_options(PO.SolveOptions(), max_iterations=10)
    r = Solver().solve(p, o, ws, precision=precision, quality=True)
    assert r.usable() and r.quality is not None and not r.qualityGateRejected
    orc = oracle_mod.Oracle(p, threads=4)
    orc.solve(PO.to_c(o))
    ref = orc.quality(PO.quality_to_c(o))
    rel = 1e-7 if precision == PO.Precision.Fp64 else 2e-3  # the mixed state is within the 1e-6 cost bar
    np.testing.assert_allclose(r.quality["track_rms_before"], ref["track_rms_before"], rtol=1e-12)
    np.testing.assert_allclose(r.quality["track_rms_after"], ref["track_rms_after"], rtol=rel)
    np.testing.assert_array_equal(r.quality["track_valid"], ref["track_valid"])
    _same_report(r.quality, ref, rel)
    _say(f"kept {r.quality['optimized_tracks']}/{r.quality['total_tracks']}, rms "
         f"{r.quality['mean_rms_before']:.4f} -> {r.quality['mean_rms_after']:.4f} px")


def test_quality_requires_a_solve(ws):
    _say("case quality guards")
    lib = _abi.load_library()
    ws.upload(problem())
    rep = _abi.QualityReportC()
    q = PO.quality_to_c(options())
    import ctypes as C

    assert lib.pb_hip_quality(ws.handle, C.byref(q), C.byref(rep), None, None, None) == _abi.PB_ERR_INVALID


def test_outcome_of_gpu_solve_matches_oracle_outcome(ws, oracle_mod):
    # id-keyed outcome comparison (comparePhotoAlignmentOutcomes) of the GPU and the oracle solve, ids shuffled
    from plabundle_amd import photo_alignment as PA

    _say("case outcome comparison")
    p = S.ring_problem(seed=9, num_cameras=16, num_points=600, max_views=6)
    o = S.ring_options(PO.SolveOptions(), max_iterations=10)
    pa = PA.PhotoAlignmentProblem(p, [f"IMG_{i:03d}" for i in range(p.num_cameras)],
                                  [f"tie-{i}" for i in range(p.num_tracks)])
    gpu = PA.makePhotoAlignmentOutcome(pa, Solver().solve(p, o, ws, precision=PO.Precision.Fp64, quality=True))
    orc = oracle_mod.Oracle(p, threads=4)
    rep = orc.solve(PO.to_c(o))
    st = orc.state()
    q = orc.quality(PO.quality_to_c(o))
    ref = PA.PhotoAlignmentOutcome(
        terminationStatus=rep["solve_status"], terminationReason="oracle", requestedBackend=gpu.requestedBackend,
        usedBackend=gpu.usedBackend, solutionUsable=bool(q["solution_usable"]),
        meanRmsBefore=q["mean_rms_before"], meanRmsAfter=q["mean_rms_after"], camera_ids=pa.cameraIds,
        cam_rotation=st["cam_rotation"], cam_center=st["cam_center"], cam_intrinsics=st["cam_intrinsics"],
        cam_image_size=p.cam_image_size, track_ids=pa.trackIds, track_point=st["track_point"],
        track_valid=q["track_valid"], track_rms_before=q["track_rms_before"], track_rms_after=q["track_rms_after"])
    assert PA.validatePhotoAlignmentOutcome(ref) == (True, "")
    perm_c = np.random.default_rng(1).permutation(p.num_cameras)
    perm_t = np.random.default_rng(2).permutation(p.num_tracks)
    shuffled = PA.PhotoAlignmentOutcome(**{k: v for k, v in gpu.__dict__.items()})
    shuffled.camera_ids = [gpu.camera_ids[i] for i in perm_c]
    for name in ("cam_rotation", "cam_center", "cam_intrinsics", "cam_image_size"):
        setattr(shuffled, name, getattr(gpu, name)[perm_c])
    shuffled.track_ids = [gpu.track_ids[i] for i in perm_t]
    for name in ("track_point", "track_valid", "track_rms_before", "track_rms_after"):
        setattr(shuffled, name, getattr(gpu, name)[perm_t])
    tol = PA.PhotoAlignmentComparisonTolerance(rms=1e-8, cameraCenterMeters=1e-7, cameraRotationDegrees=1e-5,
                                               pointMeters=1e-6, requireTerminationReasonMatch=False)
    c = PA.comparePhotoAlignmentOutcomes(ref, shuffled, tol)
    _say(f"outcome: center {c['maximum_camera_center_difference_meters']:.2e} m, rotation "
         f"{c['maximum_camera_rotation_difference_degrees']:.2e} deg, point {c['maximum_point_difference_meters']:.2e} m")
    assert c["comparable"] and c["equivalent"], c["message"]
