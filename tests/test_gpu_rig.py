"""GPU parity of camera-rig solves (captures and sensors refined instead of camera poses) against the FP64 oracle,
through the C ABI.

The device assembles the camera-level Schur complement and maps it onto the rig units (S_rig = T^T S_cam T); the
oracle assembles the reference's per-observation capture and sensor terms directly
(BundleAdjustPlaMatrixAssembly.cpp:207-284), so the two reach the rig system by different routes. Bars: block
numbering bit-exact; FP64 rig systems within 1e-10 of the oracle's (relative to the matrix scale); FP64 solves with
the oracle's iteration and accepted-step counts, final cost within 1e-9 and refined rig, cameras and points within
1e-7 m; mixed within 1e-6 of the oracle's final cost.
"""
import sys

import numpy as np
import pytest

from plabundle_amd import _abi
from plabundle_amd import options as PO
from plabundle_amd import synthetic as S
from plabundle_amd.problem import GaugePolicy
from plabundle_amd.solver import Solver, SolverWorkspace

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


def _small_fixed_rig():
    # 5 captures x 2 sensors with a fixed capture (2), a fixed sensor (0) and a fixed camera (0, which holds
    # capture 0); camera 4 = (capture 2, sensor 0) has both parts fixed, so it is a fixed camera too and the
    # explicit gauge holds. Captures 0 and 2 are the generator's unperturbed ones.
    p = S.rig_problem(seed=31, num_captures=5, num_points=200, sensor_yaw_deg=(0.0, 10.0), fixed_sensors=(0,))
    p.rig_capture_fixed = np.array([0, 0, 1, 0, 0], dtype=np.int32)
    p.fixed_cameras = np.array([0], dtype=np.int32)
    p.gauge_policy = GaugePolicy.RequireExplicitGauge
    return p


CASES = {
    "auto-gauge 16x3": lambda: S.rig_problem(),
    "fixed parts 5x2": _small_fixed_rig,
    "all sensors free 8x3": lambda: S.rig_problem(seed=9, num_captures=8, num_points=300, fixed_sensors=()),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_rig_masks_bit_exact(ws, oracle_mod, case):
    _say(f"case rig masks {case}")
    p = CASES[case]()
    o = PO.to_c(S.rig_options(PO.SolveOptions(), max_iterations=3), PO.Precision.Fp64)
    orc = oracle_mod.Oracle(p, threads=2)
    assert orc.prepare(o)["prepared"]
    ws.upload(p)
    ws.evaluate_cost(o)
    a, b = ws.masks(p), orc.masks()
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    a, b = ws.rig_masks(p), orc.rig_masks()
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("damping", [0.0, 1e-2])
def test_rig_reduced_system_fp64(ws, oracle_mod, case, damping):
    _say(f"case rig reduced fp64 {case} damping={damping}")
    p = CASES[case]()
    co = PO.to_c(S.rig_options(PO.SolveOptions()), PO.Precision.Fp64)
    orc = oracle_mod.Oracle(p, threads=2)
    prep = orc.prepare(co)
    P = prep["primary_block_count"]
    A0, b0, c0, _ = orc.reduced_system(P, damping)
    ws.upload(p)
    A, b, c = ws.reduced_system(co, damping, P)
    np.testing.assert_allclose(A, A0, atol=1e-10 * np.abs(A0).max(), rtol=0)
    np.testing.assert_allclose(b, b0, atol=1e-10 * np.abs(b0).max(), rtol=0)
    assert abs(c - c0) <= 1e-12 * c0


def _check_state(r, orc, p, tol):
    st = orc.state()
    np.testing.assert_allclose(r.cam_center, st["cam_center"], atol=tol, rtol=0)
    np.testing.assert_allclose(r.cam_rotation, st["cam_rotation"], atol=tol / 10.0, rtol=0)
    np.testing.assert_allclose(r.track_point, st["track_point"], atol=tol, rtol=0)
    rig = orc.rig_state()
    for k in rig:
        np.testing.assert_allclose(r.refinedRig[k], rig[k], atol=tol, rtol=0, err_msg=k)


@pytest.mark.parametrize("solver", [PO.ReducedSolver.Dense, PO.ReducedSolver.Pcg])
@pytest.mark.parametrize("case", sorted(CASES))
def test_rig_solve_fp64_matches_oracle(ws, oracle_mod, case, solver):
    _say(f"case rig solve fp64 {case} {solver.name}")
    p = CASES[case]()
    o = S.rig_options(PO.SolveOptions(), max_iterations=15)
    o.backend.reducedSolver = solver
    pcg = solver == PO.ReducedSolver.Pcg
    if pcg:
        # a rig's sensors couple every capture: block-Jacobi PCG needs more than the reference cap of 30 P
        # iterations to reach the path's 1e-12 tolerance, and a step it does not reach is rejected
        o.backend.pcgMaxIterations = 20000
    o.solver.maxIterations = 40 if pcg else 15
    orc = oracle_mod.Oracle(p, threads=4)
    ref = orc.solve(PO.to_c(o))
    r = Solver().solve(p, o, ws, precision=PO.Precision.Fp64)
    assert r.usable(), r.backendMessage
    assert r.report["online_schur_used"] == 0
    assert r.report["camera_block_count"] == ref["camera_block_count"]
    assert r.finalCost < 0.05 * r.initialCost
    _check_composition(r, p)
    if pcg:
        # The oracle solves every step exactly (dense Cholesky). With every sensor free and only the captures of
        # the automatic gauge's two cameras held, a rigid motion commuting with those captures' relative pose
        # (a screw about its axis) moves the world and the rig frame together without changing any camera: the
        # undamped system is singular, the factorisation rejects that step while CG still returns one, and the
        # two runs settle at different points of the same minimum. Where the gauge is complete both converge to
        # the same state.
        assert ref["converged"] and r.report["converged"]
        assert abs(r.finalCost - ref["final_cost"]) <= 1e-9 * ref["final_cost"]
        if case != "all sensors free 8x3":
            _check_state(r, orc, p, 1e-5)
        return
    assert r.report["iterations"] == ref["iterations"]
    assert r.report["accepted_steps"] == ref["accepted_steps"]
    assert abs(r.finalCost - ref["final_cost"]) <= 1e-9 * ref["final_cost"]
    _check_state(r, orc, p, 1e-7)


def _check_composition(r, p):
    # the refined cameras are the composition of the refined rig
    rig = r.refinedRig
    Rk = rig["capture_rotation"].reshape(-1, 3, 3)[p.cam_rig_capture]
    Rs = rig["sensor_rotation"].reshape(-1, 3, 3)[p.cam_rig_sensor]
    np.testing.assert_allclose(r.cam_rotation.reshape(-1, 3, 3), Rk @ Rs, atol=1e-14)
    Cc = rig["capture_center"][p.cam_rig_capture] + np.einsum("nij,nj->ni", Rk, rig["sensor_center"][p.cam_rig_sensor])
    np.testing.assert_allclose(r.cam_center, Cc, atol=1e-12)


def test_rig_mixed_within_north_star_tolerance(ws, oracle_mod):
    _say("case rig mixed")
    p = S.rig_problem()
    o = S.rig_options(PO.SolveOptions(), max_iterations=30)
    ref = oracle_mod.Oracle(p, threads=4).solve(PO.to_c(o))
    assert ref["converged"]
    r = Solver().solve(p, o, ws, precision=PO.Precision.Mixed)
    assert r.usable(), r.backendMessage
    assert abs(r.finalCost - ref["final_cost"]) <= 1e-6 * ref["final_cost"]


def test_identity_rig_matches_the_plain_assembled_solve(ws):
    # one capture per camera with a fixed identity sensor is the plain problem on the assembled path
    _say("case identity rig")
    p = S.ring_problem(seed=1, num_cameras=16, num_points=400)
    q = p.copy()
    n = p.num_cameras
    q.rig_capture_rotation = p.cam_rotation.copy()
    q.rig_capture_center = p.cam_center.copy()
    q.rig_capture_fixed = np.zeros(n, dtype=np.int32)
    q.rig_sensor_rotation = np.eye(3).reshape(1, 9).copy()
    q.rig_sensor_center = np.zeros((1, 3))
    q.rig_sensor_fixed = np.ones(1, dtype=np.int32)
    q.cam_rig_capture = np.arange(n, dtype=np.int32)
    q.cam_rig_sensor = np.zeros(n, dtype=np.int32)
    o = S.ring_options(PO.SolveOptions(), max_iterations=15)
    o.solver.useReferenceOnlineSchur = False
    o.calibration.refineSharedFocalLength = True
    o.calibration.refineSharedRadialDistortion = True
    a = Solver().solve(p, o, ws, precision=PO.Precision.Fp64)
    b = Solver().solve(q, o, ws, precision=PO.Precision.Fp64)
    assert a.report["iterations"] == b.report["iterations"]
    assert abs(a.finalCost - b.finalCost) <= 1e-12 * a.finalCost
    np.testing.assert_allclose(b.cam_center, a.cam_center, atol=1e-10)
    np.testing.assert_allclose(b.refinedRig["capture_center"], a.cam_center, atol=1e-10)


def test_rig_unsupported_paths_are_refused(ws):
    _say("case rig refusals")
    p = S.rig_problem(seed=31, num_captures=5, num_points=150, sensor_yaw_deg=(0.0, 10.0))
    o = S.rig_options(PO.SolveOptions(), max_iterations=3)
    o.backend.reducedSolver = PO.ReducedSolver.Band
    r = Solver().solve(p, o, ws, precision=PO.Precision.Fp64)
    assert r.abi_status == _abi.PB_ERR_UNSUPPORTED, r.backendMessage
    o.backend.reducedSolver = PO.ReducedSolver.Auto
    r = Solver().solve(p, o, ws, precision=PO.Precision.Fp64)
    assert r.usable(), r.backendMessage
    with pytest.raises(RuntimeError, match="rig"):
        ws.backward(p, dl_dcenter=np.zeros((p.num_cameras, 3)))
    results, summary = Solver().solve_batch([p, p], o, ws, precision=PO.Precision.Fp64)
    assert summary["abi_status"] == _abi.PB_ERR_UNSUPPORTED
