"""Oracle pins for camera rigs (CPU).

The rig branch of the frame-BA path (BundleAdjustPlaMatrixProblem.cpp:157-199 block numbering,
BundleAdjustPlaMatrixAssembly.cpp:207-284 capture and sensor Jacobians, :1152-1208 rig step and camera
composition, Conversion.cpp:83-142 default groups and fixed cameras) is pinned by:
  - hand-written block numbering of a small rig (fixed captures, fixed sensors, fixed cameras, cameras without
    residuals, the gauge scale camera's capture);
  - the analytic directional decrease along every capture and sensor parameter against central differences of
    the objective through the rig step and the composition (a wrong lever-arm sign or rig axis fails it);
  - a rig of one capture per camera with a fixed identity sensor, which is the plain assembled-Schur problem:
    its solve must be bitwise the plain one;
  - the conversion rules and invalid rigs.
"""
import numpy as np
import pytest

from plabundle_amd import options as PO
from plabundle_amd import synthetic as S
from plabundle_amd.problem import GaugePolicy


def drop_cameras(problem, cameras):
    """Copy of `problem` without the observations of `cameras` (tracks keep their order)."""
    out = problem.copy()
    keep = ~np.isin(problem.obs_camera, np.asarray(cameras))
    off = problem.track_offset
    counts = np.array([keep[off[t]:off[t + 1]].sum() for t in range(problem.num_tracks)])
    out.track_offset = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    out.obs_camera = problem.obs_camera[keep].copy()
    out.obs_uv = problem.obs_uv[keep].copy()
    out.obs_weight = problem.obs_weight[keep].copy()
    out.obs_scale = problem.obs_scale[keep].copy()
    return out


def small_rig():
    # 5 captures x 2 sensors: camera 2k + s is (capture k, sensor s)
    return S.rig_problem(seed=31, num_captures=5, num_points=200, sensor_yaw_deg=(0.0, 10.0), fixed_sensors=(0,))


def test_rig_block_numbering_known_answer(oracle_mod):
    p = drop_cameras(small_rig(), [8, 9])  # capture 4 has no residual
    p.rig_capture_fixed = np.array([0, 1, 0, 0, 0], dtype=np.int32)  # capture 1 fixed
    p.fixed_cameras = np.array([0, 5], dtype=np.int32)  # camera 0 fixes capture 0, camera 5 fixes capture 2
    p.gauge_policy = GaugePolicy.RequireExplicitGauge
    opts = S.rig_options(PO.SolveOptions(), max_iterations=3)
    orc = oracle_mod.Oracle(p, threads=2)
    prep = orc.prepare(PO.to_c(opts))
    assert prep["prepared"], prep["message"]
    m = orc.rig_masks()
    # only capture 3 is refined (block 0); sensor 0 is fixed, sensor 1 follows the captures (block 1)
    np.testing.assert_array_equal(m["capture_block"], [-1, -1, -1, -1, -1, -1, 0, 0, -1, -1])
    np.testing.assert_array_equal(m["sensor_block"], [-1, 1, -1, 1, -1, 1, -1, 1, -1, 1])
    masks = orc.masks()
    np.testing.assert_array_equal(masks["camera_block"], np.full(10, -1))
    # one calibration group per sensor by default (Conversion.cpp:83-98); intrinsic blocks after the rig blocks
    np.testing.assert_array_equal(masks["calibration_group_by_camera"], np.tile([0, 1], 5))
    np.testing.assert_array_equal(masks["intrinsic_block_by_camera"], np.tile([2, 3], 5))
    assert prep["camera_block_count"] == 2 and prep["primary_block_count"] == 4


def test_automatic_gauge_fixes_the_scale_cameras_capture(oracle_mod):
    p = small_rig()
    opts = S.rig_options(PO.SolveOptions(), max_iterations=3, intrinsics=False)
    orc = oracle_mod.Oracle(p, threads=2)
    prep = orc.prepare(PO.to_c(opts))
    assert prep["gauge_anchor_camera"] == 0
    scale = prep["gauge_scale_camera"]
    assert scale > 1
    m = orc.rig_masks()
    kb = m["capture_block"]
    fixed_captures = {0, int(p.cam_rig_capture[scale])}
    want, nxt = [], 0
    for k in range(5):
        want.append(-1 if k in fixed_captures else nxt)
        nxt += k not in fixed_captures
    np.testing.assert_array_equal(kb, np.repeat(want, 2))
    np.testing.assert_array_equal(m["sensor_block"], np.tile([-1, nxt], 5))


def test_cameras_with_both_parts_fixed_count_as_fixed(oracle_mod):
    # every capture fixed and sensor 0 fixed: the cameras of sensor 0 are fixed cameras (Conversion.cpp:120-142),
    # so the explicit gauge is satisfied without listing any
    p = small_rig()
    p.rig_capture_fixed = np.ones(5, dtype=np.int32)
    p.gauge_policy = GaugePolicy.RequireExplicitGauge
    orc = oracle_mod.Oracle(p, threads=2)
    prep = orc.prepare(PO.to_c(S.rig_options(PO.SolveOptions(), intrinsics=False)))
    assert prep["prepared"], prep["message"]
    m = orc.rig_masks()
    np.testing.assert_array_equal(m["capture_block"], np.full(10, -1))
    np.testing.assert_array_equal(m["sensor_block"], np.tile([-1, 0], 5))


@pytest.mark.parametrize("loss", [PO.ImageRobustLoss.LeastSquares, PO.ImageRobustLoss.Huber])
def test_rig_directional_decrease_matches_central_differences(oracle_mod, loss):
    p = small_rig()
    opts = S.rig_options(PO.SolveOptions(), max_iterations=3)
    opts.solver.imageRobustLoss = loss
    orc = oracle_mod.Oracle(p, threads=2)
    prep = orc.prepare(PO.to_c(opts))
    P, Ne = prep["primary_block_count"], prep["track_block_count"]
    m = orc.rig_masks()
    cap_block = int(m["capture_block"][m["capture_block"] >= 0][0])
    sen_block = int(m["sensor_block"][m["sensor_block"] >= 0][0])
    zero = np.zeros(Ne * 3)

    def fd(step, h):
        return -(orc.cost_after_step(step, zero, h) - orc.cost_after_step(step, zero, -h)) / (2 * h)

    for block in (cap_block, sen_block):
        for comp in range(6):
            step = np.zeros(P * 13)
            step[block * 13 + comp] = 1e-3
            want = orc.decrease(step, zero)
            assert abs(want) > 0
            assert fd(step, 1e-2) == pytest.approx(want, rel=2e-6, abs=1e-9 * abs(want) + 1e-12), (block, comp)
    # and along the solved step (captures, sensors, intrinsics and points together)
    primary, elim, dd, _, _ = orc.compute_step(P, Ne, 1e-3)
    assert orc.decrease(primary, elim) == pytest.approx(dd, rel=1e-12)
    h = 1e-5
    got = -(orc.cost_after_step(primary, elim, h) - orc.cost_after_step(primary, elim, -h)) / (2 * h)
    assert got == pytest.approx(dd, rel=2e-5)


def _one_capture_per_camera(p):
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
    return q


@pytest.mark.parametrize("radial", [False, True])
def test_identity_rig_is_the_plain_assembled_problem(oracle_mod, radial):
    # one capture per camera and a fixed identity sensor: the capture Jacobian is the camera's (zero lever arm),
    # the composition returns the capture pose unchanged, so the rig solve is bitwise the plain path-A solve
    p = S.ring_problem(seed=1, num_cameras=16, num_points=400)
    q = _one_capture_per_camera(p)
    o = S.ring_options(PO.SolveOptions(), max_iterations=15)
    o.solver.useReferenceOnlineSchur = False
    o.calibration.refineSharedFocalLength = True
    o.calibration.refineSharedRadialDistortion = radial
    a = oracle_mod.Oracle(p, threads=2)
    ra = a.solve(PO.to_c(o))
    b = oracle_mod.Oracle(q, threads=2)
    rb = b.solve(PO.to_c(o))
    assert rb["online_schur_used"] == 0
    assert ra["iterations"] == rb["iterations"] and ra["accepted_steps"] == rb["accepted_steps"]
    assert ra["final_cost"] == rb["final_cost"]
    sa, sb = a.state(), b.state()
    for k in ("cam_rotation", "cam_center", "track_point"):
        np.testing.assert_array_equal(sa[k], sb[k])
    rig = b.rig_state()
    np.testing.assert_array_equal(rig["capture_rotation"], sa["cam_rotation"])
    np.testing.assert_array_equal(rig["capture_center"], sa["cam_center"])


def test_rig_solve_reaches_the_noise_floor_and_recovers_the_rig(oracle_mod):
    p = S.rig_problem()
    orc = oracle_mod.Oracle(p, threads=4)
    rep = orc.solve(PO.to_c(S.rig_options(PO.SolveOptions(), max_iterations=30)))
    assert rep["converged"], rep["message"]
    n_res = 2 * p.num_observations
    # Huber 1 px with N(0, 0.5 px) pixels: about 0.5 * 0.25 per residual at the minimum
    assert rep["final_cost"] < 0.2 * n_res
    rig = orc.rig_state()
    assert np.abs(rig["sensor_center"] - p.meta["truth_sensor_center"]).max() < 5e-3
    st = orc.state()
    # the refined cameras are the composition of the refined rig
    R = rig["capture_rotation"].reshape(-1, 3, 3)[p.cam_rig_capture] @ rig["sensor_rotation"].reshape(-1, 3, 3)[p.cam_rig_sensor]
    Cc = rig["capture_center"][p.cam_rig_capture] + np.einsum(
        "nij,nj->ni", rig["capture_rotation"].reshape(-1, 3, 3)[p.cam_rig_capture], rig["sensor_center"][p.cam_rig_sensor])
    np.testing.assert_allclose(st["cam_rotation"].reshape(-1, 3, 3), R, atol=1e-14)
    np.testing.assert_allclose(st["cam_center"], Cc, atol=1e-12)


def test_invalid_rigs_are_rejected(oracle_mod):
    opts = PO.to_c(S.rig_options(PO.SolveOptions(), intrinsics=False))
    p = small_rig()
    p.cam_rig_capture = p.cam_rig_capture.copy()
    p.cam_rig_capture[3] = -1
    rep = oracle_mod.Oracle(p, threads=1).prepare(opts)
    assert not rep["prepared"] and "binding" in rep["message"]
    p = small_rig()
    p.rig_sensor_rotation = p.rig_sensor_rotation.copy()
    p.rig_sensor_rotation[1, 0] *= 1.01
    rep = oracle_mod.Oracle(p, threads=1).prepare(opts)
    assert not rep["prepared"] and "rigid transform" in rep["message"]
