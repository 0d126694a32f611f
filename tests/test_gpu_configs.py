"""Every benchmark configuration at its stated size, on the GPU through the C ABI.

  C1  50 cameras on a 10 m ring, 2000 points, at most 8 views, Huber 1 px, 10 LM iterations: against the oracle run
      in the test (same iterations and final cost to 1e-9, states to 1e-7 m, masks bit-exact).
  C2  the 1.5M-observation aerial block, solved to the reference's convergence stop: against the committed converged
      oracle result (tests/golden: final cost and iteration count, every camera and a 2000-track point sample).
  C3  the 15M-observation survey block (10k cameras, 2M points, full Metashape calibration over 8 groups, 50 control
      points, 10k LiDAR planes; path A): properties that hold at any size - accepted objectives never increase, the
      mixed and FP64 solves agree, and the control points' residuals drop.
  C4  16 windows drawn from the full batched-window distribution (cameras U{8..64}, points U{500..8000}): window by
      window against the oracle.
  C5  the adjoint on a 20-camera / 2000-point converged sub-problem: against the dense implicit-function reference
      and against central differences (h = 1e-4 px) of full GPU re-solves.
"""
import json
import os
import sys
import time

import numpy as np
import pytest

from plabundle_amd import _abi
from plabundle_amd import options as PO
from plabundle_amd import synthetic as S
from plabundle_amd.solver import Solver, SolverWorkspace, SolveStatus

from adjoint_reference import adjoint_reference

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ws():
    assert _abi.load_library().pb_hip_available(0), "no gfx950 device"
    w = SolverWorkspace(0)
    yield w
    w.close()


def _state_diff(ref, r, index=None):
    # rotation angle between R and R' from ||R - R'||_F = 2 sqrt(2) sin(angle / 2): unlike
    # acos((tr(R^T R') - 1) / 2) it keeps its precision near zero (acos has a 4e-8 rad floor in FP64)
    fro = np.linalg.norm((ref["cam_rotation"] - r.cam_rotation).reshape(-1, 9), axis=1)
    rot = np.degrees(2.0 * np.arcsin(np.clip(fro / (2.0 * np.sqrt(2.0)), 0.0, 1.0)))
    X = r.track_point if index is None else r.track_point[index]
    V = r.track_valid if index is None else r.track_valid[index]
    both = (ref["track_valid"] != 0) & (V != 0)
    return (float(np.linalg.norm(ref["cam_center"] - r.cam_center, axis=1).max()), float(rot.max()),
            float(np.linalg.norm(ref["track_point"][both] - X[both], axis=1).max()),
            int(np.count_nonzero((ref["track_valid"] != 0) != (V != 0))))


def test_c1_ring_matches_oracle(ws, oracle_mod):
    _say("case C1: 50 cameras, 2000 points, <= 8 views, Huber, 10 iterations")
    p = S.ring_problem(seed=1, num_cameras=50, num_points=2000, max_views=8)
    o = S.ring_options(PO.SolveOptions(), max_iterations=10)
    orc = oracle_mod.Oracle(p, threads=8)
    ref = orc.solve(PO.to_c(o))
    st = orc.state()
    masks = orc.masks()
    r = Solver().solve(p, o, ws, precision=PO.Precision.Fp64)
    assert r.usable(), r.backendMessage
    _say(f"  gpu {r.report['iterations']} it cost {r.finalCost:.12g}; oracle {ref['iterations']} it "
         f"{ref['final_cost']:.12g} ({r.report['linear_solver_name']})")
    assert p.num_observations > 10_000
    assert r.status == SolveStatus(ref["solve_status"])
    assert r.report["iterations"] == ref["iterations"]
    assert abs(r.finalCost - ref["final_cost"]) <= 1e-9 * ref["final_cost"]
    c, rdeg, pt, mm = _state_diff(st, r)
    _say(f"  states: centre {c:.2e} m, rotation {rdeg:.2e} deg, point {pt:.2e} m, mismatches {mm}")
    assert c <= 1e-7 and rdeg <= 1e-6 and pt <= 1e-6 and mm == 0
    gm = ws.masks(p)
    for k in gm:
        np.testing.assert_array_equal(gm[k], masks[k], err_msg=k)
    m = Solver().solve(p, o, ws, precision=PO.Precision.Mixed)
    rel = abs(m.finalCost - ref["final_cost"]) / ref["final_cost"]
    _say(f"  mixed rel cost error {rel:.2e} ({m.report['message']})")
    assert m.usable() and rel <= 1e-6


@pytest.fixture(scope="module")
def aerial():
    golden = json.load(open(os.path.join(GOLDEN, "aerial_oracle.json")))
    state = dict(np.load(os.path.join(GOLDEN, "aerial_oracle_state.npz")))
    t = time.time()
    p = S.aerial_problem()
    _say(f"  generated the aerial block in {time.time() - t:.1f}s")
    import hashlib

    h = hashlib.sha256()
    for a in (p.cam_rotation, p.cam_center, p.cam_intrinsics, p.track_point, p.track_offset, p.obs_camera, p.obs_uv):
        h.update(np.ascontiguousarray(a).tobytes())
    assert h.hexdigest() == golden["fingerprint"], "the generator no longer makes the golden problem"
    return p, golden, state


def test_c2_aerial_block_converges_to_the_oracle(ws, aerial):
    _say("case C2: 1.5M observations to the reference's convergence stop")
    p, golden, gst = aerial
    conv = golden["converged_run"]
    o = S.aerial_options(PO.SolveOptions())
    o.solver.maxIterations = 400
    r = Solver().solve(p, o, ws, precision=PO.Precision.Fp64)
    assert r.usable(), r.backendMessage
    rel = abs(r.finalCost - conv["final_cost"]) / conv["final_cost"]
    c, rdeg, pt, mm = _state_diff(gst, r, gst["point_index"])
    _say(f"  fp64: {r.report['iterations']} it ({r.status.name}) in {r.report['solve_seconds']:.3f}s, rel cost "
         f"{rel:.2e}; centre {c:.2e} m, rotation {rdeg:.2e} deg, point {pt:.2e} m, mismatches {mm}")
    assert r.status == SolveStatus.Success and r.report["iterations"] == conv["iterations"]
    assert rel <= 1e-9
    assert c <= 1e-4 and rdeg <= 1e-4 and pt <= 1e-4 and mm == 0
    m = Solver().solve(p, o, ws, precision=PO.Precision.Mixed)
    assert m.usable(), m.backendMessage
    relm = abs(m.finalCost - conv["final_cost"]) / conv["final_cost"]
    c, rdeg, pt, mm = _state_diff(gst, m, gst["point_index"])
    _say(f"  mixed: {m.report['iterations']} it ({m.status.name}) in {m.report['solve_seconds']:.3f}s, rel cost "
         f"{relm:.2e}; centre {c:.2e} m, rotation {rdeg:.2e} deg, point {pt:.2e} m, mismatches {mm}")
    assert m.status == SolveStatus.Success and relm <= 1e-6
    assert c <= 1e-2 and rdeg <= 1e-3 and pt <= 1e-2 and mm == 0


def test_c3_survey_block_full_size_properties(ws):
    _say("case C3: survey block, 15M observations, path A")
    t = time.time()
    p = S.large_aerial_problem()
    _say(f"  generated {p.num_cameras} cameras, {p.num_tracks} tracks, {p.num_observations} observations in "
         f"{time.time() - t:.1f}s")
    assert p.num_cameras == 10_000 and p.num_observations > 14_000_000
    costs = []

    def progress(sm):
        costs.append(sm.averageRmsPixels)
        return True

    o = S.large_aerial_options(PO.SolveOptions())
    o.solver.progressCallback = progress
    # the reference's budget; at this size the reduced system (60k rows) goes to PCG with the truncated stop
    o.backend.pcgRelativeTolerance, o.backend.pcgModelEta = 1e-3, 0.1
    o.backend.mixedSwitchDecrease = -1.0
    m = Solver().solve(p, o, ws, precision=PO.Precision.Mixed)
    assert m.usable(), m.backendMessage
    mixed_costs = np.array(costs)
    costs.clear()
    o.backend.pcgMaxIterations = max(200, 3 * m.report["primary_block_count"] // 5)
    r = Solver().solve(p, o, ws, precision=PO.Precision.Fp64)
    assert r.usable(), r.backendMessage
    rel = abs(m.finalCost - r.finalCost) / r.finalCost
    _say(f"  mixed {m.report['iterations']} it {m.finalCost:.10g} in {m.report['solve_seconds']:.2f}s; fp64 same stop "
         f"{r.report['iterations']} it {r.finalCost:.10g}; rel {rel:.2e}")
    assert m.report["online_schur_used"] == 0 and r.report["online_schur_used"] == 0
    # the objective after every trial never increases (rejected trials keep it), and drops overall
    for seq in (mixed_costs, np.array(costs)):
        assert np.all(np.diff(seq) <= 1e-12 * seq[:-1])
    assert m.finalCost < m.initialCost and r.finalCost < r.initialCost
    assert rel <= 1e-6

    def cp_rms(X):
        d = X[p.cp_track] - p.cp_point
        return float(np.sqrt(np.mean(np.sum(d * d, axis=1))))

    before, after = cp_rms(p.track_point), cp_rms(r.track_point)
    _say(f"  control-point RMS {before:.4f} m -> {after:.4f} m (mixed {cp_rms(m.track_point):.4f} m)")
    assert after < 0.5 * before and cp_rms(m.track_point) < 0.5 * before


def test_c4_sixteen_windows_from_the_full_distribution(ws, oracle_mod):
    _say("case C4: 16 windows, cameras U{8..64}, points U{500..8000}")
    wins = S.batch_windows(seed=4, num_windows=16)
    cams = [w.num_cameras for w in wins]
    pts = [w.num_tracks for w in wins]
    _say(f"  cameras {min(cams)}..{max(cams)}, points {min(pts)}..{max(pts)}")
    assert max(cams) - min(cams) >= 20 and max(pts) - min(pts) >= 2000
    o = S.window_options(PO.SolveOptions(), max_iterations=20)
    res, summary = Solver().solve_batch(wins, o, ws, precision=PO.Precision.Fp64)
    assert summary["abi_status"] == _abi.PB_OK, summary.get("message")
    co = PO.to_c(o)
    worst = 0.0
    for w, (p, r) in enumerate(zip(wins, res)):
        orc = oracle_mod.Oracle(p, threads=8)
        ref = orc.solve(co)
        st = orc.state()
        rel = abs(r.finalCost - ref["final_cost"]) / ref["final_cost"]
        worst = max(worst, rel)
        assert r.report["iterations"] == ref["iterations"], w
        assert r.report["solve_status"] == ref["solve_status"], w
        assert rel <= 1e-9, (w, rel)
        c, rdeg, pt, mm = _state_diff(st, r)
        assert c <= 1e-6 and pt <= 1e-5 and mm == 0, (w, c, pt, mm)
    _say(f"  worst window rel cost error {worst:.2e}")


# The adjoint is a Gauss-Newton adjoint: it leaves out the residual-weighted second derivatives of the
# projection, so against finite differences of full re-solves it is exact only where the converged residuals
# vanish. With 0.5 px noise the omitted term measured 0.29 % of the largest gradient; noise-free it must agree to
# the finite-difference accuracy.
@pytest.mark.parametrize("noise_px,fd_tolerance", [(0.5, 1e-2), (0.0, 1e-4)])
def test_c5_adjoint_on_a_converged_subproblem(ws, oracle_mod, noise_px, fd_tolerance):
    _say(f"case C5: adjoint, 20 cameras / 2000 points, Huber, noise {noise_px} px, FD h = 1e-4 px")
    p = S.ring_problem(seed=5, num_cameras=20, num_points=2000, max_views=8, noise_px=noise_px)
    o = S.ring_options(PO.SolveOptions(), max_iterations=100)
    o.backend.stepRmsTolerance = 1e-13
    solver = Solver()
    r = solver.solve(p, o, ws, precision=PO.Precision.Fp64)
    assert r.usable(), r.backendMessage
    rng = np.random.default_rng(5)
    T = p.meta["truth_center"] + rng.normal(0.0, 0.05, r.cam_center.shape)
    g = ws.backward(p, 2.0 * (r.cam_center - T))
    # dense implicit-function reference from the oracle's Jacobians at the GPU's converged state
    copts = PO.to_c(o, PO.Precision.Fp64)
    orc = oracle_mod.Oracle(p, threads=8)
    prep = orc.prepare(copts)
    masks = orc.masks()
    orc.close()
    state = {"cam_rotation": r.cam_rotation, "cam_center": r.cam_center, "cam_intrinsics": r.cam_intrinsics,
             "track_point": r.track_point}
    ref = adjoint_reference(oracle_mod, p, state, masks, prep, copts, 2.0 * (r.cam_center - T), None, {})
    err = np.abs(g - ref).max() / np.abs(ref).max()
    _say(f"  max |gpu - dense reference| / max |reference| = {err:.2e}")
    assert err <= 1e-7
    # central differences of full re-solves: the observations with the largest gradients
    h = 1e-4
    order = np.argsort(-np.abs(g).ravel())[:4]
    for flat in order:
        i, k = divmod(int(flat), 2)
        vals = []
        for sgn in (1.0, -1.0):
            q = p.copy()
            q.obs_uv[i, k] += sgn * h
            rq = solver.solve(q, o, ws, precision=PO.Precision.Fp64)
            assert rq.usable()
            vals.append(float(np.sum((rq.cam_center - T) ** 2)))
        fd = (vals[0] - vals[1]) / (2 * h)
        _say(f"  obs {i} axis {k}: adjoint {g[i, k]:.6e}, finite difference {fd:.6e}")
        assert abs(fd - g[i, k]) <= fd_tolerance * np.abs(g).max()


def test_c5_rotation_gradient_noise_free(ws):
    """dl/dR through pb_hip_backward_pose: l = sum_c <A_c, R_c> + |C - T|^2 on the noise-free C5 problem, where the
    Gauss-Newton adjoint is exact, against central differences of full re-solves."""
    _say("case C5: adjoint with a rotation loss, 20 cameras / 2000 points, noise-free, FD h = 1e-4 px")
    p = S.ring_problem(seed=5, num_cameras=20, num_points=2000, max_views=8, noise_px=0.0)
    o = S.ring_options(PO.SolveOptions(), max_iterations=100)
    o.backend.stepRmsTolerance = 1e-13
    solver = Solver()
    r = solver.solve(p, o, ws, precision=PO.Precision.Fp64)
    assert r.usable(), r.backendMessage
    rng = np.random.default_rng(6)
    A = rng.normal(0.0, 1.0, (p.num_cameras, 9))
    T = p.meta["truth_center"] + rng.normal(0.0, 0.05, r.cam_center.shape)

    def loss(res):
        return float(np.sum(A * res.cam_rotation.reshape(-1, 9)) + np.sum((res.cam_center - T) ** 2))

    g = ws.backward(p, 2.0 * (r.cam_center - T), dl_drotation=A)
    g_center_only = ws.backward(p, 2.0 * (r.cam_center - T))
    assert np.abs(g - g_center_only).max() > 0.1 * np.abs(g).max()  # the rotation term matters
    h = 1e-4
    order = np.argsort(-np.abs(g).ravel())[:4]
    for flat in order:
        i, k = divmod(int(flat), 2)
        vals = []
        for sgn in (1.0, -1.0):
            q = p.copy()
            q.obs_uv[i, k] += sgn * h
            rq = solver.solve(q, o, ws, precision=PO.Precision.Fp64)
            assert rq.usable()
            vals.append(loss(rq))
        fd = (vals[0] - vals[1]) / (2 * h)
        _say(f"  obs {i} axis {k}: adjoint {g[i, k]:.6e}, finite difference {fd:.6e}")
        assert abs(fd - g[i, k]) <= 1e-4 * np.abs(g).max()


def test_adjoint_ridge_path_on_an_aerial_block(ws, oracle_mod):
    """Cauchy loss: an observation whose residual exceeds the loss scale sits on the branch where the curvature
    along the residual is negative and clamped, so it adds a rank-1 term; a two-view track with both residuals
    above 1 px leaves its point block singular, and the adjoint gives it a 1e-9 ridge. Such tracks are planted in
    a 120-camera strip block; at convergence Cauchy moves the whole disagreement onto one observation, so the
    adjoint is taken after two LM iterations, where both residuals are still large. The check is the linear
    algebra of the ridge path: the GPU adjoint (band route) against the CPU implicit-function reference with the
    same ridge rule, at the same state."""
    _say("case adjoint ridge path: aerial strips, 120 cameras, Cauchy, planted two-view disagreements")
    p = S.aerial_problem(seed=7, cameras_per_strip=(40, 40, 40), num_points=12_000)
    # two-view tracks whose observations disagree by several pixels: at convergence both residuals stay above the
    # Cauchy scale (1 px), where the along-residual curvature is negative and clamped, so each observation adds a
    # rank-1 term and the point block is singular
    two = np.nonzero(np.diff(p.track_offset) == 2)[0][:12]
    assert len(two) == 12
    for t in two:
        o0 = p.track_offset[t]
        p.obs_uv[o0] += (7.0, -5.0)
        p.obs_uv[o0 + 1] += (-7.0, 5.0)
    o = S.aerial_options(PO.SolveOptions())
    c = o.calibration
    c.refineSharedFocalLength = c.refineSharedRadialDistortion = c.refineSharedHighOrderDistortion = False
    o.solver.maxIterations = 2
    r = Solver().solve(p, o, ws, precision=PO.Precision.Fp64)
    assert r.report["iterations"] == 2 and r.report["linear_solver_name"] == "band-cr-fp64", r.backendMessage
    rng = np.random.default_rng(8)
    T = r.cam_center + rng.normal(0.0, 0.05, r.cam_center.shape)
    g = ws.backward(p, 2.0 * (r.cam_center - T))
    copts = PO.to_c(o, PO.Precision.Fp64)
    orc = oracle_mod.Oracle(p, threads=8)
    prep = orc.prepare(copts)
    masks = orc.masks()
    orc.close()
    state = {"cam_rotation": r.cam_rotation, "cam_center": r.cam_center, "cam_intrinsics": r.cam_intrinsics,
             "track_point": r.track_point}
    stats = {}
    t = time.time()
    ref = adjoint_reference(oracle_mod, p, state, masks, prep, copts, 2.0 * (r.cam_center - T), None, {},
                            stats=stats)
    err = np.abs(g - ref).max() / np.abs(ref).max()
    _say(f"  reference in {time.time() - t:.1f}s, ridged point blocks {stats['ridged_point_blocks']}, "
         f"max |gpu - reference| / max |reference| = {err:.2e}")
    assert stats["ridged_point_blocks"] > 0
    # a ridged block has a condition number up to 1e9, so two FP64 evaluations in different orders agree to
    # about 1e-7..1e-6 of the largest gradient
    assert err <= 5e-6
