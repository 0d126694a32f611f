"""GPU parity of the survey constraints (control points, LiDAR point-to-plane) and of the assembled path A
against the FP64 oracle, through the C ABI.

Bars as for the unconstrained path: block numbering (gauge from control points) bit-exact; FP64 reduced systems
within 1e-10 of the oracle's; FP64 solves within 1e-9 relative final cost with the same iteration and accepted
step counts; mixed solves within 1e-6; the adjoint's control-point gradients within 1e-7 of the dense reference
and within finite-difference accuracy of GPU re-solves."""
import ctypes as C
import sys

import numpy as np
import pytest

from plabundle_amd import _abi
from plabundle_amd import options as PO
from plabundle_amd import synthetic as S
from plabundle_amd.problem import GaugePolicy
from plabundle_amd.solver import Solver, SolverWorkspace, SolveStatus

from adjoint_reference import adjoint_reference

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


def _case(kind, n=14, pts=400, seed=7, intrinsics=False):
    p = S.ring_problem(seed=seed, num_cameras=n, num_points=pts, max_views=6)
    if kind == "control":
        p = S.add_survey_constraints(p, seed=3, num_control=10, kinds=(0, 1, 2), control_sigma=0.02)
    elif kind == "planes":
        p = S.add_survey_constraints(p, seed=3, num_control=6, num_planes=40, kinds=(0, 1, 2), control_sigma=0.02)
    o = S.ring_options(PO.SolveOptions(), max_iterations=15)
    o.solver.useReferenceOnlineSchur = kind != "pathA"
    if intrinsics:
        o.calibration.refineSharedFocalLength = True
        o.calibration.refineSharedRadialDistortion = True
    return p, o


@pytest.mark.parametrize("kind", ["control", "planes", "pathA"])
@pytest.mark.parametrize("damping", [0.0, 1e-2])
def test_reduced_system_fp64(ws, oracle_mod, kind, damping):
    _say(f"case constraints reduced fp64 kind={kind} damping={damping}")
    p, o = _case(kind, intrinsics=True)
    co = PO.to_c(o, PO.Precision.Fp64)
    orc = oracle_mod.Oracle(p, threads=2)
    prep = orc.prepare(co)
    P = prep["primary_block_count"]
    A0, b0, c0, _ = orc.reduced_system(P, damping)
    assert ws.upload(p) == _abi.PB_OK, ws.last_error()
    A, b, c = ws.reduced_system(co, damping, P)
    scale = np.abs(A0).max()
    np.testing.assert_allclose(A, A0, atol=1e-10 * scale, rtol=0)
    np.testing.assert_allclose(b, b0, atol=1e-10 * np.abs(b0).max(), rtol=0)
    assert abs(c - c0) <= 1e-12 * c0


@pytest.mark.parametrize("kind,intrinsics", [("control", False), ("control", True), ("planes", False),
                                             ("planes", True), ("pathA", True)])
def test_full_solve_fp64_matches_oracle(ws, oracle_mod, kind, intrinsics):
    _say(f"case constraints solve fp64 kind={kind} intrinsics={intrinsics}")
    p, o = _case(kind, intrinsics=intrinsics)
    r = Solver().solve(p, o, ws, precision=PO.Precision.Fp64)
    orc = oracle_mod.Oracle(p, threads=4)
    ref = orc.solve(PO.to_c(o))
    st = orc.state()
    _say(f"  gpu {r.report['iterations']} it / {r.finalCost:.12g}; oracle {ref['iterations']} it / "
         f"{ref['final_cost']:.12g}; online={r.report['online_schur_used']}")
    assert r.report["online_schur_used"] == ref["online_schur_used"] == (1 if kind == "control" else 0)
    assert r.status == SolveStatus(ref["solve_status"])
    assert r.report["iterations"] == ref["iterations"]
    assert r.report["accepted_steps"] == ref["accepted_steps"]
    assert r.report["objective_evaluations"] == ref["objective_evaluations"]
    assert abs(r.finalCost - ref["final_cost"]) <= 1e-9 * ref["final_cost"]
    np.testing.assert_allclose(r.cam_center, st["cam_center"], atol=1e-7)
    np.testing.assert_allclose(r.track_point, st["track_point"], atol=1e-6)
    np.testing.assert_allclose(r.cam_intrinsics, st["cam_intrinsics"], rtol=1e-7, atol=1e-9)


@pytest.mark.parametrize("kind", ["control", "planes"])
def test_full_solve_mixed_within_north_star_tolerance(ws, oracle_mod, kind):
    _say(f"case constraints solve mixed kind={kind}")
    p = S.aerial_problem(seed=2, cameras_per_strip=(20, 20, 20), num_points=4000)
    truth = p.meta["truth_points"]
    p = S.add_survey_constraints(p, seed=9, num_control=12, num_planes=0 if kind == "control" else 300,
                                 kinds=(0, 1, 2), control_sigma=0.05, truth_points=truth)
    o = S.aerial_options(PO.SolveOptions())
    r = Solver().solve(p, o, ws, precision=PO.Precision.Mixed)
    ref = oracle_mod.Oracle(p, threads=4).solve(PO.to_c(o))
    assert r.usable(), r.backendMessage
    rel = abs(r.finalCost - ref["final_cost"]) / ref["final_cost"]
    _say(f"  mixed rel cost error {rel:.3e} iterations {r.report['iterations']} vs {ref['iterations']}")
    assert rel <= 1e-6


def test_control_point_gauge_masks_bit_exact(ws, oracle_mod):
    _say("case control-point gauge masks")
    p = S.ring_problem(seed=5, num_cameras=10, num_points=150, max_views=6)
    p.fixed_cameras = None
    p.gauge_policy = GaugePolicy.AutoAnchor
    p = S.add_survey_constraints(p, seed=2, num_control=5, kinds=(0,), control_sigma=0.01)
    o = S.ring_options(PO.SolveOptions(), max_iterations=12)
    co = PO.to_c(o, PO.Precision.Fp64)
    orc = oracle_mod.Oracle(p, threads=2)
    prep = orc.prepare(co)
    assert prep["gauge_scale_camera"] == -1  # three non-collinear targets: nothing fixed
    assert ws.upload(p) == _abi.PB_OK
    ws.evaluate_cost(co)
    a, b = ws.masks(p), orc.masks()
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    assert (a["camera_block"] >= 0).all()
    r = Solver().solve(p, o, ws, precision=PO.Precision.Fp64)
    ref = orc.solve(co)
    assert r.report["iterations"] == ref["iterations"]
    assert abs(r.finalCost - ref["final_cost"]) <= 1e-9 * ref["final_cost"]


def test_invalid_constraints_rejected(ws):
    _say("case invalid constraints")
    p, o = _case("planes")
    for mutate in (lambda q: q.cp_point.__setitem__((0, 1), np.nan),
                   lambda q: q.cp_weight.__setitem__(1, 0.0),
                   lambda q: q.cp_matrix.__setitem__(1, np.diag([1.0, -1.0, 1.0]).reshape(9)),
                   lambda q: q.lp_normal.__setitem__(0, 0.0),
                   lambda q: q.lp_track.__setitem__(0, -1)):
        q = p.copy()
        mutate(q)
        assert ws.upload(q) == _abi.PB_ERR_INVALID
    assert ws.upload(p) == _abi.PB_OK
    o.constraints.laserPlaneWeight = -1.0
    r = Solver().solve(p, o, ws, precision=PO.Precision.Fp64)
    assert r.status == SolveStatus.InvalidInput


def _adjoint_setup(online):
    p0 = S.ring_problem(seed=11, num_cameras=8, num_points=60, max_views=5)
    p = S.add_survey_constraints(p0, seed=5, num_control=4, num_planes=0 if online else 6, kinds=(0, 1, 2),
                                 control_sigma=0.02)
    o = PO.SolveOptions()
    o.solver.maxIterations = 80
    o.solver.useReferenceOnlineSchur = online
    o.backend.stepRmsTolerance = 1e-13
    return p, o


@pytest.mark.parametrize("online", [True, False])
def test_backward_control_points_match_dense_reference(ws, oracle_mod, online):
    _say(f"case adjoint control points online={online}")
    p, o = _adjoint_setup(online)
    r = Solver().solve(p, o, ws, precision=PO.Precision.Fp64)
    assert r.usable(), r.backendMessage
    rng = np.random.default_rng(4)
    T = r.cam_center + rng.normal(0.0, 0.05, r.cam_center.shape)
    Y = r.track_point + rng.normal(0.0, 0.05, r.track_point.shape)
    dl_dC, dl_dX = 2.0 * (r.cam_center - T), 2.0 * (r.track_point - Y)
    g, gc = ws.backward(p, dl_dC, dl_dX, with_control=True)
    copts = PO.to_c(o, PO.Precision.Fp64)
    orc = oracle_mod.Oracle(p, threads=2)
    prep = orc.prepare(copts)
    masks = orc.masks()
    orc.close()
    state = {"cam_rotation": r.cam_rotation, "cam_center": r.cam_center, "cam_intrinsics": r.cam_intrinsics,
             "track_point": r.track_point}
    ref, refc = adjoint_reference(oracle_mod, p, state, masks, prep, copts, dl_dC, dl_dX, with_control=True)
    err = np.abs(g - ref).max() / np.abs(ref).max()
    errc = np.abs(gc - refc).max() / np.abs(refc).max()
    _say(f"  obs {err:.3e} control {errc:.3e}")
    assert err <= 1e-7 and errc <= 1e-7
    # finite differences of GPU re-solves for two control coordinates
    solver = Solver()
    h = 1e-4
    for i, k in ((0, 0), (2, 2)):
        vals = []
        for sgn in (1.0, -1.0):
            q = p.copy()
            q.cp_point[i, k] += sgn * h
            rq = solver.solve(q, o, ws, precision=PO.Precision.Fp64)
            assert rq.usable()
            vals.append(float(np.sum((rq.cam_center - T) ** 2) + np.sum((rq.track_point - Y) ** 2)))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - gc[i, k]) <= 2e-3 * np.abs(gc).max(), (i, k, fd, gc[i, k])


def test_survey_block_matches_oracle(ws, oracle_mod):
    # the large aerial block's setup at small size: strips with a full Metashape calibration over 8 groups,
    # 2 % outliers, 50 control points (datum from them, no fixed camera) and 1000 LiDAR planes (path A); 20 LM
    # iterations, which this block needs all of (it converges slowly in its weakly observed calibration slots)
    _say("case survey block (small) fp64 + mixed")
    p = S.large_aerial_problem(strips=4, cameras_per_strip=40, num_points=20000, num_control=50, num_planes=1000)
    o = S.large_aerial_options(PO.SolveOptions())
    ref = oracle_mod.Oracle(p, threads=8).solve(PO.to_c(o))
    r = Solver().solve(p, o, ws, precision=PO.Precision.Fp64)
    _say(f"  fp64 gpu {r.report['iterations']} it / {r.finalCost:.12g}; oracle {ref['iterations']} it / "
         f"{ref['final_cost']:.12g}")
    assert r.report["online_schur_used"] == 0 and ref["online_schur_used"] == 0
    assert r.report["accepted_steps"] == ref["accepted_steps"]
    assert abs(r.finalCost - ref["final_cost"]) <= 1e-9 * ref["final_cost"]
    m = Solver().solve(p, o, ws, precision=PO.Precision.Mixed)
    rel = abs(m.finalCost - ref["final_cost"]) / ref["final_cost"]
    _say(f"  mixed rel cost error after 20 iterations {rel:.3e} ({m.report['message']})")
    # the FP32 assembly hands over to FP64 at its rounding floor, so even at the budget the mixed solve meets the
    # 1e-6 bar (measured 1.9e-10)
    assert m.usable() and rel <= 1e-6
