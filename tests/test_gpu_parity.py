"""GPU parity of the MI355X backend against the FP64 oracle, through the C ABI.

Bars: integer block numbering bit-exact; FP64 path reduced systems within 1e-10 of the
oracle's (relative to the matrix scale) and final costs within 1e-9 relative with the same
iteration count; mixed path final cost within 1e-6 relative.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from plabundle_amd import _abi
from plabundle_amd import options as PO
from plabundle_amd import synthetic as S
from plabundle_amd.solver import Solver, SolverWorkspace, SolveStatus

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


def _ring(intrinsics=False, n=10, pts=200, seed=7):
    p = S.ring_problem(seed=seed, num_cameras=n, num_points=pts, max_views=6)
    o = S.ring_options(PO.SolveOptions(), max_iterations=10)
    if intrinsics:
        o.calibration.refineSharedFocalLength = True
        o.calibration.refineSharedRadialDistortion = True
    return p, o


def _aerial(n=(20, 20, 20), pts=4000, seed=2):
    p = S.aerial_problem(seed=seed, cameras_per_strip=n, num_points=pts)
    o = S.aerial_options(PO.SolveOptions())
    return p, o


@pytest.mark.parametrize("intrinsics", [False, True])
def test_masks_bit_exact(ws, oracle_mod, intrinsics):
    _say(f"case masks intrinsics={intrinsics}")
    p, o = _ring(intrinsics)
    orc = oracle_mod.Oracle(p, threads=2)
    orc.prepare(PO.to_c(o))
    ws.upload(p)
    ws.evaluate_cost(PO.to_c(o, PO.Precision.Fp64))
    a, b = ws.masks(p), orc.masks()
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


@pytest.mark.parametrize("intrinsics", [False, True])
@pytest.mark.parametrize("damping", [0.0, 1e-2])
def test_reduced_system_fp64(ws, oracle_mod, intrinsics, damping):
    _say(f"case reduced fp64 intrinsics={intrinsics} damping={damping}")
    p, o = _ring(intrinsics)
    co = PO.to_c(o, PO.Precision.Fp64)
    orc = oracle_mod.Oracle(p, threads=2)
    prep = orc.prepare(co)
    P = prep["primary_block_count"]
    A0, b0, c0, _ = orc.reduced_system(P, damping)
    ws.upload(p)
    A, b, c = ws.reduced_system(co, damping, P)
    scale = np.abs(A0).max()
    np.testing.assert_allclose(A, A0, atol=1e-10 * scale, rtol=0)
    np.testing.assert_allclose(b, b0, atol=1e-10 * np.abs(b0).max(), rtol=0)
    assert abs(c - c0) <= 1e-12 * c0


def test_reduced_system_mixed(ws, oracle_mod):
    _say("case reduced mixed")
    p, o = _ring(True)
    co = PO.to_c(o, PO.Precision.Mixed)
    orc = oracle_mod.Oracle(p, threads=2)
    prep = orc.prepare(co)
    P = prep["primary_block_count"]
    A0, b0, c0, _ = orc.reduced_system(P, 0.0)
    ws.upload(p)
    A, b, c = ws.reduced_system(co, 0.0, P)
    np.testing.assert_allclose(A, A0, atol=2e-5 * np.abs(A0).max(), rtol=0)
    # the Schur right-hand side correction sums FP32 cross blocks: ~1e-7 relative per term
    np.testing.assert_allclose(b, b0, atol=1e-7 * np.abs(b0).max(), rtol=0)
    assert abs(c - c0) <= 1e-12 * c0  # cost stays FP64


def test_native_three_camera_scene(ws, oracle_mod):
    _say("case native three cameras")
    p = S.native_three_cameras()
    o = PO.SolveOptions()
    o.solver.maxIterations = 8
    o.solver.useReferenceOnlineSchur = True  # as the reference scene (test_native_problem.cpp:84)
    r = Solver().solve(p, o, ws, precision=PO.Precision.Fp64)
    assert r.report["online_schur_used"] == 1 and r.usable(), r.backendMessage
    assert r.finalCost < r.initialCost
    ref = oracle_mod.Oracle(p, threads=1).solve(PO.to_c(o))
    assert r.report["iterations"] == ref["iterations"]
    np.testing.assert_allclose(r.cam_center[2], [2.0, 0.0, 0.0], atol=1e-6)


@pytest.mark.parametrize("intrinsics", [False, True])
def test_full_solve_fp64_matches_oracle(ws, oracle_mod, intrinsics):
    _say(f"case solve fp64 ring intrinsics={intrinsics}")
    p, o = _ring(intrinsics, n=16, pts=500)
    r = Solver().solve(p, o, ws, precision=PO.Precision.Fp64)
    orc = oracle_mod.Oracle(p, threads=4)
    ref = orc.solve(PO.to_c(o))
    st = orc.state()
    assert r.status == SolveStatus(ref["solve_status"])
    assert r.report["iterations"] == ref["iterations"]
    assert r.report["accepted_steps"] == ref["accepted_steps"]
    assert abs(r.finalCost - ref["final_cost"]) <= 1e-9 * ref["final_cost"]
    np.testing.assert_allclose(r.cam_center, st["cam_center"], atol=1e-7)
    np.testing.assert_allclose(r.track_point, st["track_point"], atol=1e-6)
    np.testing.assert_allclose(r.cam_intrinsics, st["cam_intrinsics"], rtol=1e-7, atol=1e-9)
    np.testing.assert_array_equal(r.track_valid, st["track_valid"])


def _track_length_mix(seed=5):
    # a ring scene whose tracks cover both point linearisation kernels: long tracks (more than a wave's 64 lanes
    # of observations) and tracks that see one camera twice go one thread per point, short tracks one lane per
    # observation
    p = S.ring_problem(seed=seed, num_cameras=72, num_points=240, max_views=72)
    off = p.track_offset
    cams, uvs, lens = [], [], []
    for t in range(len(off) - 1):
        c = p.obs_camera[off[t]:off[t + 1]]
        uv = p.obs_uv[off[t]:off[t + 1]]
        if t % 3 == 1:
            c, uv = c[:6], uv[:6]
        elif t % 3 == 2:  # a second observation in the first camera, half a pixel away
            c = np.concatenate([c[:5], c[:1]])
            uv = np.concatenate([uv[:5], uv[:1] + 0.5])
        cams.append(c)
        uvs.append(uv)
        lens.append(len(c))
    q = p.copy()
    q.track_offset = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    q.obs_camera = np.concatenate(cams).astype(np.int32)
    q.obs_uv = np.ascontiguousarray(np.concatenate(uvs))
    q.obs_weight = np.ones(len(q.obs_camera))
    q.obs_scale = np.ones(len(q.obs_camera))
    assert max(lens) > 64
    return q


@pytest.mark.parametrize("intrinsics", [False, True])
def test_long_and_repeated_camera_tracks_match_oracle(ws, oracle_mod, intrinsics):
    _say(f"case track length mix intrinsics={intrinsics}")
    p = _track_length_mix()
    o = S.ring_options(PO.SolveOptions(), max_iterations=8)
    if intrinsics:
        o.calibration.refineSharedFocalLength = True
        o.calibration.refineSharedRadialDistortion = True
    r = Solver().solve(p, o, ws, precision=PO.Precision.Fp64)
    orc = oracle_mod.Oracle(p, threads=8)
    ref = orc.solve(PO.to_c(o))
    st = orc.state()
    orc.close()
    _say(f"  {r.report['iterations']} it, cost {r.finalCost:.12g} vs {ref['final_cost']:.12g}")
    assert r.usable(), r.backendMessage
    assert r.report["iterations"] == ref["iterations"]
    assert r.report["accepted_steps"] == ref["accepted_steps"]
    assert abs(r.finalCost - ref["final_cost"]) <= 1e-9 * ref["final_cost"]
    np.testing.assert_allclose(r.cam_center, st["cam_center"], atol=1e-7)
    np.testing.assert_allclose(r.track_point, st["track_point"], atol=1e-6)
    np.testing.assert_allclose(r.cam_intrinsics, st["cam_intrinsics"], rtol=1e-7, atol=1e-9)


@pytest.mark.parametrize("route", [PO.ReducedSolver.Dense, PO.ReducedSolver.Pcg])
def test_full_solve_mixed_within_north_star_tolerance(ws, oracle_mod, route):
    _say(f"case solve mixed aerial-small route={route.name}")
    p, o = _aerial()
    o.backend.reducedSolver = route
    r = Solver().solve(p, o, ws, precision=PO.Precision.Mixed)
    ref = oracle_mod.Oracle(p, threads=4).solve(PO.to_c(o))
    assert r.usable(), r.backendMessage
    rel = abs(r.finalCost - ref["final_cost"]) / ref["final_cost"]
    _say(f"mixed rel cost error {rel:.3e} iterations {r.report['iterations']} vs {ref['iterations']} "
         f"({r.report['linear_solver_name']}, seeds {r.report['mixed_seed_steps']})")
    assert rel <= 1e-6


def test_aerial_fp64_matches_oracle(ws, oracle_mod):
    _say("case solve fp64 aerial-small")
    p, o = _aerial()
    r = Solver().solve(p, o, ws, precision=PO.Precision.Fp64)
    ref = oracle_mod.Oracle(p, threads=4).solve(PO.to_c(o))
    rel = abs(r.finalCost - ref["final_cost"]) / ref["final_cost"]
    _say(f"fp64 rel cost error {rel:.3e}")
    assert r.report["iterations"] == ref["iterations"]
    assert rel <= 1e-9


def _state_differences(a, b):
    """Largest centre distance (m), rotation angle (deg) and point distance (m) over tracks valid in both (the
    metrics of comparePhotoAlignmentOutcomes, src/photo_alignment.cpp:225-318), and valid-flag mismatches."""
    tr = np.einsum("ij,ij->i", a["cam_rotation"], b["cam_rotation"])
    rot = np.degrees(np.arccos(np.clip(0.5 * (tr - 1.0), -1.0, 1.0)))
    va, vb = a["track_valid"] != 0, b["track_valid"] != 0
    both = va & vb
    return (float(np.linalg.norm(a["cam_center"] - b["cam_center"], axis=1).max()), float(rot.max()),
            float(np.linalg.norm(a["track_point"][both] - b["track_point"][both], axis=1).max()),
            int(np.count_nonzero(va != vb)))


def _state(r):
    return {"cam_rotation": r.cam_rotation, "cam_center": r.cam_center, "track_point": r.track_point,
            "track_valid": r.track_valid}


def test_aerial_states_within_stated_tolerances(ws, oracle_mod):
    # FP64 (reference stop) against the oracle's refined state: 1e-6 m / 1e-5 deg; mixed against an FP64 solve with
    # the same truncated-Newton stop (the precision effect alone): 0.1 m, same valid tracks
    _say("case refined states aerial-small")
    p, o = _aerial()
    orc = oracle_mod.Oracle(p, threads=4)
    orc.solve(PO.to_c(o))
    ost = orc.state()
    r64 = Solver().solve(p, o, ws, precision=PO.Precision.Fp64)
    c, rdeg, pt, mm = _state_differences(ost, _state(r64))
    _say(f"  fp64 vs oracle: centre {c:.2e} m, rotation {rdeg:.2e} deg, point {pt:.2e} m, mismatches {mm}")
    assert c <= 1e-6 and rdeg <= 1e-5 and pt <= 1e-5 and mm == 0
    o.backend.reducedSolver = PO.ReducedSolver.Pcg  # the truncated-Newton PCG stop of the FP32-assembly mode
    rm = Solver().solve(p, o, ws, precision=PO.Precision.Mixed)
    oi = S.aerial_options(PO.SolveOptions())
    oi.backend.reducedSolver = PO.ReducedSolver.Pcg
    oi.backend.pcgModelEta = 0.1
    oi.backend.pcgRelativeTolerance = 1e-3
    oi.backend.pcgMaxIterations = max(200, 3 * rm.report["primary_block_count"] // 5)
    ri = Solver().solve(p, oi, ws, precision=PO.Precision.Fp64)
    c, rdeg, pt, mm = _state_differences(_state(ri), _state(rm))
    _say(f"  mixed vs fp64 same stop: centre {c:.2e} m, rotation {rdeg:.2e} deg, point {pt:.2e} m, mismatches {mm}")
    assert c <= 0.1 and pt <= 0.1 and mm == 0


def test_cancel_publishes_nothing(ws):
    _say("case cancel")
    p, o = _ring()
    r = Solver().solve(p, o, ws, precision=PO.Precision.Fp64, cancel=True)
    assert r.status == SolveStatus.Cancelled and not r.usable()
    np.testing.assert_array_equal(r.cam_center, p.cam_center)
    np.testing.assert_array_equal(r.track_point, p.track_point)


def test_invalid_problem_rejected(ws):
    _say("case invalid")
    p, o = _ring()
    bad = p.copy()
    bad.track_offset = bad.track_offset.copy()
    bad.track_offset[0] = 1
    assert ws.upload(bad) == _abi.PB_ERR_INVALID
    p2 = p.copy()
    p2.fixed_cameras = np.array([0], dtype=np.int32)
    r = Solver().solve(p2, o, ws, precision=PO.Precision.Fp64)
    assert r.status == SolveStatus.UnsupportedConfiguration


@pytest.mark.parametrize("precision", [PO.Precision.Fp64, PO.Precision.Mixed])
def test_pcg_graph_replay_matches_direct_launches(ws, precision):
    # the PCG iterations replayed from the captured graph (absolute iteration from the device counter) must give
    # bitwise the solve of one-by-one launches, also on a second solve (graph reused) and after the stop
    # parameters change (graph recaptured)
    _say(f"case PCG graph replay vs direct launches precision={precision.name}")
    p, o = _aerial(n=(12, 12, 12), pts=2000)
    o.backend.reducedSolver = PO.ReducedSolver.Pcg
    o.backend.pcgGraphs = False
    direct = Solver().solve(p, o, ws, precision=precision)
    assert direct.usable(), direct.backendMessage
    o.backend.pcgGraphs = True
    for rep in range(2):
        r = Solver().solve(p, o, ws, precision=precision)
        assert r.report["iterations"] == direct.report["iterations"]
        assert r.report["linear_iterations"] == direct.report["linear_iterations"]
        assert r.finalCost == direct.finalCost
        np.testing.assert_array_equal(r.cam_center, direct.cam_center)
        np.testing.assert_array_equal(r.cam_intrinsics, direct.cam_intrinsics)
        np.testing.assert_array_equal(r.track_point, direct.track_point)
    o.backend.pcgMaxIterations = 37  # odd cap: the stop falls inside a replay
    o.backend.pcgGraphs = False
    a = Solver().solve(p, o, ws, precision=precision)
    o.backend.pcgGraphs = True
    b = Solver().solve(p, o, ws, precision=precision)
    _say(f"  capped: {a.report['linear_iterations']} PCG iterations, cost {a.finalCost:.12g}")
    assert a.report["linear_iterations"] == b.report["linear_iterations"]
    assert a.finalCost == b.finalCost
    np.testing.assert_array_equal(a.track_point, b.track_point)


@pytest.mark.parametrize("route", [PO.ReducedSolver.Dense, PO.ReducedSolver.Pcg])
def test_auto_anchor_with_a_gauge_scale_camera_matches_oracle(ws, oracle_mod, route):
    # AutoAnchor without an absolute scale: camera 0 fixed, the farthest camera moves on the sphere of the baseline
    # through its two tangent columns (Assembly.cpp:69-128, 204-206, 295-306)
    _say(f"case AutoAnchor scale camera route={route.name}")
    from plabundle_amd.problem import GaugePolicy

    p = S.ring_problem(seed=23, num_cameras=14, num_points=400, max_views=6)
    p.fixed_cameras = None
    p.gauge_policy = GaugePolicy.AutoAnchor
    o = S.ring_options(PO.SolveOptions(), max_iterations=12)
    o.backend.reducedSolver = route
    orc = oracle_mod.Oracle(p, threads=4)
    ref = orc.solve(PO.to_c(o))
    st = orc.state()
    assert ref["gauge_scale_camera"] > 0 and ref["gauge_baseline"] > 0
    r = Solver().solve(p, o, ws, precision=PO.Precision.Fp64)
    assert r.usable(), r.backendMessage
    rep = r.report
    _say(f"  scale camera {rep['gauge_scale_camera']} baseline {rep['gauge_baseline']:.6f}; gpu {rep['iterations']} it "
         f"{r.finalCost:.12g}, oracle {ref['iterations']} it {ref['final_cost']:.12g}")
    assert (rep["gauge_anchor_camera"], rep["gauge_scale_camera"]) == (ref["gauge_anchor_camera"],
                                                                         ref["gauge_scale_camera"])
    assert rep["gauge_baseline"] == pytest.approx(ref["gauge_baseline"], rel=1e-14)
    assert rep["iterations"] == ref["iterations"] and rep["accepted_steps"] == ref["accepted_steps"]
    assert abs(r.finalCost - ref["final_cost"]) <= 1e-9 * ref["final_cost"]
    np.testing.assert_allclose(r.cam_center, st["cam_center"], atol=1e-7)
    # the scale camera stays on the sphere of the baseline around the anchor
    a, sc = rep["gauge_anchor_camera"], rep["gauge_scale_camera"]
    assert np.linalg.norm(r.cam_center[sc] - r.cam_center[a]) == pytest.approx(rep["gauge_baseline"], rel=1e-12)
    masks = orc.masks()
    gm = ws.masks(p)
    for k in gm:
        np.testing.assert_array_equal(gm[k], masks[k], err_msg=k)


def test_in_place_edit_between_solves_on_one_workspace(ws):
    # a caller that edits the problem's arrays in place between two solves on one workspace must get the solve of
    # the edited problem (the workspace re-uploads), identical to a fresh workspace's
    _say("case in-place edit between solves")
    p, o = _ring(n=12, pts=300, seed=11)
    first = Solver().solve(p, o, ws, precision=PO.Precision.Fp64)
    assert first.usable(), first.backendMessage
    p.cam_center[5] += np.array([0.3, -0.2, 0.1])
    p.track_point[:40] *= 1.01
    again = Solver().solve(p, o, ws, precision=PO.Precision.Fp64)
    fresh_ws = SolverWorkspace(0)
    try:
        fresh = Solver().solve(p, o, fresh_ws, precision=PO.Precision.Fp64)
    finally:
        fresh_ws.close()
    _say(f"  first {first.finalCost:.12g} ({first.report['initial_cost']:.6g} initial), edited "
         f"{again.finalCost:.12g} ({again.report['initial_cost']:.6g} initial), fresh {fresh.finalCost:.12g}")
    assert again.report["initial_cost"] != first.report["initial_cost"]
    assert again.report["iterations"] == fresh.report["iterations"]
    assert again.finalCost == fresh.finalCost
    np.testing.assert_array_equal(again.cam_center, fresh.cam_center)
    np.testing.assert_array_equal(again.track_point, fresh.track_point)
