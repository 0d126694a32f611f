"""GPU parity of the band route (block-tridiagonal cyclic reduction of the reduced camera system, pb_band.hip).

The band route is an exact FP64 solve like the dense Cholesky, so on strip blocks whose camera graph it fits it
must reproduce the dense route and the FP64 oracle: the same LM trajectory (iterations, accepted steps), final cost
to 1e-10, refined state to 1e-7 m. Cases cover the three border shapes (no refined intrinsics: no border; shared
focal + Brown: one unit per group; full Metashape calibration: three units per group), the FP32 seed and mixed
modes, the adjoint, and the refusal of a camera graph that is too wide.
"""
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
    assert _abi.load_library().pb_hip_available(0), "no gfx950 device"
    w = SolverWorkspace(0)
    yield w
    w.close()


@pytest.fixture(scope="module")
def strips():
    # 3 strips of 40 cameras (120 pose units: bandwidth below 21, at least six super-blocks)
    return S.aerial_problem(seed=7, cameras_per_strip=(40, 40, 40), num_points=12_000)


@pytest.fixture(scope="module")
def strips_metashape():
    return S.aerial_problem(seed=8, cameras_per_strip=(30, 30, 30), num_points=9_000, metashape=True)


def _options(kind, max_iterations=25):
    o = S.aerial_options(PO.SolveOptions())
    o.solver.maxIterations = max_iterations
    c = o.calibration
    if kind == "none":
        c.refineSharedFocalLength = c.refineSharedRadialDistortion = c.refineSharedHighOrderDistortion = False
    elif kind == "full":
        c.refineSharedFocalLength = c.refineSharedFocalAspectRatio = c.refineSharedPrincipalPoint = True
        c.refineSharedRadialDistortion = c.refineSharedHighOrderDistortion = True
        c.refineSharedMetashapeParameters = True
    return o


def _route(o, route):
    o.backend.reducedSolver = route
    return o


def _state_gap(a, b):
    c = float(np.linalg.norm(a.cam_center - b.cam_center, axis=1).max())
    both = (a.track_valid != 0) & (b.track_valid != 0)
    pt = float(np.linalg.norm(a.track_point[both] - b.track_point[both], axis=1).max())
    return c, pt


@pytest.mark.parametrize("kind", ["none", "brown", "full"])
def test_band_matches_dense_and_oracle_fp64(ws, oracle_mod, strips, strips_metashape, kind):
    p = strips_metashape if kind == "full" else strips
    _say(f"case band fp64 kind={kind}: {p.cam_center.shape[0]} cameras, {p.num_observations} observations")
    o = _options(kind)
    rb = Solver().solve(p, _route(o, PO.ReducedSolver.Band), ws, precision=PO.Precision.Fp64)
    assert rb.usable(), rb.backendMessage
    assert rb.report["linear_solver_name"] == "band-cr-fp64"
    rd = Solver().solve(p, _route(o, PO.ReducedSolver.Dense), ws, precision=PO.Precision.Fp64)
    assert rd.usable(), rd.backendMessage
    orc = oracle_mod.Oracle(p, threads=8)
    ref = orc.solve(PO.to_c(o, PO.Precision.Fp64))
    orc.close()
    rel_bd = abs(rb.finalCost - rd.finalCost) / rd.finalCost
    rel_bo = abs(rb.finalCost - ref["final_cost"]) / ref["final_cost"]
    c, pt = _state_gap(rb, rd)
    _say(f"  band {rb.report['iterations']} it {rb.finalCost:.12g} ({rb.report['solve_seconds']:.3f}s), dense "
         f"{rd.report['iterations']} it ({rd.report['solve_seconds']:.3f}s), oracle {ref['iterations']} it; "
         f"band-dense {rel_bd:.2e}, band-oracle {rel_bo:.2e}, centre {c:.2e} m, point {pt:.2e} m")
    assert rb.report["iterations"] == rd.report["iterations"] == ref["iterations"]
    assert rb.report["accepted_steps"] == rd.report["accepted_steps"]
    assert rel_bd <= 1e-10 and rel_bo <= 1e-9
    assert c <= 1e-7 and pt <= 1e-7


def test_auto_takes_the_band_route_on_strips(ws, strips):
    _say("case band auto route")
    r = Solver().solve(strips, _options("brown", 3), ws, precision=PO.Precision.Fp64)
    assert r.usable(), r.backendMessage
    assert r.report["linear_solver_name"] == "band-cr-fp64"


@pytest.mark.parametrize("precision", [PO.Precision.Fp32Seed, PO.Precision.Mixed])
def test_band_seed_and_mixed_within_tolerance(ws, oracle_mod, strips, precision):
    _say(f"case band precision={precision.name}")
    o = _options("brown", 60)
    orc = oracle_mod.Oracle(strips, threads=8)
    ref = orc.solve(PO.to_c(o, PO.Precision.Fp64))
    orc.close()
    r = Solver().solve(strips, _route(o, PO.ReducedSolver.Band), ws, precision=precision)
    assert r.usable(), r.backendMessage
    rel = abs(r.finalCost - ref["final_cost"]) / ref["final_cost"]
    _say(f"  {r.report['iterations']} it, {r.report['linear_solver_name']}, seeds {r.report['mixed_seed_steps']}, "
         f"fallbacks {r.report['dense_fallbacks']}, rel cost {rel:.2e} ({r.report['message']})")
    assert rel <= 1e-6


def test_band_adjoint_matches_dense(ws, strips):
    _say("case band adjoint")
    o = _options("brown", 60)
    o.backend.stepRmsTolerance = 1e-12
    grads = {}
    for route in (PO.ReducedSolver.Dense, PO.ReducedSolver.Band):
        r = Solver().solve(strips, _route(o, route), ws, precision=PO.Precision.Fp64)
        assert r.usable(), r.backendMessage
        rng = np.random.default_rng(3)
        T = r.cam_center + rng.normal(0.0, 0.05, r.cam_center.shape)
        grads[route] = ws.backward(strips, 2.0 * (r.cam_center - T))
    gd, gb = grads[PO.ReducedSolver.Dense], grads[PO.ReducedSolver.Band]
    err = np.abs(gb - gd).max() / np.abs(gd).max()
    _say(f"  max |band - dense| / max |dense| = {err:.2e}")
    assert err <= 1e-8


def test_band_route_refuses_a_wide_camera_graph(ws):
    _say("case band refuses the ring")
    p = S.ring_problem(seed=1, num_cameras=50, num_points=2000, max_views=8)
    o = _route(S.ring_options(PO.SolveOptions(), max_iterations=2), PO.ReducedSolver.Band)
    r = Solver().solve(p, o, ws, precision=PO.Precision.Fp64)
    assert r.status == SolveStatus.UnsupportedConfiguration
    assert "band route" in r.backendMessage
    # auto falls back to the dense route
    ra = Solver().solve(p, S.ring_options(PO.SolveOptions(), max_iterations=2), ws, precision=PO.Precision.Fp64)
    assert ra.usable() and ra.report["linear_solver_name"] == "dense-cholesky-fp64"
