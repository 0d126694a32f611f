"""GPU parity of the full frame calibration (all 13 slots of a Metashape-convention group refined: three 6-wide
units per group) and of the reference calibration transition prior, against the FP64 oracle through the C ABI.

Bars as for the other GPU parity tests: FP64 reduced systems within 1e-10 of the oracle's, FP64 solves within
1e-9 relative final cost with the same iteration count, mixed solves within 1e-6."""
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


def _case(n=12, pts=300, seed=8, transition=False, previous=None, iters=25):
    p = S.ring_problem(seed=seed, num_cameras=n, num_points=pts, max_views=6, metashape=True)
    o = S.ring_options(PO.SolveOptions(), max_iterations=iters)
    c = o.calibration
    c.refineSharedFocalLength = c.refineSharedFocalAspectRatio = c.refineSharedPrincipalPoint = True
    c.refineSharedRadialDistortion = c.refineSharedHighOrderDistortion = True
    c.refineSharedMetashapeParameters = True
    c.useReferenceCalibrationTransitionPrior = transition
    if previous is not None:
        c.referencePreviousIntrinsicParameterMask = list(previous)
    return p, o


@pytest.mark.parametrize("damping", [0.0, 1e-2])
def test_thirteen_slot_reduced_system_fp64(ws, oracle_mod, damping):
    _say(f"case 13-slot reduced fp64 damping={damping}")
    p, o = _case()
    co = PO.to_c(o, PO.Precision.Fp64)
    assert co.shared_intrinsic_mask == 0x1FFF
    orc = oracle_mod.Oracle(p, threads=2)
    P = orc.prepare(co)["primary_block_count"]
    A0, b0, c0, _ = orc.reduced_system(P, damping)
    assert ws.upload(p) == _abi.PB_OK
    A, b, c = ws.reduced_system(co, damping, P)
    np.testing.assert_allclose(A, A0, atol=1e-10 * np.abs(A0).max(), rtol=0)
    np.testing.assert_allclose(b, b0, atol=1e-10 * np.abs(b0).max(), rtol=0)
    assert abs(c - c0) <= 1e-12 * c0
    a, m = ws.masks(p), orc.masks()
    for k in a:
        np.testing.assert_array_equal(a[k], m[k], err_msg=k)


@pytest.mark.parametrize("transition", [False, True])
def test_thirteen_slot_solve_fp64_matches_oracle(ws, oracle_mod, transition):
    _say(f"case 13-slot solve fp64 transition={transition}")
    previous = [True, True, True, True, True, True, False, True, True, False, False, False, False]
    p, o = _case(transition=transition, previous=previous)
    r = Solver().solve(p, o, ws, precision=PO.Precision.Fp64)
    orc = oracle_mod.Oracle(p, threads=4)
    ref = orc.solve(PO.to_c(o))
    st = orc.state()
    _say(f"  gpu {r.report['iterations']} it / {r.finalCost:.12g}; oracle {ref['iterations']} it / "
         f"{ref['final_cost']:.12g}")
    assert r.status == SolveStatus(ref["solve_status"])
    assert r.report["iterations"] == ref["iterations"]
    assert abs(r.finalCost - ref["final_cost"]) <= 1e-9 * ref["final_cost"]
    np.testing.assert_allclose(r.cam_center, st["cam_center"], atol=1e-7)
    np.testing.assert_allclose(r.cam_intrinsics, st["cam_intrinsics"], rtol=1e-7, atol=1e-9)


def test_thirteen_slot_solve_mixed_within_north_star_tolerance(ws, oracle_mod):
    _say("case 13-slot solve mixed")
    p, o = _case(n=16, pts=600)
    r = Solver().solve(p, o, ws, precision=PO.Precision.Mixed)
    ref = oracle_mod.Oracle(p, threads=4).solve(PO.to_c(o))
    assert r.usable(), r.backendMessage
    rel = abs(r.finalCost - ref["final_cost"]) / ref["final_cost"]
    _say(f"  mixed rel cost error {rel:.3e} iterations {r.report['iterations']} vs {ref['iterations']}")
    assert rel <= 1e-6


def test_structure_follows_calibration_options(ws, oracle_mod):
    # the intrinsic groups depend on the prior options: a second solve with other sigmas on the same upload
    # must use them (the prepared structure is rebuilt)
    _say("case calibration options change between solves")
    p, o = _case(iters=6)
    assert ws.upload(p) == _abi.PB_OK
    r1 = Solver().solve(p, o, ws, precision=PO.Precision.Fp64)
    o.calibration.sharedRadialK1PriorSigma = 1e-4
    o.calibration.sharedFocalPriorSigma = 1e-3
    r2 = Solver().solve(p, o, ws, precision=PO.Precision.Fp64)
    ref = oracle_mod.Oracle(p, threads=4).solve(PO.to_c(o))
    assert abs(r2.finalCost - ref["final_cost"]) <= 1e-9 * ref["final_cost"]
    assert abs(r1.finalCost - r2.finalCost) > 1e-9 * r2.finalCost
