"""GPU adjoint (pb_hip_backward) against the dense implicit-function reference (tests/adjoint_reference.py,
itself pinned by finite differences of oracle solves) and against finite differences of GPU solves.

Loss: l = sum_c |C_c - T_c|^2 (+ sum_t |X_t - Y_t|^2 for the point case) over the refined state of an FP64 solve
run to a step RMS of 1e-13."""
import sys

import numpy as np
import pytest

from plabundle_amd import options as PO
from plabundle_amd import synthetic as S
from plabundle_amd.solver import Solver, SolverWorkspace

from adjoint_reference import adjoint_reference

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ws():
    w = SolverWorkspace(0)
    yield w
    w.close()


def _setup(loss, intrinsics=False):
    p = S.ring_problem(seed=11, num_cameras=8, num_points=60, max_views=5)
    o = PO.SolveOptions()
    o.solver.imageRobustLoss = loss
    o.solver.useReferenceOnlineSchur = True
    o.solver.maxIterations = 80
    o.backend.stepRmsTolerance = 1e-13
    if intrinsics:
        o.calibration.refineSharedFocalLength = True
        o.calibration.refineSharedRadialDistortion = True
    return p, o


def _priors(p, copts, prep):
    """Inverse prior variances of the shared intrinsics (identical ring cameras: the reference values are the
    cameras' own), as the oracle assembles them."""
    out = {}
    mask = copts.shared_intrinsic_mask
    if not mask:
        return out
    f = p.cam_intrinsics[0, 0]
    sigma = {0: copts.focal_prior_sigma * f, 4: copts.prior_sigma[0], 5: copts.prior_sigma[1],
             6: copts.prior_sigma[2], 7: copts.prior_sigma[3], 8: copts.prior_sigma[4]}
    for gi in range(prep["intrinsic_block_count"]):
        blk = prep["camera_block_count"] + gi
        for q, sg in sigma.items():
            if (mask >> q) & 1:
                out[blk * 13 + q] = 1.0 / sg**2
    return out


@pytest.mark.parametrize("loss, intrinsics, points", [
    (PO.ImageRobustLoss.LeastSquares, False, False),
    (PO.ImageRobustLoss.Huber, False, True),
    (PO.ImageRobustLoss.Cauchy, False, False),
    (PO.ImageRobustLoss.Huber, True, False),
])
def test_backward_matches_dense_reference(ws, oracle_mod, loss, intrinsics, points):
    _say(f"adjoint vs dense reference: loss={loss.name} intrinsics={intrinsics} points={points}")
    p, o = _setup(loss, intrinsics)
    r = Solver().solve(p, o, ws, precision=PO.Precision.Fp64)
    assert r.usable(), r.backendMessage
    rng = np.random.default_rng(5)
    dl_dC = 2.0 * (r.cam_center - (r.cam_center + rng.normal(0.0, 0.05, r.cam_center.shape)))
    dl_dX = 2.0 * rng.normal(0.0, 0.01, r.track_point.shape) if points else None
    g = ws.backward(p, dl_dC, dl_dX)
    copts = PO.to_c(o, PO.Precision.Fp64)
    orc = oracle_mod.Oracle(p, threads=2)
    prep = orc.prepare(copts)
    masks = orc.masks()
    orc.close()
    state = {"cam_rotation": r.cam_rotation, "cam_center": r.cam_center, "cam_intrinsics": r.cam_intrinsics,
             "track_point": r.track_point}
    ref = adjoint_reference(oracle_mod, p, state, masks, prep, copts, dl_dC, dl_dX, _priors(p, copts, prep))
    scale = np.abs(ref).max()
    assert scale > 0
    err = np.abs(g - ref).max() / scale
    _say(f"  max |gpu - reference| / max |reference| = {err:.3e}")
    assert err <= 1e-7


def test_backward_matches_gpu_finite_differences(ws):
    _say("adjoint vs finite differences of GPU solves (least squares)")
    p, o = _setup(PO.ImageRobustLoss.LeastSquares)
    solver = Solver()
    r = solver.solve(p, o, ws, precision=PO.Precision.Fp64)
    assert r.usable()
    rng = np.random.default_rng(9)
    T = r.cam_center + rng.normal(0.0, 0.05, r.cam_center.shape)
    g = ws.backward(p, 2.0 * (r.cam_center - T))
    h = 1e-3
    for i in rng.choice(p.num_observations, size=3, replace=False):
        for k in range(2):
            vals = []
            for sgn in (1.0, -1.0):
                q = p.copy()
                q.obs_uv[i, k] += sgn * h
                rq = solver.solve(q, o, ws, precision=PO.Precision.Fp64)
                assert rq.usable()
                vals.append(float(np.sum((rq.cam_center - T) ** 2)))
            fd = (vals[0] - vals[1]) / (2 * h)
            assert abs(fd - g[i, k]) <= 2e-3 * np.abs(g).max(), (i, k, fd, g[i, k])


def test_backward_needs_a_solve(ws):
    import ctypes as C

    from plabundle_amd import _abi

    p, o = _setup(PO.ImageRobustLoss.LeastSquares)
    fresh = SolverWorkspace(0)
    fresh.upload(p)
    out = np.zeros((p.num_observations, 2))
    lib = _abi.load_library()
    assert lib.pb_hip_backward(fresh.handle, None, None, _abi.ptr(out, C.c_double), None) == _abi.PB_ERR_INVALID
    Solver().solve(p, o, fresh, precision=PO.Precision.Fp64)
    ctrl = np.full(3, 7.0)
    # a problem without control points leaves dl_dcontrol untouched
    assert lib.pb_hip_backward(fresh.handle, None, None, _abi.ptr(out, C.c_double),
                               _abi.ptr(ctrl, C.c_double)) == _abi.PB_OK
    np.testing.assert_array_equal(ctrl, 7.0)
    fresh.close()
