"""Pins the dense implicit-function adjoint (tests/adjoint_reference.py) against central finite differences of
tightly converged oracle solves: l = sum_c |C_c - T_c|^2 over the refined camera centres."""
import numpy as np
import pytest

from plabundle_amd import options as PO
from plabundle_amd import synthetic as S
from adjoint_reference import adjoint_reference


def _setup(loss):
    p = S.ring_problem(seed=11, num_cameras=8, num_points=60, max_views=5)
    opts = PO.SolveOptions()
    opts.solver.imageRobustLoss = loss
    opts.solver.maxIterations = 80
    opts.backend.stepRmsTolerance = 1e-13
    return p, PO.to_c(opts)


def _solve(O, p, copts):
    orc = O.Oracle(p, threads=2)
    rep = orc.solve(copts)
    assert rep["solution_usable"]
    st = orc.state()
    prep = orc.prepare(copts)
    masks = orc.masks()
    orc.close()
    return st, prep, masks, rep


# Measured on this problem: least squares and Huber agree to about 6e-4 of the largest gradient entry for
# h = 1e-2, 1e-3 and 1e-4 alike (so the gap is the projection's second derivatives that the adjoint leaves out,
# not difference noise); Cauchy to about 4.5e-2, because the negative curvature of its outliers is clamped.
@pytest.mark.parametrize("loss, tol", [(PO.ImageRobustLoss.LeastSquares, 2e-3), (PO.ImageRobustLoss.Huber, 2e-3),
                                       (PO.ImageRobustLoss.Cauchy, 8e-2)])
def test_implicit_function_gradient_matches_finite_differences(oracle_mod, loss, tol):
    O = oracle_mod
    p, copts = _setup(loss)
    st, prep, masks, rep = _solve(O, p, copts)
    rng = np.random.default_rng(3)
    T = st["cam_center"] + rng.normal(0.0, 0.05, st["cam_center"].shape)
    dl_dC = 2.0 * (st["cam_center"] - T)
    grad = adjoint_reference(O, p, st, masks, prep, copts, dl_dC)

    def loss_at(q):
        s2, *_ = _solve(O, q, copts)
        return float(np.sum((s2["cam_center"] - T) ** 2))

    h = 1e-3
    picks = rng.choice(p.num_observations, size=6, replace=False)
    for i in picks:
        for k in range(2):
            qp, qm = p.copy(), p.copy()
            qp.obs_uv[i, k] += h
            qm.obs_uv[i, k] -= h
            fd = (loss_at(qp) - loss_at(qm)) / (2 * h)
            assert abs(fd - grad[i, k]) <= tol * np.abs(grad).max(), (i, k, fd, grad[i, k])


@pytest.mark.parametrize("online", [True, False])
def test_control_point_gradient_matches_finite_differences(oracle_mod, online):
    # control points (three uncertainty forms) and, on path A, laser planes: the adjoint adds their exact Huber
    # weight matrices to H, and dl/dP = J^T M J lambda_X; pinned by central differences of re-solves
    O = oracle_mod
    p0 = S.ring_problem(seed=11, num_cameras=8, num_points=60, max_views=5)
    p = S.add_survey_constraints(p0, seed=5, num_control=4, num_planes=0 if online else 6, kinds=(0, 1, 2),
                                 control_sigma=0.02)
    opts = PO.SolveOptions()
    opts.solver.maxIterations = 80
    opts.solver.useReferenceOnlineSchur = online
    opts.backend.stepRmsTolerance = 1e-13
    copts = PO.to_c(opts)
    st, prep, masks, rep = _solve(O, p, copts)
    assert rep["online_schur_used"] == (1 if online else 0)
    rng = np.random.default_rng(4)
    T = st["cam_center"] + rng.normal(0.0, 0.05, st["cam_center"].shape)
    Y = st["track_point"] + rng.normal(0.0, 0.05, st["track_point"].shape)
    dl_dC = 2.0 * (st["cam_center"] - T)
    dl_dX = 2.0 * (st["track_point"] - Y)
    grad, dctl = adjoint_reference(O, p, st, masks, prep, copts, dl_dC, dl_dX, with_control=True)

    def loss_at(q):
        s2, *_ = _solve(O, q, copts)
        return float(np.sum((s2["cam_center"] - T) ** 2) + np.sum((s2["track_point"] - Y) ** 2))

    h = 1e-4
    scale = np.abs(dctl).max()
    assert scale > 0
    for i in range(p.num_control_points):
        for k in range(3):
            qp, qm = p.copy(), p.copy()
            qp.cp_point[i, k] += h
            qm.cp_point[i, k] -= h
            fd = (loss_at(qp) - loss_at(qm)) / (2 * h)
            assert abs(fd - dctl[i, k]) <= 2e-3 * scale, (i, k, fd, dctl[i, k])
    # observation gradients of the constrained problem
    for i in rng.choice(p.num_observations, size=3, replace=False):
        qp, qm = p.copy(), p.copy()
        qp.obs_uv[i, 0] += 1e-3
        qm.obs_uv[i, 0] -= 1e-3
        fd = (loss_at(qp) - loss_at(qm)) / 2e-3
        assert abs(fd - grad[i, 0]) <= 2e-3 * np.abs(grad).max(), (i, fd, grad[i, 0])
