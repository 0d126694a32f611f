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
