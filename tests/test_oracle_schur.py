"""Pins the oracle's online point Schur (A6-A10) against dense linear algebra.

The full normal equations are assembled here in numpy from the oracle's
per-observation linearisations and the block numbering it reports; the dense
Schur complement, dense solves and central differences of the objective are
then compared with the oracle's reduced system, step and directional decrease.
"""
import numpy as np
import pytest

from plabundle_amd import options as PO
from plabundle_amd import synthetic as S

P13 = 13


def _problem(intrinsics: bool):
    p = S.ring_problem(seed=7, num_cameras=8, num_points=40, max_views=5)
    opts = PO.SolveOptions()
    opts.solver.imageRobustLoss = PO.ImageRobustLoss.Huber
    opts.solver.useReferenceOnlineSchur = True
    if intrinsics:
        opts.calibration.refineSharedFocalLength = True
        opts.calibration.refineSharedRadialDistortion = True
    return p, PO.to_c(opts)


def _full_system(O, p, copts, oracle):
    prep = oracle.prepare(copts)
    assert prep["prepared"]
    masks = oracle.masks()
    P = prep["primary_block_count"]
    Ne = prep["track_block_count"]
    n = P * P13 + 3 * Ne
    H = np.zeros((n, n))
    g = np.zeros(n)
    mask = copts.shared_intrinsic_mask
    for t in range(p.num_tracks):
        if not masks["active_track"][t]:
            continue
        X = p.track_point[t]
        for i in range(p.track_offset[t], p.track_offset[t + 1]):
            c = p.obs_camera[i]
            lin = O.linearize(p.cam_rotation[c], p.cam_center[c], p.cam_intrinsics[c], 0, X, p.obs_uv[i, 0],
                              p.obs_uv[i, 1], p.obs_weight[i], p.obs_scale[i], copts.robust_loss,
                              copts.robust_loss_scale, True, mask if masks["intrinsic_block_by_camera"][c] >= 0 else 0)
            J = np.zeros((2, n))
            cb = masks["camera_block"][c]
            if cb >= 0:
                J[:, cb * P13: cb * P13 + 6] = lin["Jpose"]
            ib = masks["intrinsic_block_by_camera"][c]
            if ib >= 0:
                J[:, ib * P13: ib * P13 + P13] = lin["Jint"]
            tb = masks["track_block"][t]
            if tb >= 0:
                J[:, P * P13 + 3 * tb: P * P13 + 3 * tb + 3] = lin["Jp"]
            H += lin["weight"] * J.T @ J
            g += lin["weight"] * J.T @ lin["residual"]
    # intrinsic priors (BundleAdjustPlaMatrixModel.cpp:150-171): identical cameras -> value == prior
    if mask:
        f = p.cam_intrinsics[0, 0]
        sigma = {0: copts.focal_prior_sigma * f, 4: copts.prior_sigma[0], 5: copts.prior_sigma[1],
                 6: copts.prior_sigma[2], 7: copts.prior_sigma[3], 8: copts.prior_sigma[4]}
        for gi in range(prep["intrinsic_block_count"]):
            blk = prep["camera_block_count"] + gi
            for q, sg in sigma.items():
                if (mask >> q) & 1:
                    H[blk * P13 + q, blk * P13 + q] += 1.0 / sg**2
    return H, g, P, Ne, prep


def _reduce(H, g, P, damping):
    m = P * P13
    Hcc, Hcp, Hpp = H[:m, :m], H[:m, m:], H[m:, m:].copy()
    gc, gp = g[:m], g[m:]
    Ne = Hpp.shape[0] // 3
    Hinv = np.zeros_like(Hpp)
    for k in range(Ne):
        blk = Hpp[3 * k:3 * k + 3, 3 * k:3 * k + 3].copy()
        blk[np.diag_indices(3)] *= 1.0 + damping
        Hinv[3 * k:3 * k + 3, 3 * k:3 * k + 3] = np.linalg.inv(blk)
    Sm = Hcc - Hcp @ Hinv @ Hcp.T
    rhs = -(gc - Hcp @ Hinv @ gp)
    d = np.diag(Sm).copy()
    d = np.where(d == 0.0, 1.0, d * (1.0 + damping))
    Sm[np.diag_indices(m)] = d
    return Sm, rhs, Hinv


@pytest.mark.parametrize("intrinsics", [False, True])
@pytest.mark.parametrize("damping", [0.0, 1e-2])
def test_reduced_system_matches_dense_schur(oracle_mod, intrinsics, damping):
    p, copts = _problem(intrinsics)
    orc = oracle_mod.Oracle(p, threads=3)
    H, g, P, Ne, _ = _full_system(oracle_mod, p, copts, orc)
    Sm, rhs, _ = _reduce(H, g, P, damping)
    A, b, cost, singular = orc.reduced_system(P, damping)
    assert not singular
    scale = np.abs(Sm).max()
    np.testing.assert_allclose(A, Sm, atol=1e-10 * scale, rtol=0)
    np.testing.assert_allclose(b, rhs, atol=1e-10 * np.abs(rhs).max(), rtol=0)
    assert abs(cost - orc.initial_cost()) <= 1e-12 * cost


@pytest.mark.parametrize("intrinsics", [False, True])
def test_step_matches_dense_full_solve(oracle_mod, intrinsics):
    p, copts = _problem(intrinsics)
    orc = oracle_mod.Oracle(p, threads=2)
    H, g, P, Ne, _ = _full_system(oracle_mod, p, copts, orc)
    m = P * P13
    Hf = H.copy()
    zero = np.diag(Hf) == 0.0
    Hf[zero, zero] = 1.0
    delta = np.linalg.solve(Hf, -g)
    primary, elim, dd, rg, rp = orc.compute_step(P, Ne, 0.0)
    np.testing.assert_allclose(primary, delta[:m], atol=1e-8 * np.abs(delta).max())
    np.testing.assert_allclose(elim, delta[m:], atol=1e-8 * np.abs(delta).max())
    # directional decrease = -g . delta (ReferenceSchur.cpp:1004-1007,1060-1062)
    assert abs(dd - (-g @ delta)) <= 1e-8 * abs(dd)
    # step RMS over pose/intrinsic and point entries (BundleAdjustPlaMatrix.cpp:77-158)
    cams = [b for b in range(P) if b < P - (1 if intrinsics else 0)]
    glob = [primary[b * P13: b * P13 + 6] for b in cams]
    if intrinsics:
        on = [q for q in range(13) if (copts.shared_intrinsic_mask >> q) & 1]
        glob.append(primary[(P - 1) * P13 + np.array(on)])
    gv = np.concatenate(glob)
    assert abs(rg - np.sqrt(np.mean(gv**2))) <= 1e-12 * rg
    assert abs(rp - np.sqrt(np.mean(elim**2))) <= 1e-12 * rp


@pytest.mark.parametrize("damping", [0.0, 1e-3])
def test_damped_step_is_reduced_solve_plus_back_substitution(oracle_mod, damping):
    p, copts = _problem(True)
    orc = oracle_mod.Oracle(p, threads=1)
    H, g, P, Ne, _ = _full_system(oracle_mod, p, copts, orc)
    m = P * P13
    Sm, rhs, Hinv = _reduce(H, g, P, damping)
    dc = np.linalg.solve(Sm, rhs)
    dp = Hinv @ (-g[m:] - H[m:, :m] @ dc)
    primary, elim, dd, _, _ = orc.compute_step(P, Ne, damping)
    np.testing.assert_allclose(primary, dc, atol=1e-8 * np.abs(dc).max())
    np.testing.assert_allclose(elim, dp, atol=1e-8 * np.abs(dp).max())


def test_directional_derivative_matches_objective(oracle_mod):
    # d/dalpha F(retract(state, alpha * step)) at 0 equals -dd: the retraction and point scale are
    # consistent with the Jacobians the step was computed from.
    p, copts = _problem(True)
    orc = oracle_mod.Oracle(p, threads=2)
    prep = orc.prepare(copts)
    P, Ne = prep["primary_block_count"], prep["track_block_count"]
    primary, elim, dd, _, _ = orc.compute_step(P, Ne, 0.0)
    h = 1e-5
    fd = (orc.cost_after_step(primary, elim, h) - orc.cost_after_step(primary, elim, -h)) / (2 * h)
    assert abs(fd + dd) <= 1e-5 * abs(dd)
    assert orc.cost_after_step(primary, elim, 1.0) < orc.initial_cost()
