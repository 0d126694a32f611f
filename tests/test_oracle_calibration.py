"""Pins the oracle's full frame calibration (all 13 slots of a Metashape-convention group: f, aspect, cx, cy,
k1..k4, p1..p4, skew b2) and the reference calibration transition prior (BundleAdjustPlaMatrixModel.cpp:173-225).

  * the reduced system with 13 refined slots against a dense Schur complement assembled in numpy from the
    oracle's per-observation Jacobians (themselves pinned by central differences) and the group priors;
  * the analytic directional decrease against central differences of the objective along the step;
  * the transition weights against their closed forms on a 4000 x 3000 px, f = 3000 px group (normalised
    half-diagonal r = sqrt((2000/3000)^2 + (1500/3000)^2)), the unchanged-mask case (every weight but the
    aspect's zero) and the image-size fallback (twice the reference principal point).
"""
import numpy as np
import pytest

from plabundle_amd import options as PO
from plabundle_amd import synthetic as S

P13 = 13
ALL = [True] * 13


def _opts(all_slots=True, transition=False, previous=None):
    o = S.ring_options(PO.SolveOptions(), max_iterations=12)
    c = o.calibration
    c.refineSharedFocalLength = True
    c.refineSharedFocalAspectRatio = all_slots
    c.refineSharedPrincipalPoint = all_slots
    c.refineSharedRadialDistortion = True
    c.refineSharedHighOrderDistortion = True
    c.refineSharedMetashapeParameters = all_slots
    c.useReferenceCalibrationTransitionPrior = transition
    if previous is not None:
        c.referencePreviousIntrinsicParameterMask = list(previous)
    return o


def _dense(O, p, copts, orc):
    prep = orc.prepare(copts)
    assert prep["prepared"]
    masks = orc.masks()
    groups = orc.groups()
    P, Ne = prep["primary_block_count"], prep["track_block_count"]
    n = P * P13 + 3 * Ne
    H = np.zeros((n, n))
    g = np.zeros(n)
    conv = p.cam_convention if p.cam_convention is not None else np.zeros(p.num_cameras, dtype=np.int32)
    for t in range(p.num_tracks):
        if not masks["active_track"][t]:
            continue
        for i in range(p.track_offset[t], p.track_offset[t + 1]):
            c = p.obs_camera[i]
            ib = masks["intrinsic_block_by_camera"][c]
            lin = O.linearize(p.cam_rotation[c], p.cam_center[c], p.cam_intrinsics[c], int(conv[c]),
                              p.track_point[t], p.obs_uv[i, 0], p.obs_uv[i, 1], p.obs_weight[i], p.obs_scale[i],
                              copts.robust_loss, copts.robust_loss_scale, True,
                              int(groups["enabled"][masks["calibration_group_by_camera"][c]]) if ib >= 0 else 0)
            J = np.zeros((2, n))
            cb = masks["camera_block"][c]
            if cb >= 0:
                J[:, cb * P13: cb * P13 + 6] = lin["Jpose"]
            if ib >= 0:
                J[:, ib * P13: ib * P13 + P13] = lin["Jint"]
            tb = masks["track_block"][t]
            if tb >= 0:
                J[:, P * P13 + 3 * tb: P * P13 + 3 * tb + 3] = lin["Jp"]
            H += lin["weight"] * J.T @ J
            g += lin["weight"] * J.T @ lin["residual"]
    for gi in range(prep["intrinsic_block_count"]):
        blk = prep["camera_block_count"] + gi
        for q in range(13):
            if (int(groups["enabled"][gi]) >> q) & 1:
                w = groups["inv_sigma"][gi, q]
                H[blk * P13 + q, blk * P13 + q] += w * w
                g[blk * P13 + q] += w * w * (groups["params"][gi, q] - groups["prior"][gi, q])
    return H, g, P, Ne


def test_thirteen_slot_reduced_system_matches_dense_schur(oracle_mod):
    p = S.ring_problem(seed=8, num_cameras=9, num_points=60, max_views=5, metashape=True)
    copts = PO.to_c(_opts())
    assert copts.shared_intrinsic_mask == 0x1FFF
    orc = oracle_mod.Oracle(p, threads=2)
    H, g, P, Ne = _dense(oracle_mod, p, copts, orc)
    assert orc.groups()["enabled"][0] == 0x1FFF  # Metashape cameras support every slot
    m = P * P13
    for damping in (0.0, 1e-2):
        Hcc, Hcp, Hpp = H[:m, :m], H[:m, m:], H[m:, m:].copy()
        Hinv = np.zeros_like(Hpp)
        for k in range(Ne):
            blk = Hpp[3 * k:3 * k + 3, 3 * k:3 * k + 3].copy()
            blk[np.diag_indices(3)] *= 1.0 + damping
            Hinv[3 * k:3 * k + 3, 3 * k:3 * k + 3] = np.linalg.inv(blk)
        Sm = Hcc - Hcp @ Hinv @ Hcp.T
        rhs = -(g[:m] - Hcp @ Hinv @ g[m:])
        d = np.diag(Sm).copy()
        Sm[np.diag_indices(m)] = np.where(d == 0.0, 1.0, d * (1.0 + damping))
        A, b, _, singular = orc.reduced_system(P, damping)
        assert not singular
        np.testing.assert_allclose(A, Sm, atol=1e-10 * np.abs(Sm).max(), rtol=0)
        np.testing.assert_allclose(b, rhs, atol=1e-10 * np.abs(rhs).max(), rtol=0)


def test_thirteen_slot_directional_decrease_and_solve(oracle_mod):
    p = S.ring_problem(seed=8, num_cameras=12, num_points=200, max_views=6, metashape=True)
    copts = PO.to_c(_opts())
    orc = oracle_mod.Oracle(p, threads=2)
    prep = orc.prepare(copts)
    primary, elim, dd, _, _ = orc.compute_step(prep["primary_block_count"], prep["track_block_count"], 0.0)
    h = 1e-5
    fd = (orc.cost_after_step(primary, elim, h) - orc.cost_after_step(primary, elim, -h)) / (2 * h)
    assert abs(fd + dd) <= 1e-5 * abs(dd)
    o = _opts()
    o.solver.maxIterations = 25  # the weakly observed slots converge linearly on this ring (17 iterations)
    rep = orc.solve(PO.to_c(o))
    assert rep["solve_status"] == 1 and rep["final_cost"] < 0.01 * rep["initial_cost"]


def _expected_transition(enabled, previous, f=3000.0, W=4000.0, H=3000.0, aspect=1.0):
    r = np.hypot(0.5 * W / f, 0.5 * H / (f * aspect))
    base = np.array([1 / r, 0, 1, 1, 1 / (r**3 * f), 1 / (r**5 * f), 1 / (r**7 * f), 1 / (r**2 * f),
                     1 / (r**2 * f), 1 / r, 1 / (r**9 * f), 1, 1])
    changed = any(enabled[i] != previous[i] for i in range(13) if i != 1)
    w = np.zeros(13)
    if changed:
        for i in range(13):
            if i != 1 and enabled[i]:
                w[i] = 1.0 / (base[i] * (1e6 if previous[i] else 0.01))
    return w


def test_transition_prior_weights(oracle_mod):
    p = S.ring_problem(seed=8, num_cameras=9, num_points=60, max_views=5, metashape=True)
    previous = [True, True, False, False, True, False, False, False, False, False, False, False, False]
    copts = PO.to_c(_opts(transition=True, previous=previous))
    orc = oracle_mod.Oracle(p, threads=1)
    assert orc.prepare(copts)["prepared"]
    gr = orc.groups()
    enabled = [True] * 13
    expected = _expected_transition(enabled, previous, aspect=np.exp(p.cam_intrinsics[0, 1]))
    np.testing.assert_allclose(np.delete(gr["inv_sigma"][0], 1), np.delete(expected, 1), rtol=1e-12)
    # the aspect keeps its weak prior 1 / aspect sigma
    assert gr["inv_sigma"][0, 1] == pytest.approx(1.0 / copts.aspect_prior_sigma)
    # unchanged refined set: every transition weight is zero
    same = PO.to_c(_opts(transition=True, previous=ALL))
    orc2 = oracle_mod.Oracle(p, threads=1)
    orc2.prepare(same)
    w = orc2.groups()["inv_sigma"][0]
    assert np.all(np.delete(w, 1) == 0.0) and w[1] > 0.0
    # without image sizes the half-diagonal comes from twice the reference principal point (here the same)
    q = p.copy()
    q.cam_image_size = None
    orc3 = oracle_mod.Oracle(q, threads=1)
    orc3.prepare(copts)
    np.testing.assert_allclose(orc3.groups()["inv_sigma"][0], gr["inv_sigma"][0], rtol=1e-12)
    q.ref_intrinsics = q.cam_intrinsics.copy()
    q.ref_intrinsics[:, 2] = 1000.0  # reference cx -> width 2000 px
    orc4 = oracle_mod.Oracle(q, threads=1)
    orc4.prepare(copts)
    exp4 = _expected_transition(enabled, previous, W=2000.0, aspect=np.exp(p.cam_intrinsics[0, 1]))
    np.testing.assert_allclose(np.delete(orc4.groups()["inv_sigma"][0], 1), np.delete(exp4, 1), rtol=1e-12)


def test_transition_prior_objective(oracle_mod):
    # the prior terms use the transition weights: the objective changes by 1/2 sum (w (theta - prior))^2
    p = S.ring_problem(seed=8, num_cameras=9, num_points=60, max_views=5, metashape=True)
    q = p.copy()
    q.cam_intrinsics = q.cam_intrinsics.copy()
    q.cam_intrinsics[:, 0] *= 1.01      # initial focal away from the reference focal
    q.cam_intrinsics[:, 4] += 2e-3      # k1 away from the reference k1
    q.ref_intrinsics = p.cam_intrinsics.copy()
    previous = [True] + [False] * 12
    copts = PO.to_c(_opts(transition=True, previous=previous))
    orc = oracle_mod.Oracle(q, threads=1)
    orc.prepare(copts)
    gr = orc.groups()
    off = PO.to_c(_opts(transition=True, previous=ALL))  # all transition weights zero
    orc0 = oracle_mod.Oracle(q, threads=1)
    orc0.prepare(off)
    g0 = orc0.groups()
    w = gr["inv_sigma"][0]
    res = w * (gr["params"][0] - gr["prior"][0])
    res0 = g0["inv_sigma"][0] * (g0["params"][0] - g0["prior"][0])
    assert abs(res[0]) > 0 and abs(res[4]) > 0
    np.testing.assert_allclose(orc.initial_cost() - orc0.initial_cost(), 0.5 * (res @ res - res0 @ res0),
                               rtol=1e-9)


def test_reduced_solve_with_zero_entries_in_intrinsic_blocks(oracle_mod):
    # eight 13-slot groups on a survey block: some entries of the intrinsic blocks are exactly zero; the envelope
    # Cholesky of the reduced solve must leave them out of its profile (a store of a zero outside the profile
    # once overwrote a neighbouring row and made the factorisation fail)
    p = S.large_aerial_problem(strips=3, cameras_per_strip=24, num_points=6000, num_control=20, num_planes=200)
    copts = PO.to_c(S.large_aerial_options(PO.SolveOptions()))
    orc = oracle_mod.Oracle(p, threads=4)
    prep = orc.prepare(copts)
    assert prep["prepared"] and prep["intrinsic_block_count"] == 8
    P, Ne = prep["primary_block_count"], prep["track_block_count"]
    for damping in (0.0, 1e-3):
        A, b, _, _ = orc.reduced_system(P, damping)
        assert (A == 0.0).sum() > 0
        primary, _, _, _, _ = orc.compute_step(P, Ne, damping)
        assert np.linalg.norm(A @ primary - b) <= 1e-10 * np.linalg.norm(b)
