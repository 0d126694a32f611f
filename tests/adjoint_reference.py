"""Dense implicit-function reference of the adjoint (test infrastructure).

At a converged state theta*(y) the optimality condition g(theta, y) = sum_i J_i^T w_n r_i = 0 holds; differentiating
it gives H dtheta/dy_i = J_i^T M_i with H = sum_i J_i^T M_i J_i and M_i = d(w_n r)/dr = w_n I + 4 w^2 rho''(s) r r^T
(w = weight / scale^2, s = w |r|^2, w_n = 2 w rho'(s)). For a loss l(theta*): dl/dy_i = M_i J_i lambda with
H lambda = dl/dtheta. Everything here is dense numpy over the oracle's per-observation Jacobians (reference
parameterisation: pose rows [rotation 3, centre 3], points stepped as X + s dp with s = |X|/1000 / atan(|X|/1000))."""
import numpy as np

P13 = 13


def robust_curvature(s, loss, d):
    if loss == 1:  # Huber: 1/2 s inside, d sqrt(s) - d^2/2 outside
        return 0.0 if s <= d * d else -0.25 * d / s**1.5
    if loss == 2:  # Cauchy: d^2/2 log(1 + s/d^2)
        return -0.5 / (d * d * (1.0 + s / (d * d)) ** 2)
    return 0.0


def weight_matrix(r, w_obs, w_n, loss, d):
    n2 = float(r @ r)
    if n2 == 0.0:
        return w_n * np.eye(2)
    c = 4.0 * w_obs**2 * robust_curvature(w_obs * n2, loss, d)
    mu = max(0.0, w_n + c * n2)
    P = np.outer(r, r) / n2
    return w_n * (np.eye(2) - P) + mu * P


def constraint_rows(O, problem, copts):
    """Survey constraints as (track, J (rows x 3), target, delta, control index or -1): J = sqrt(w) A for a control
    point, sqrt(w) n^T for a laser plane; delta is the Huber threshold (BundleAdjustPlaMatrixConstraints.cpp)."""
    out = []
    for i in range(problem.num_laser_planes):
        sw = np.sqrt(copts.laser_plane_weight * problem.lp_weight[i])
        out.append((int(problem.lp_track[i]), sw * problem.lp_normal[i][None, :], problem.lp_point[i],
                    copts.laser_huber_delta * sw, -1))
    for i in range(problem.num_control_points):
        A = O.control_whitening(int(problem.cp_uncertainty[i]), float(problem.cp_sigma[i]), problem.cp_matrix[i])
        sw = np.sqrt(copts.control_point_weight * problem.cp_weight[i])
        out.append((int(problem.cp_track[i]), sw * A, problem.cp_point[i], copts.control_point_huber_delta, i))
    return out


def constraint_weight_matrix(r, delta):
    """d(w r)/dr for Huber on s = |r|^2 (weight w = 1 inside, delta / sqrt(s) outside), along-r eigenvalue >= 0."""
    s = float(r @ r)
    w = 1.0 if s <= delta * delta else delta / np.sqrt(s)
    if s == 0.0:
        return w * np.eye(len(r))
    mu = max(0.0, w + 4.0 * robust_curvature(s, 1, delta) * s)
    Pm = np.outer(r, r) / s
    return w * (np.eye(len(r)) - Pm) + mu * Pm


def adjoint_reference(O, problem, state, masks, prep, copts, dl_dC, dl_dX=None, priors=None, with_control=False,
                      stats=None):
    """dl/d(u, v) per observation ([num_observations, 2]) at `state` (dict of the oracle / workspace download);
    with_control=True also returns dl/dP of every control point ([num_control_points, 3]). Problems with survey
    constraints use the default parameterisation (world-frame rotation, unscaled point step).

    H lambda = dl/dtheta is solved through the point blocks: H = [[Hcc, W], [W^T, blockdiag(Hpp)]] is assembled
    as its primary part (dense), one primary x 3 cross block and one 3 x 3 block per point, and reduced to the
    primary system S = Hcc - sum W Hpp^-1 W^T (the same solution as the dense solve of H, without forming it)."""
    ref_param = bool(copts.use_online_schur) and problem.num_control_points == 0 and problem.num_laser_planes == 0
    P = prep["primary_block_count"]
    Ne = prep["track_block_count"]
    PP = P * P13
    mask = copts.shared_intrinsic_mask
    ids, Jcs, Jps, Ms, tbs = [], [], [], [], []
    for t in range(problem.num_tracks):
        if not masks["active_track"][t]:
            continue
        X = state["track_point"][t]
        tb = masks["track_block"][t]
        for i in range(problem.track_offset[t], problem.track_offset[t + 1]):
            c = problem.obs_camera[i]
            w_obs = problem.obs_weight[i] / problem.obs_scale[i] ** 2
            if not w_obs > 0:
                continue
            lin = O.linearize(state["cam_rotation"][c], state["cam_center"][c], state["cam_intrinsics"][c], 0, X,
                              problem.obs_uv[i, 0], problem.obs_uv[i, 1], problem.obs_weight[i],
                              problem.obs_scale[i], copts.robust_loss, copts.robust_loss_scale, ref_param,
                              mask if masks["intrinsic_block_by_camera"][c] >= 0 else 0)
            assert lin is not None
            Jc = np.zeros((2, PP))
            cb = masks["camera_block"][c]
            if cb >= 0:
                Jc[:, cb * P13: cb * P13 + 6] = lin["Jpose"]
            ib = masks["intrinsic_block_by_camera"][c]
            if ib >= 0:
                Jc[:, ib * P13: ib * P13 + P13] = lin["Jint"]
            ids.append(i)
            Jcs.append(Jc)
            Jps.append(lin["Jp"] if tb >= 0 else np.zeros((2, 3)))
            Ms.append(weight_matrix(lin["residual"], w_obs, lin["weight"], copts.robust_loss, copts.robust_loss_scale))
            tbs.append(tb)
    Jc = np.array(Jcs).reshape(-1, 2, PP)
    Jp = np.array(Jps).reshape(-1, 2, 3)
    M = np.array(Ms).reshape(-1, 2, 2)
    tb = np.array(tbs, dtype=np.int64)
    MJc, MJp = M @ Jc, M @ Jp
    Hcc = np.tensordot(Jc, MJc, axes=([0, 1], [0, 1]))
    on = tb >= 0
    W = np.zeros((max(Ne, 1), PP, 3))
    Hpp = np.zeros((max(Ne, 1), 3, 3))
    np.add.at(W, tb[on], np.einsum("nki,nkj->nij", Jc[on], MJp[on]))
    np.add.at(Hpp, tb[on], np.einsum("nki,nkj->nij", Jp[on], MJp[on]))
    cons = []
    for t, Jk, target, delta, index in constraint_rows(O, problem, copts):
        b = masks["track_block"][t]
        if not masks["active_track"][t]:
            continue
        r = Jk @ (state["track_point"][t] - target)
        Mk = constraint_weight_matrix(r, delta)
        if b >= 0:
            Hpp[b] += Jk.T @ Mk @ Jk
        cons.append((b, Jk, Mk, index))
    for idx, h in (priors or {}).items():
        Hcc[idx, idx] += h
    gc = np.zeros(PP)
    for c in range(problem.num_cameras):
        cb = masks["camera_block"][c]
        if cb >= 0:
            gc[cb * P13 + 3: cb * P13 + 6] = dl_dC[c]
    gp = np.zeros((max(Ne, 1), 3))
    if dl_dX is not None:
        from oracle import oracle as Oo

        for t in range(problem.num_tracks):
            b = masks["track_block"][t]
            if b >= 0:
                sc = Oo.point_scale(state["track_point"][t]) if ref_param else 1.0
                gp[b] = sc * dl_dX[t]
    # a point block that the 3x3 cofactor inverse rejects (|det| <= 1e-15 max|m|^3, rank-deficient weight matrices
    # of observations on the clamped branch of the loss) gets a ridge of 1e-9 of its largest diagonal entry, the
    # rule of the GPU's adjoint linearisation
    Hb = Hpp[:Ne].copy()
    if Ne:
        mx = np.abs(Hb).reshape(Ne, 9).max(axis=1)
        det = np.linalg.det(Hb)
        sing = ~(np.abs(det) > 1e-15 * mx**3)
        if np.any(sing):
            ridge = 1e-9 * np.abs(np.diagonal(Hb[sing], axis1=1, axis2=2)).max(axis=1)
            Hb[sing] += ridge[:, None, None] * np.eye(3)[None]
        if stats is not None:
            stats["ridged_point_blocks"] = int(np.count_nonzero(sing))
    Hinv = np.linalg.inv(Hb) if Ne else np.zeros((0, 3, 3))
    V = W[:Ne] @ Hinv                                  # W Hpp^-1 per point
    S = Hcc - np.tensordot(V, W[:Ne], axes=([0, 2], [0, 2]))
    rhs = gc - np.einsum("tij,tj->i", V, gp[:Ne])
    keep = np.diag(Hcc) != 0.0
    lam_c = np.zeros(PP)
    lam_c[keep] = np.linalg.solve(S[np.ix_(keep, keep)], rhs[keep])
    lam_p = np.einsum("tij,tj->ti", Hinv, gp[:Ne] - np.einsum("tij,i->tj", W[:Ne], lam_c))
    out = np.zeros((problem.num_observations, 2))
    lp = np.zeros((len(tb), 3))
    lp[on] = lam_p[tb[on]]
    vals = np.einsum("nij,nj->ni", M, Jc @ lam_c + np.einsum("nkj,nj->nk", Jp, lp))
    out[np.array(ids, dtype=np.int64)] = vals
    if not with_control:
        return out
    dctl = np.zeros((problem.num_control_points, 3))
    for b, Jk, Mk, index in cons:
        if index >= 0 and b >= 0:
            dctl[index] = Jk.T @ (Mk @ (Jk @ lam_p[b]))
    return out, dctl
