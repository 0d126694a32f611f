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


def adjoint_reference(O, problem, state, masks, prep, copts, dl_dC, dl_dX=None, priors=None):
    """dl/d(u, v) per observation ([num_observations, 2]) at `state` (dict of the oracle / workspace download)."""
    P = prep["primary_block_count"]
    Ne = prep["track_block_count"]
    n = P * P13 + 3 * Ne
    H = np.zeros((n, n))
    rows = []
    mask = copts.shared_intrinsic_mask
    for t in range(problem.num_tracks):
        if not masks["active_track"][t]:
            continue
        X = state["track_point"][t]
        for i in range(problem.track_offset[t], problem.track_offset[t + 1]):
            c = problem.obs_camera[i]
            w_obs = problem.obs_weight[i] / problem.obs_scale[i] ** 2
            if not w_obs > 0:
                continue
            lin = O.linearize(state["cam_rotation"][c], state["cam_center"][c], state["cam_intrinsics"][c], 0, X,
                              problem.obs_uv[i, 0], problem.obs_uv[i, 1], problem.obs_weight[i],
                              problem.obs_scale[i], copts.robust_loss, copts.robust_loss_scale, True,
                              mask if masks["intrinsic_block_by_camera"][c] >= 0 else 0)
            assert lin is not None
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
            M = weight_matrix(lin["residual"], w_obs, lin["weight"], copts.robust_loss, copts.robust_loss_scale)
            H += J.T @ M @ J
            rows.append((i, J, M))
    for idx, h in (priors or {}).items():
        H[idx, idx] += h
    g = np.zeros(n)
    for c in range(problem.num_cameras):
        cb = masks["camera_block"][c]
        if cb >= 0:
            g[cb * P13 + 3: cb * P13 + 6] = dl_dC[c]
    if dl_dX is not None:
        from oracle import oracle as Oo

        for t in range(problem.num_tracks):
            tb = masks["track_block"][t]
            if tb >= 0:
                g[P * P13 + 3 * tb: P * P13 + 3 * tb + 3] = Oo.point_scale(state["track_point"][t]) * dl_dX[t]
    keep = np.diag(H) != 0.0
    lam = np.zeros(n)
    lam[keep] = np.linalg.solve(H[np.ix_(keep, keep)], g[keep])
    out = np.zeros((problem.num_observations, 2))
    for i, J, M in rows:
        out[i] = M @ (J @ lam)
    return out
