"""Seeded synthetic frame-BA problems.

Workloads (sizes from the north-star configurations):
  * ``native_three_cameras`` - the reference's own online-Schur test scene
    (test/test_native_problem.cpp:51-94): three unit cameras, 24 tracks.
  * ``ring_problem`` - 50 pinhole cameras on a 10 m circle looking inward,
    4000x3000 px, f = 3000 px, one shared Brown group, 2000 points in a 4 m
    cube, <= 8 views per point, 0.5 px noise, Huber 1 px (seed 1).
  * ``aerial_problem`` - nadir aerial block: cameras in 3 strips at 300 m,
    80 % forward / 60 % side overlap, 4000x3000 px, f = 3000 px, 4 shared
    intrinsic groups, points on an fBm heightfield (+-20 m), geometric track
    lengths (min 2, mean 7.5 after the visibility cap), 0.5 px noise, 1 %
    outliers, Cauchy 1 px (seed 2; 1000 cameras / 200k points / ~1.5M obs at
    full size).

Random numbers come from numpy's PCG64 generator with the stated seed; the
same numpy build (this image) reproduces every problem bit for bit.
"""
from __future__ import annotations

import math

import numpy as np

from .problem import GaugePolicy, Problem, intrinsic_block


# --- small geometry helpers (numpy, float64) -------------------------------------------------
def exp_so3(w: np.ndarray) -> np.ndarray:
    """Rodrigues rotation matrices for axis-angle rows w (..., 3) -> (..., 3, 3)."""
    w = np.asarray(w, dtype=np.float64)
    th = np.linalg.norm(w, axis=-1)[..., None, None]
    K = np.zeros(w.shape[:-1] + (3, 3))
    K[..., 0, 1] = -w[..., 2]
    K[..., 0, 2] = w[..., 1]
    K[..., 1, 0] = w[..., 2]
    K[..., 1, 2] = -w[..., 0]
    K[..., 2, 0] = -w[..., 1]
    K[..., 2, 1] = w[..., 0]
    small = th < 1e-10
    th_safe = np.where(small, 1.0, th)
    A = np.where(small, 1.0, np.sin(th_safe) / th_safe)
    B = np.where(small, 0.5, (1.0 - np.cos(th_safe)) / th_safe**2)
    eye = np.broadcast_to(np.eye(3), K.shape)
    return eye + A * K + B * (K @ K)


def project_points(R: np.ndarray, C: np.ndarray, intr: np.ndarray, X: np.ndarray, metashape=None):
    """Projects world points X (n,3) through cameras (R (n,3,3) c2w, C (n,3), intr (n,13)).

    Returns (uv (n,2), depth (n,)). Same frame-camera model as the C ABI documents."""
    d = X - C
    xc = np.einsum("nij,ni->nj", R, d)
    z = xc[:, 2]
    x = xc[:, 0] / z
    y = xc[:, 1] / z
    fx = intr[:, 0]
    fy = fx * np.exp(intr[:, 1])
    k1, k2, k3, p1, p2, b2, k4, p3, p4 = (intr[:, i] for i in range(4, 13))
    r2 = x * x + y * y
    rad = 1 + r2 * (k1 + r2 * (k2 + r2 * (k3 + r2 * k4)))
    if metashape is None:
        metashape = np.zeros(len(x), dtype=bool)
    tx_o = 2 * p1 * x * y + p2 * (r2 + 2 * x * x)
    ty_o = p1 * (r2 + 2 * y * y) + 2 * p2 * x * y
    T = 1 + p3 * r2 + p4 * r2 * r2
    tx_m = (p1 * (r2 + 2 * x * x) + 2 * p2 * x * y) * T
    ty_m = (p2 * (r2 + 2 * y * y) + 2 * p1 * x * y) * T
    tx = np.where(metashape, tx_m, tx_o)
    ty = np.where(metashape, ty_m, ty_o)
    xd = x * rad + tx
    yd = y * rad + ty
    u = fx * xd + b2 * yd + intr[:, 2]
    v = fy * yd + intr[:, 3]
    return np.stack([u, v], axis=1), z


def look_at_rotation(center, target, up=(0.0, 0.0, 1.0)) -> np.ndarray:
    """Camera-to-world rotation whose columns are the camera right/down/forward axes."""
    f = np.asarray(target, float) - np.asarray(center, float)
    f /= np.linalg.norm(f)
    r = np.cross(f, up)
    r /= np.linalg.norm(r)
    dn = np.cross(f, r)
    return np.stack([r, dn, f], axis=1)


def _finish(rot, cen, intr, pts, track_obs, **kwargs) -> Problem:
    off = np.zeros(len(track_obs) + 1, dtype=np.int64)
    off[1:] = np.cumsum([len(t[0]) for t in track_obs])
    cam = np.concatenate([t[0] for t in track_obs]).astype(np.int32)
    uv = np.concatenate([t[1] for t in track_obs]).reshape(-1, 2)
    return Problem(
        cam_rotation=np.ascontiguousarray(rot.reshape(-1, 9)),
        cam_center=np.ascontiguousarray(cen),
        cam_intrinsics=np.ascontiguousarray(intr),
        track_point=np.ascontiguousarray(pts),
        track_offset=off,
        obs_camera=cam,
        obs_uv=np.ascontiguousarray(uv),
        obs_weight=np.ones(len(cam)),
        obs_scale=np.ones(len(cam)),
        **kwargs,
    )


# --- the reference's own online-Schur test scene ---------------------------------------------
def native_three_cameras() -> Problem:
    """test/test_native_problem.cpp:51-94: three f=100 cameras at x=0,1,2, camera 2 perturbed."""
    intr = intrinsic_block(100.0, 100.0, 50.0, 50.0)
    eye = np.eye(3)
    truth_R = np.stack([eye, eye, eye])
    truth_C = np.array([[0.0, 0, 0], [1.0, 0, 0], [2.0, 0, 0]])
    intrs = np.stack([intr] * 3)
    R = truth_R.copy()
    C = truth_C.copy()
    delta = np.array([0.001, -0.001, 0.0005, 0.02, -0.01, 0.005])
    R[2] = exp_so3(delta[:3]) @ R[2]
    C[2] = C[2] + delta[3:]
    tracks = []
    pts = []
    for index in range(24):
        truth = np.array([0.5 * (index % 4) - 0.75, 0.3 * ((index // 4) % 3) - 0.3, 5.0 + 0.2 * (index % 5)])
        pts.append(truth + np.array([0.03, -0.02, 0.04]))
        uv, _ = project_points(truth_R, truth_C, intrs, np.repeat(truth[None], 3, axis=0))
        tracks.append((np.arange(3), uv))
    return _finish(R, C, intrs, np.array(pts), tracks, fixed_cameras=np.array([0, 1], dtype=np.int32),
                   gauge_policy=GaugePolicy.RequireExplicitGauge,
                   meta={"workload": "native_three_cameras", "seed": None})


# --- ring of cameras around a cube (configuration 1) ------------------------------------------
def ring_problem(seed: int = 1, num_cameras: int = 50, num_points: int = 2000, max_views: int = 8,
                 noise_px: float = 0.5, rot_sigma: float = 0.01, center_sigma: float = 0.05,
                 point_sigma: float = 0.1, metashape: bool = False) -> Problem:
    """Cameras on a 10 m circle looking inward at points in a 4 m cube, one shared Brown group. With
    metashape=True the cameras use the Metashape convention with skew b2, k4, p3 and p4 as well (the full
    frame calibration)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    W, H, f = 4000, 3000, 3000.0
    ang = 2 * np.pi * np.arange(num_cameras) / num_cameras
    cen = np.stack([10 * np.cos(ang), 10 * np.sin(ang), np.zeros(num_cameras)], axis=1)
    rot = np.stack([look_at_rotation(c, (0, 0, 0)) for c in cen])
    dist = rng.normal(0.0, 1e-2, size=4)  # one shared Brown group: k1, k2, p1, p2 ~ N(0, sigma=1e-2)
    extra = rng.normal(0.0, [0.5, 1e-3, 1e-2, 1e-2]) if metashape else np.zeros(4)  # b2 (px), k4, p3, p4
    intr1 = intrinsic_block(f, 1.002 * f if metashape else f, W / 2, H / 2, k1=dist[0], k2=dist[1], p1=dist[2],
                            p2=dist[3], b2=extra[0], k4=extra[1], p3=extra[2], p4=extra[3])
    intr = np.repeat(intr1[None], num_cameras, axis=0)
    conv = np.full(num_cameras, 1 if metashape else 0, dtype=np.int32)
    truth = rng.uniform(-2.0, 2.0, size=(num_points, 3))
    tracks, pts, truth_index = [], [], []
    for p in range(num_points):
        X = np.repeat(truth[p][None], num_cameras, axis=0)
        uv, z = project_points(rot, cen, intr, X, metashape=conv == 1)
        vis = np.nonzero((z > 0) & (uv[:, 0] >= 0) & (uv[:, 0] < W) & (uv[:, 1] >= 0) & (uv[:, 1] < H))[0]
        vis = rng.permutation(vis)[:max_views]
        if len(vis) < 2:
            continue
        vis = np.sort(vis)
        obs = uv[vis] + rng.normal(0.0, noise_px, size=(len(vis), 2))
        tracks.append((vis, obs))
        pts.append(truth[p] + rng.normal(0.0, point_sigma, size=3))
        truth_index.append(p)
    # perturb every camera but the two gauge cameras 0 and 1 (as the reference benchmark does)
    dR = exp_so3(rng.normal(0.0, rot_sigma, size=(num_cameras, 3)))
    dC = rng.normal(0.0, center_sigma, size=(num_cameras, 3))
    dR[:2] = np.eye(3)
    dC[:2] = 0.0
    rot0 = dR @ rot
    cen0 = cen + dC
    return _finish(rot0, cen0, intr, np.array(pts), tracks,
                   cam_image_size=np.tile(np.array([W, H], dtype=np.int32), (num_cameras, 1)),
                   cam_group=np.zeros(num_cameras, dtype=np.int32),
                   cam_convention=conv if metashape else None,
                   fixed_cameras=np.array([0, 1], dtype=np.int32),
                   gauge_policy=GaugePolicy.RequireExplicitGauge,
                   meta={"workload": "ring", "seed": seed, "truth_rotation": rot, "truth_center": cen,
                         "truth_points": truth[np.array(truth_index, dtype=np.int64)]})


def rig_problem(seed: int = 6, num_captures: int = 16, num_points: int = 600, max_views: int = 8,
                sensor_yaw_deg=(0.0, 14.0, -14.0), sensor_offset: float = 0.3, noise_px: float = 0.5,
                capture_rot_sigma: float = 0.01, capture_center_sigma: float = 0.05, sensor_rot_sigma: float = 0.004,
                sensor_center_sigma: float = 0.02, fixed_sensors=(0,), distinct_groups: bool = False) -> Problem:
    """A multi-camera rig on a 10 m circle: each capture is a rig pose looking inward at points in a 4 m cube, and
    every capture carries the same sensors (cameras at sensor_yaw_deg about the rig's down axis, spaced
    sensor_offset m apart along the rig's right axis). Every camera is bound to its (capture, sensor). Capture
    and sensor poses are perturbed (fixed sensors stay at the truth), observations get N(0, noise_px) pixels.
    The cameras' own pose arrays hold the composed initial poses. With distinct_groups each sensor gets its own
    calibration group id explicitly (otherwise the rig's default grouping applies: one group per sensor)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    W, H, f = 4000, 3000, 3000.0
    S = len(sensor_yaw_deg)
    ang = 2 * np.pi * np.arange(num_captures) / num_captures
    capC = np.stack([10 * np.cos(ang), 10 * np.sin(ang), 0.3 * np.sin(3 * ang)], axis=1)
    capR = np.stack([look_at_rotation(c, (0, 0, 0)) for c in capC])
    senR = np.stack([exp_so3(np.array([0.0, np.deg2rad(y), 0.0])) for y in sensor_yaw_deg])
    senC = np.array([[sensor_offset * (k - (S - 1) / 2), 0.02 * k, 0.0] for k in range(S)])
    nc = num_captures * S
    cam_cap = np.repeat(np.arange(num_captures), S).astype(np.int32)
    cam_sen = np.tile(np.arange(S), num_captures).astype(np.int32)
    rot = np.einsum("nij,njk->nik", capR[cam_cap], senR[cam_sen])
    cen = capC[cam_cap] + np.einsum("nij,nj->ni", capR[cam_cap], senC[cam_sen])
    # radial k1, k2 ~ N(0, 1e-2); tangential p1, p2 ~ N(0, 2e-3), inside the default bound |p| <= 0.02
    dist = rng.normal(0.0, [1e-2, 1e-2, 2e-3, 2e-3])
    intr1 = intrinsic_block(f, f, W / 2, H / 2, k1=dist[0], k2=dist[1], p1=dist[2], p2=dist[3])
    intr = np.repeat(intr1[None], nc, axis=0)
    truth = rng.uniform(-2.0, 2.0, size=(num_points, 3))
    tracks, pts, truth_index = [], [], []
    for p in range(num_points):
        X = np.repeat(truth[p][None], nc, axis=0)
        uv, z = project_points(rot, cen, intr, X)
        vis = np.nonzero((z > 0) & (uv[:, 0] >= 0) & (uv[:, 0] < W) & (uv[:, 1] >= 0) & (uv[:, 1] < H))[0]
        vis = rng.permutation(vis)[:max_views]
        if len(vis) < 2:
            continue
        vis = np.sort(vis)
        tracks.append((vis, uv[vis] + rng.normal(0.0, noise_px, size=(len(vis), 2))))
        pts.append(truth[p] + rng.normal(0.0, 0.1, size=3))
        truth_index.append(p)
    fixed_s = np.zeros(S, dtype=np.int32)
    fixed_s[list(fixed_sensors)] = 1
    # captures 0 and num_captures / 2 stay at the truth: the automatic gauge fixes the capture of camera 0 and
    # the capture of the camera farthest from it
    dRk = exp_so3(rng.normal(0.0, capture_rot_sigma, size=(num_captures, 3)))
    dCk = rng.normal(0.0, capture_center_sigma, size=(num_captures, 3))
    dRk[[0, num_captures // 2]] = np.eye(3)
    dCk[[0, num_captures // 2]] = 0.0
    capR0 = dRk @ capR
    capC0 = capC + dCk
    dRs = exp_so3(rng.normal(0.0, sensor_rot_sigma, size=(S, 3)))
    dCs = rng.normal(0.0, sensor_center_sigma, size=(S, 3))
    dRs[fixed_s == 1] = np.eye(3)
    dCs[fixed_s == 1] = 0.0
    senR0 = dRs @ senR
    senC0 = senC + dCs
    rot0 = np.einsum("nij,njk->nik", capR0[cam_cap], senR0[cam_sen])
    cen0 = capC0[cam_cap] + np.einsum("nij,nj->ni", capR0[cam_cap], senC0[cam_sen])
    return _finish(rot0, cen0, intr, np.array(pts), tracks,
                   cam_image_size=np.tile(np.array([W, H], dtype=np.int32), (nc, 1)),
                   cam_group=(cam_sen.copy() if distinct_groups else None),
                   gauge_policy=GaugePolicy.AutoAnchor,
                   rig_capture_rotation=np.ascontiguousarray(capR0.reshape(-1, 9)),
                   rig_capture_center=np.ascontiguousarray(capC0),
                   rig_capture_fixed=np.zeros(num_captures, dtype=np.int32),
                   rig_sensor_rotation=np.ascontiguousarray(senR0.reshape(-1, 9)),
                   rig_sensor_center=np.ascontiguousarray(senC0),
                   rig_sensor_fixed=fixed_s,
                   cam_rig_capture=cam_cap,
                   cam_rig_sensor=cam_sen,
                   meta={"workload": "rig", "seed": seed, "truth_capture_rotation": capR,
                         "truth_capture_center": capC, "truth_sensor_rotation": senR, "truth_sensor_center": senC,
                         "truth_points": truth[np.array(truth_index, dtype=np.int64)]})


def rig_options(options=None, max_iterations: int = 10, intrinsics: bool = True):
    """Options of the rig workload: Huber 1 px and, with intrinsics, the shared focal length and radial
    distortion refined per group. A rig always takes the assembled Schur path with the default
    parameterisation (Assembly.cpp:24-29), whatever useReferenceOnlineSchur says."""
    from .options import ImageRobustLoss, SolveOptions

    o = options or SolveOptions()
    o.solver.imageRobustLoss = ImageRobustLoss.Huber
    o.solver.imageRobustLossScalePixels = 1.0
    o.solver.maxIterations = max_iterations
    o.solver.useReferenceOnlineSchur = True
    if intrinsics:
        o.calibration.refineSharedFocalLength = True
        o.calibration.refineSharedRadialDistortion = True
    return o


def add_survey_constraints(problem: Problem, seed: int = 11, num_control: int = 6, num_planes: int = 0,
                           control_sigma: float = 0.02, plane_noise: float = 0.01, kinds=(0,),
                           truth_points=None) -> Problem:
    """Control points and LiDAR point-to-plane constraints on distinct tracks of a copy of `problem`.

    Control points are surveyed truth positions plus N(0, control_sigma) per axis with the uncertainty forms in
    `kinds` (cycled: 0 isotropic sigma, 1 covariance, 2 square-root information of a random SPD matrix whose
    axis sigmas are about control_sigma). Laser planes pass within N(0, plane_noise) of the truth point with a
    random unit normal. Truth points default to problem.meta["truth_points"] (one per track)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = problem.copy()
    truth = np.asarray(truth_points if truth_points is not None else problem.meta["truth_points"], dtype=np.float64)
    order = rng.permutation(problem.num_tracks)
    cp_tracks = np.sort(order[:num_control]).astype(np.int32)
    lp_tracks = np.sort(order[num_control:num_control + num_planes]).astype(np.int32)
    if num_control:
        n = num_control
        kind = np.array([kinds[i % len(kinds)] for i in range(n)], dtype=np.int32)
        point = truth[cp_tracks] + rng.normal(0.0, control_sigma, size=(n, 3))
        matrix = np.zeros((n, 9))
        for i in range(n):
            Q = exp_so3(rng.normal(0.0, 0.5, size=3))
            sig = control_sigma * rng.uniform(0.7, 1.5, size=3)
            cov = Q @ np.diag(sig**2) @ Q.T
            if kind[i] == 1:
                matrix[i] = cov.reshape(9)
            elif kind[i] == 2:
                matrix[i] = (np.diag(1.0 / sig) @ Q.T).reshape(9)  # W^T W = cov^-1
        out.cp_track = cp_tracks
        out.cp_point = point
        out.cp_sigma = np.full(n, control_sigma)
        out.cp_weight = rng.uniform(0.5, 2.0, size=n)
        out.cp_uncertainty = kind
        out.cp_matrix = matrix
    if num_planes:
        n = num_planes
        nrm = rng.normal(size=(n, 3))
        nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
        tang = np.cross(nrm, rng.normal(size=(n, 3)))
        out.lp_track = lp_tracks
        out.lp_normal = nrm
        out.lp_point = truth[lp_tracks] + tang + nrm * rng.normal(0.0, plane_noise, size=(n, 1))
        out.lp_weight = rng.uniform(50.0, 200.0, size=n)
    return out


# --- aerial strip block (configuration 2) ------------------------------------------------------
def _fbm_height(x, y, rng_params):
    h = np.zeros_like(x)
    for amp, kx, ky, ph in rng_params:
        h += amp * np.sin(kx * x + ky * y + ph)
    return h


def aerial_problem(seed: int = 2, cameras_per_strip=(334, 333, 333), num_points: int = 200_000,
                   mean_track_length: float = 7.5, outlier_fraction: float = 0.01, noise_px: float = 0.5,
                   num_groups: int = 4, rot_sigma: float = 0.002, center_sigma: float = 0.2,
                   point_sigma: float = 0.2, focal_sigma: float = 0.002, distortion_sigma: float = 0.01,
                   altitude: float = 300.0, metashape: bool = False, strip_window: int | None = None,
                   chunk: int | None = None) -> Problem:
    """Aerial strip block: strips flown along +x at `altitude` with 80 % forward and 60 % side overlap, points
    on an fBm terrain, tracks of geometric length, outliers uniform over the image.

    metashape=True gives every group a full Metashape frame calibration (aspect, principal point, skew b2, k1..k4,
    p1..p4 besides the focal length). strip_window limits the candidate cameras of a point to the strips within
    that many strips of it, and chunk bounds the points drawn per batch (both needed for blocks of many strips;
    the defaults keep the draws of the three-strip block)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    W, H, f = 4000, 3000, 3000.0
    along = altitude * H / f          # 300 m footprint along track (image v axis)
    across = altitude * W / f         # 400 m footprint across track (image u axis)
    base = 0.2 * along                # 80 % forward overlap
    spacing = 0.4 * across            # 60 % side overlap
    strips = len(cameras_per_strip)
    n_cam = int(sum(cameras_per_strip))
    # cameras: all strips fly +x; camera right axis = +y world, down = +x world, forward = -z
    nadir = np.array([[0.0, 1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, -1.0]])
    cen, strip_of = [], []
    for s, n in enumerate(cameras_per_strip):
        for k in range(n):
            cen.append([k * base, s * spacing, altitude])
            strip_of.append(s)
    cen = np.array(cen)
    cen += rng.normal(0.0, 1.0, size=cen.shape) * np.array([2.0, 2.0, 1.0])  # flight irregularity
    att = rng.normal(0.0, 0.02, size=(n_cam, 3))                            # attitude irregularity
    rot = exp_so3(att) @ nadir
    group = (np.arange(n_cam) * num_groups) // n_cam
    # per-group truth intrinsics
    g_f = f * (1.0 + rng.normal(0.0, focal_sigma, size=num_groups))
    g_d = rng.normal(0.0, distortion_sigma, size=(num_groups, 2))           # k1, k2
    g_t = rng.normal(0.0, 1e-4, size=(num_groups, 2))                       # p1, p2
    if metashape:
        # aspect (1e-4), principal offsets (2 px), skew b2 (0.5 px), k3, k4 (1e-3), p3, p4 (1e-2)
        g_m = rng.normal(0.0, [1e-4, 2.0, 2.0, 0.5, 1e-3, 1e-3, 1e-2, 1e-2], size=(num_groups, 8))
        intr_true = np.stack([intrinsic_block(g_f[g], g_f[g] * np.exp(g_m[g, 0]), W / 2 + g_m[g, 1],
                                              H / 2 + g_m[g, 2], k1=g_d[g, 0], k2=g_d[g, 1], k3=g_m[g, 4],
                                              p1=g_t[g, 0], p2=g_t[g, 1], b2=g_m[g, 3], k4=g_m[g, 5],
                                              p3=g_m[g, 6], p4=g_m[g, 7]) for g in group])
    else:
        intr_true = np.stack([intrinsic_block(g_f[g], g_f[g], W / 2, H / 2, k1=g_d[g, 0], k2=g_d[g, 1],
                                              p1=g_t[g, 0], p2=g_t[g, 1]) for g in group])
    conv = np.full(n_cam, 1 if metashape else 0, dtype=np.int32)
    intr_init = np.repeat(intrinsic_block(f, f, W / 2, H / 2)[None], n_cam, axis=0)
    # terrain
    octaves = []
    for o in range(6):
        amp = 20.0 * 0.55**o
        k = 2 * np.pi / (2000.0 * 0.5**o)
        th = rng.uniform(0, 2 * np.pi)
        octaves.append((amp, k * np.cos(th), k * np.sin(th), rng.uniform(0, 2 * np.pi)))
    norm = sum(a for a, *_ in octaves) / 20.0
    octaves = [(a / norm, kx, ky, ph) for a, kx, ky, ph in octaves]
    x_lo, x_hi = cen[:, 0].min() - along / 2, cen[:, 0].max() + along / 2
    y_lo, y_hi = -across / 2, (strips - 1) * spacing + across / 2
    # Track length L = min(2 + G, visible views) with G geometric; q is solved (once, on the first batch)
    # so that the expected L over the batch equals the target mean.
    q = None
    half = int(np.ceil(along / base)) + 2
    win = 2 * half + 1
    strip_start = np.concatenate([[0], np.cumsum(cameras_per_strip)])
    pts_true, cams_l, uv_l, len_l = [], [], [], []
    have = 0
    ns = strips if strip_window is None else min(strips, 2 * strip_window + 1)
    while have < num_points:
        m = int((num_points - have) * 1.2) + 64
        if chunk is not None:
            m = min(m, chunk)
        px = rng.uniform(x_lo, x_hi, size=m)
        py = rng.uniform(y_lo, y_hi, size=m)
        pz = _fbm_height(px, py, octaves)
        X = np.stack([px, py, pz], axis=1)
        # candidate cameras: a window of along-track indices in every strip (or in the strips near the point)
        kc = np.floor(px / base).astype(np.int64)
        cand = np.full((m, ns * win), -1, dtype=np.int64)
        s0 = None if strip_window is None else np.rint(py / spacing).astype(np.int64) - strip_window
        for j in range(ns):
            k = kc[:, None] + np.arange(-half, half + 1)[None, :]
            if s0 is None:
                valid = (k >= 0) & (k < cameras_per_strip[j])
                cand[:, j * win:(j + 1) * win] = np.where(valid, strip_start[j] + k, -1)
            else:
                sj = s0 + j
                ok_s = (sj >= 0) & (sj < strips)
                sjc = np.clip(sj, 0, strips - 1)
                per = np.asarray(cameras_per_strip)[sjc]
                valid = ok_s[:, None] & (k >= 0) & (k < per[:, None])
                cand[:, j * win:(j + 1) * win] = np.where(valid, strip_start[sjc][:, None] + k, -1)
        flat = cand.reshape(-1)
        ok = flat >= 0
        cidx = np.where(ok, flat, 0)
        Xr = np.repeat(X, ns * win, axis=0)
        uv, z = project_points(rot[cidx], cen[cidx], intr_true[cidx], Xr, metashape=conv[cidx] == 1)
        vis = ok & (z > 0) & (uv[:, 0] >= 0) & (uv[:, 0] < W) & (uv[:, 1] >= 0) & (uv[:, 1] < H)
        vis = vis.reshape(m, -1)
        uv = uv.reshape(m, -1, 2)
        nvis = vis.sum(axis=1)
        if q is None:
            n2 = nvis[nvis >= 2].astype(np.float64)

            def expected(qq):
                return float(np.mean(2.0 + qq * (1.0 - qq ** (n2 - 2.0)) / (1.0 - qq)))

            lo_q, hi_q = 1e-6, 1.0 - 1e-9
            for _ in range(200):
                mid = 0.5 * (lo_q + hi_q)
                lo_q, hi_q = (mid, hi_q) if expected(mid) < mean_track_length else (lo_q, mid)
            q = 0.5 * (lo_q + hi_q)
        extra = rng.geometric(1.0 - q, size=m) - 1
        keys = np.where(vis, rng.random(size=vis.shape), 2.0)
        order = np.argsort(keys, axis=1)
        L = np.minimum(2 + extra, nvis)
        noise = rng.normal(0.0, noise_px, size=uv.shape)
        outl = rng.random(size=vis.shape) < outlier_fraction
        outl_uv = rng.uniform(0, 1, size=uv.shape) * np.array([W, H])
        # every visible point up to the target count: the first L views of its random order, by camera index
        rows = np.nonzero(nvis >= 2)[0][: num_points - have]
        Lr = L[rows]
        o_r = order[rows]
        sel = np.arange(o_r.shape[1])[None, :] < Lr[:, None]
        cam_r = np.take_along_axis(cand[rows], o_r, axis=1)
        srt = np.argsort(np.where(sel, cam_r, np.iinfo(np.int64).max), axis=1)
        col = np.take_along_axis(o_r, srt, axis=1)[np.take_along_axis(sel, srt, axis=1)]
        rr = np.repeat(rows, Lr)
        obs = uv[rr, col] + noise[rr, col]
        o_mask = outl[rr, col]
        obs[o_mask] = outl_uv[rr, col][o_mask]
        pts_true.append(X[rows])
        cams_l.append(cand[rr, col])
        uv_l.append(obs)
        len_l.append(Lr)
        have += len(rows)
    pts_true = np.concatenate(pts_true)
    lens = np.concatenate(len_l)
    split = np.cumsum(lens)[:-1]
    tracks = list(zip(np.split(np.concatenate(cams_l), split), np.split(np.concatenate(uv_l), split)))
    dR = exp_so3(rng.normal(0.0, rot_sigma, size=(n_cam, 3)))
    dC = rng.normal(0.0, center_sigma, size=(n_cam, 3))
    dR[:2] = np.eye(3)
    dC[:2] = 0.0
    rot0 = dR @ rot
    cen0 = cen + dC
    pts0 = pts_true + rng.normal(0.0, point_sigma, size=pts_true.shape)
    return _finish(rot0, cen0, intr_init, pts0, tracks,
                   cam_image_size=np.tile(np.array([W, H], dtype=np.int32), (n_cam, 1)),
                   cam_group=group.astype(np.int32),
                   cam_convention=conv if metashape else None,
                   fixed_cameras=np.array([0, 1], dtype=np.int32),
                   gauge_policy=GaugePolicy.RequireExplicitGauge,
                   meta={"workload": "aerial", "seed": seed, "geometric_q": q, "truth_rotation": rot, "truth_center": cen,
                         "truth_points": pts_true, "truth_intrinsics": intr_true})


def large_aerial_problem(seed: int = 3, strips: int = 20, cameras_per_strip: int = 500, num_points: int = 2_000_000,
                         num_control: int = 50, num_planes: int = 10_000, num_groups: int = 8,
                         outlier_fraction: float = 0.02, mean_track_length: float = 7.5) -> Problem:
    """Large aerial block with the full survey setup: a strip block (80 % / 60 % overlap) with a full Metashape
    calibration per group, 0.5 px noise plus outliers, surveyed control points (sigma 0.02 m, three uncertainty
    forms) and LiDAR point-to-plane constraints. The datum comes from the control points (no fixed camera).
    Defaults: 20 strips x 500 cameras, 8 groups, 2M points, about 15M observations, 50 control points, 10k
    planes."""
    p = aerial_problem(seed=seed, cameras_per_strip=(cameras_per_strip,) * strips, num_points=num_points,
                       mean_track_length=mean_track_length, outlier_fraction=outlier_fraction,
                       num_groups=num_groups, metashape=True, strip_window=2, chunk=200_000)
    p.fixed_cameras = None
    p.gauge_policy = GaugePolicy.AutoAnchor
    return add_survey_constraints(p, seed=seed + 100, num_control=num_control, num_planes=num_planes,
                                  kinds=(0, 1, 2), control_sigma=0.02)


def large_aerial_options(options=None):
    """Solve options of the large aerial block: Cauchy 1 px, every calibration slot refined per group."""
    from .options import ImageRobustLoss, SolveOptions

    o = options or SolveOptions()
    o.solver.imageRobustLoss = ImageRobustLoss.Cauchy
    o.solver.imageRobustLossScalePixels = 1.0
    o.solver.useReferenceOnlineSchur = True
    c = o.calibration
    c.refineSharedFocalLength = c.refineSharedFocalAspectRatio = c.refineSharedPrincipalPoint = True
    c.refineSharedRadialDistortion = c.refineSharedHighOrderDistortion = c.refineSharedMetashapeParameters = True
    return o


def aerial_options(options=None):
    """Solve options of the aerial workload: Cauchy 1 px, shared focal + Brown radial/tangential refined."""
    from .options import ImageRobustLoss, SolveOptions

    o = options or SolveOptions()
    o.solver.imageRobustLoss = ImageRobustLoss.Cauchy
    o.solver.imageRobustLossScalePixels = 1.0
    o.solver.useReferenceOnlineSchur = True
    o.calibration.refineSharedFocalLength = True
    o.calibration.refineSharedRadialDistortion = True
    o.calibration.refineSharedHighOrderDistortion = True
    return o


def ring_options(options=None, max_iterations: int = 10):
    """Solve options of the ring workload: Huber 1 px, 10 iterations, intrinsics fixed."""
    from .options import ImageRobustLoss, SolveOptions

    o = options or SolveOptions()
    o.solver.imageRobustLoss = ImageRobustLoss.Huber
    o.solver.imageRobustLossScalePixels = 1.0
    o.solver.maxIterations = max_iterations
    o.solver.useReferenceOnlineSchur = True
    return o


# --- batched local windows (configuration 4) ----------------------------------------------------
def window_problem(rng: np.random.Generator, num_cameras: int, num_points: int, track_range=(2, 12),
                   noise_px: float = 0.5, rot_sigma: float = 0.01, center_sigma: float = 0.05,
                   point_sigma: float = 0.1, arc_step_deg: float = 3.0) -> Problem:
    """One local window: cameras on an arc of a 10 m circle (arc_step_deg apart, heights ~ N(0, 0.5 m)) looking
    at a 4 m cube of points, 4000x3000 px, f = 3000 px, one shared Brown group (k1, k2, p1, p2 ~ N(0, 1e-2)).
    Track length L ~ U{track_range} capped by the visible cameras, which are drawn at random; 0.5 px noise;
    cameras 0 and 1 fixed (explicit gauge), the others and the points perturbed as the ring workload."""
    W, H, f = 4000, 3000, 3000.0
    step = np.deg2rad(arc_step_deg)
    ang = rng.uniform(0.0, 2 * np.pi) + step * (np.arange(num_cameras) - 0.5 * (num_cameras - 1))
    cen = np.stack([10 * np.cos(ang), 10 * np.sin(ang), rng.normal(0.0, 0.5, num_cameras)], axis=1)
    rot = np.stack([look_at_rotation(c, (0, 0, 0)) for c in cen])
    dist = rng.normal(0.0, 1e-2, size=4)
    intr1 = intrinsic_block(f, f, W / 2, H / 2, k1=dist[0], k2=dist[1], p1=dist[2], p2=dist[3])
    intr = np.repeat(intr1[None], num_cameras, axis=0)
    truth = rng.uniform(-2.0, 2.0, size=(num_points, 3))
    # every point in every camera: (points x cameras)
    pc = np.repeat(np.arange(num_cameras)[None, :], num_points, axis=0).reshape(-1)
    uv, z = project_points(rot[pc], cen[pc], intr[pc], np.repeat(truth, num_cameras, axis=0))
    vis = ((z > 0) & (uv[:, 0] >= 0) & (uv[:, 0] < W) & (uv[:, 1] >= 0) & (uv[:, 1] < H)).reshape(num_points, -1)
    uv = uv.reshape(num_points, num_cameras, 2)
    L = np.minimum(rng.integers(track_range[0], track_range[1] + 1, size=num_points), vis.sum(axis=1))
    keys = np.where(vis, rng.random(size=vis.shape), 2.0)
    order = np.argsort(keys, axis=1)
    noise = rng.normal(0.0, noise_px, size=uv.shape)
    keep = np.nonzero(L >= 2)[0]
    # the first L visible cameras of each kept point in the random order, then in camera order
    m = min(int(track_range[1]), num_cameras)
    first = order[keep, :m]
    chosen = np.arange(m)[None, :] < L[keep][:, None]
    picked = np.sort(np.where(chosen, first, num_cameras), axis=1)
    valid = picked < num_cameras
    rows = np.repeat(keep, valid.sum(axis=1))
    cam = picked[valid].astype(np.int32)
    obs_uv = uv[rows, cam] + noise[rows, cam]
    off = np.zeros(len(keep) + 1, dtype=np.int64)
    off[1:] = np.cumsum(valid.sum(axis=1))
    truth = truth[keep]
    dR = exp_so3(rng.normal(0.0, rot_sigma, size=(num_cameras, 3)))
    dC = rng.normal(0.0, center_sigma, size=(num_cameras, 3))
    dR[:2] = np.eye(3)
    dC[:2] = 0.0
    pts0 = truth + rng.normal(0.0, point_sigma, size=truth.shape)
    return Problem(
        cam_rotation=np.ascontiguousarray((dR @ rot).reshape(-1, 9)),
        cam_center=np.ascontiguousarray(cen + dC),
        cam_intrinsics=np.ascontiguousarray(intr),
        track_point=np.ascontiguousarray(pts0),
        track_offset=off,
        obs_camera=cam,
        obs_uv=np.ascontiguousarray(obs_uv),
        obs_weight=np.ones(len(cam)),
        obs_scale=np.ones(len(cam)),
        cam_image_size=np.tile(np.array([W, H], dtype=np.int32), (num_cameras, 1)),
        cam_group=np.zeros(num_cameras, dtype=np.int32),
        fixed_cameras=np.array([0, 1], dtype=np.int32),
        gauge_policy=GaugePolicy.RequireExplicitGauge,
        meta={"workload": "window", "truth_rotation": rot, "truth_center": cen, "truth_points": truth})


def batch_windows(seed: int = 4, num_windows: int = 512, cameras=(8, 64), points=(500, 8000),
                  track_range=(2, 12), noise_px: float = 0.5) -> list[Problem]:
    """Batched local BA: num_windows independent windows with cameras ~ U{cameras}, points ~ U{points} and
    track lengths ~ U{track_range} (seed 4; 512 windows, ~15M observations at full size)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(num_windows):
        nc = int(rng.integers(cameras[0], cameras[1] + 1))
        npts = int(rng.integers(points[0], points[1] + 1))
        out.append(window_problem(rng, nc, npts, track_range=track_range, noise_px=noise_px))
    return out


def window_options(options=None, max_iterations: int = 20):
    """Solve options of the batched windows: Huber 1 px, shared focal + Brown (k1, k2, k3, p1, p2) refined,
    the reference's default iteration budget."""
    from .options import ImageRobustLoss, SolveOptions

    o = options or SolveOptions()
    o.solver.imageRobustLoss = ImageRobustLoss.Huber
    o.solver.imageRobustLossScalePixels = 1.0
    o.solver.maxIterations = max_iterations
    o.solver.useReferenceOnlineSchur = True
    o.calibration.refineSharedFocalLength = True
    o.calibration.refineSharedRadialDistortion = True
    o.calibration.refineSharedHighOrderDistortion = True
    return o
