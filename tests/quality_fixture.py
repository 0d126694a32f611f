"""Hand-built problem with known strict reprojection RMS values for the post-solve quality checks.

Three identity-rotation pinhole cameras (f = 100 px, principal point (50, 50)) at x = 0, 1, 2 looking along +Z;
nothing is refined (no pose refinement, every track fixed), so the refined state equals the input state and
each track's strict RMS (BundleAdjustQuality.cpp:302-363) is known in closed form from the pixel offsets added
to the exact projections u = f x / z + cx, v = f y / z + cy.
"""
import math

import numpy as np

from plabundle_amd.problem import GaugePolicy, Problem, intrinsic_block

F, CX, CY = 100.0, 50.0, 50.0
CENTERS = np.array([[0.0, 0, 0], [1.0, 0, 0], [2.0, 0, 0]])


def _pixel(cam, X):
    d = np.asarray(X) - CENTERS[cam]
    return np.array([F * d[0] / d[2] + CX, F * d[1] / d[2] + CY])


# track: (point, [(camera, (du, dv), weight, scale), ...], expected strict RMS)
TRACKS = [
    # A: two cameras, offsets (0.3, 0.4) and 0 -> sqrt(0.25 / 4)
    ((0.5, 0.2, 5.0), [(0, (0.3, 0.4), 1.0, 1.0), (1, (0.0, 0.0), 1.0, 1.0)], math.sqrt(0.25 / 4)),
    # B: weight 4 with scale 2 whitens to 1 -> sqrt(2 / 6)
    ((1.0, -0.3, 6.0), [(0, (1.0, 0.0), 4.0, 2.0), (1, (0.0, 1.0), 1.0, 1.0), (2, (0.0, 0.0), 1.0, 1.0)],
     math.sqrt(2.0 / 6.0)),
    # C: one distinct camera (observed twice) -> inf; never activated
    ((0.2, 0.1, 4.0), [(0, (0.1, 0.0), 1.0, 1.0), (0, (0.0, 0.1), 1.0, 1.0)], math.inf),
    # D: behind camera 2 -> inf; never activated
    ((1.0, 0.0, -3.0), [(0, (0.0, 0.0), 1.0, 1.0), (2, (0.0, 0.0), 1.0, 1.0)], math.inf),
    # E: a zero-weight observation with a NaN pixel is skipped -> sqrt((0.25 + 0.25) / 4)
    ((1.5, 0.4, 5.5), [(0, (0.5, 0.0), 1.0, 1.0), (1, (0.0, 0.5), 1.0, 1.0), (2, (math.nan, math.nan), 0.0, 1.0)],
     math.sqrt(0.5 / 4)),
    # F: a weighted observation of a camera that does not exist -> inf (the track is still active)
    ((0.8, 0.0, 5.0), [(0, (0.1, 0.0), 1.0, 1.0), (1, (0.0, 0.1), 1.0, 1.0), (7, (0.0, 0.0), 1.0, 1.0)],
     math.inf),
    # G: 10 px offsets -> 10 / sqrt(2), removed by the point filter
    ((1.2, 0.3, 6.5), [(1, (10.0, 0.0), 1.0, 1.0), (2, (0.0, 10.0), 1.0, 1.0)], 10.0 / math.sqrt(2.0)),
]
ACTIVE = [True, True, False, False, True, True, True]


def problem() -> Problem:
    cams, uv, w, sc, off, pts = [], [], [], [], [0], []
    for X, obs, _ in TRACKS:
        pts.append(X)
        for cam, (du, dv), weight, scale in obs:
            base = _pixel(cam, X) if 0 <= cam < len(CENTERS) else np.array([CX, CY])
            cams.append(cam)
            uv.append(base + np.array([du, dv]))
            w.append(weight)
            sc.append(scale)
        off.append(len(cams))
    n = len(CENTERS)
    return Problem(
        cam_rotation=np.tile(np.eye(3).reshape(1, 9), (n, 1)),
        cam_center=CENTERS.copy(),
        cam_intrinsics=np.tile(intrinsic_block(F, F, CX, CY), (n, 1)),
        track_point=np.array(pts, dtype=np.float64),
        track_offset=np.array(off, dtype=np.int64),
        obs_camera=np.array(cams, dtype=np.int32),
        obs_uv=np.array(uv, dtype=np.float64),
        obs_weight=np.array(w, dtype=np.float64),
        obs_scale=np.array(sc, dtype=np.float64),
        fixed_tracks=np.arange(len(TRACKS), dtype=np.int32),
        gauge_policy=GaugePolicy.CallerManaged,
    )


def options():
    from plabundle_amd import options as PO

    o = PO.SolveOptions()
    o.calibration.refineCameraPose = False
    return o


def expected_rms() -> np.ndarray:
    return np.array([t[2] for t in TRACKS])
