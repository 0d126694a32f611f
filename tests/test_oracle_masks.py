"""Known-answer fixtures for the active-problem and gauge rules (A14), built by hand from
BundleAdjustPlaMatrixProblem.cpp:29-278 and BundleAdjustValidation.cpp:261-326,767-896."""
import numpy as np

from plabundle_amd import options as PO
from plabundle_amd.problem import GaugePolicy, Problem, intrinsic_block

EYE = np.eye(3).reshape(9)
INTR = intrinsic_block(100.0, 100.0, 50.0, 50.0)


def _uv(cx, X):
    # pinhole of the cameras below: identity rotation, centre (cx, 0, 0), f = 100, c = (50, 50)
    return 100.0 * (X[0] - cx) / X[2] + 50.0, 100.0 * X[1] / X[2] + 50.0


def _scene(gauge=GaugePolicy.RequireExplicitGauge, fixed=(0, 1), groups=None):
    centers = [0.0, 1.0, 2.0, 3.0, 4.0]
    cams = [(EYE, [c, 0.0, 0.0], INTR) for c in centers]
    X = np.array([0.5, 0.2, 5.0])
    good = [(k, *_uv(centers[k], X)) for k in (0, 1, 2)]
    tracks = [
        (X, good),                                                   # 0 active
        (X, [good[0]]),                                              # 1 one observation
        (X, [good[1], (1, good[1][1] + 0.1, good[1][2])]),           # 2 two observations, one camera
        ([0.5, 0.2, -5.0], good),                                    # 3 behind the cameras
        (X, [(0, 300.0, 300.0), (2, *_uv(2.0, X))]),                 # 4 initial RMS > 100 px
        ([np.nan, 0.0, 5.0], good),                                  # 5 non-finite point
        (X, [good[0], (1, np.nan, 3.0), (3, *_uv(3.0, X))]),         # 6 NaN observation skipped
        (X, [good[0], (2, *_uv(2.0, X), 0.0)]),                      # 7 zero weight -> one usable obs
        (X, [(2, *_uv(2.0, X)), (3, *_uv(3.0, X))]),                 # 8 active (cameras 2, 3)
    ]
    kw = dict(fixed_cameras=np.array(fixed, dtype=np.int32) if fixed is not None else None, gauge_policy=gauge)
    if groups is not None:
        kw["cam_group"] = np.array(groups, dtype=np.int32)
    return Problem.from_tracks(cams, tracks, **kw)


def test_active_tracks_and_block_numbering(oracle_mod):
    p = _scene()
    orc = oracle_mod.Oracle(p, threads=1)
    rep = orc.prepare(oracle_mod.default_options())
    assert rep["prepared"]
    m = orc.masks()
    assert m["active_track"].tolist() == [1, 0, 0, 0, 0, 0, 1, 0, 1]
    assert rep["rejected_initial_tracks"] == 1
    assert rep["observation_count"] == 3 + 2 + 2
    # cameras 0,1 fixed; camera 4 has no residual -> blocks for cameras 2, 3 in camera order
    assert m["camera_block"].tolist() == [-1, -1, 0, 1, -1]
    assert m["track_block"].tolist() == [0, -1, -1, -1, -1, -1, 1, -1, 2]
    assert m["intrinsic_block_by_camera"].tolist() == [-1] * 5
    assert rep["primary_block_count"] == 2


def test_fixed_tracks_keep_residuals_without_blocks(oracle_mod):
    p = _scene()
    p.fixed_tracks = np.array([6], dtype=np.int32)
    orc = oracle_mod.Oracle(p, threads=1)
    rep = orc.prepare(oracle_mod.default_options())
    m = orc.masks()
    assert m["active_track"].tolist() == [1, 0, 0, 0, 0, 0, 1, 0, 1]
    assert m["track_block"].tolist() == [0, -1, -1, -1, -1, -1, -1, -1, 1]
    assert rep["track_block_count"] == 2


def test_intrinsic_groups_in_first_appearance_order(oracle_mod):
    p = _scene(groups=[7, 3, 7, 5, 3])
    o = PO.SolveOptions()
    o.calibration.refineSharedFocalLength = True
    orc = oracle_mod.Oracle(p, threads=1)
    rep = orc.prepare(PO.to_c(o))
    m = orc.masks()
    assert m["calibration_group_by_camera"].tolist() == [0, 1, 0, 2, 1]
    # intrinsic blocks follow the two camera blocks
    assert m["intrinsic_block_by_camera"].tolist() == [2, 3, 2, 4, 3]
    assert rep["intrinsic_block_count"] == 3 and rep["primary_block_count"] == 5


def test_auto_anchor_fixes_camera_zero_and_farthest_scale_camera(oracle_mod):
    p = _scene(gauge=GaugePolicy.AutoAnchor, fixed=None)
    orc = oracle_mod.Oracle(p, threads=1)
    rep = orc.prepare(oracle_mod.default_options())
    assert rep["gauge_anchor_camera"] == 0
    assert rep["gauge_scale_camera"] == 4                       # farthest from camera 0
    assert abs(rep["gauge_baseline"] - 4.0) < 1e-15
    m = orc.masks()
    assert m["camera_block"].tolist() == [-1, 0, 1, 2, -1]      # camera 0 fixed, camera 4 no residual


def test_auto_anchor_tie_goes_to_lowest_index(oracle_mod):
    p = _scene(gauge=GaugePolicy.AutoAnchor, fixed=None)
    p.cam_center = np.array([[0.0, 0, 0], [2.0, 0, 0], [-2.0, 0, 0], [0, 2.0, 0], [1.0, 0, 0]])
    orc = oracle_mod.Oracle(p, threads=1)
    rep = orc.prepare(oracle_mod.default_options())
    assert rep["gauge_scale_camera"] == 1


def test_require_explicit_gauge_rejects_single_fixed_camera(oracle_mod):
    p = _scene(fixed=(0,))
    orc = oracle_mod.Oracle(p, threads=1)
    rep = orc.prepare(oracle_mod.default_options())
    assert not rep["prepared"]
    assert rep["solve_status"] == 5  # unsupported configuration


def test_invalid_fixed_camera_index(oracle_mod):
    p = _scene(fixed=(0, 9))
    orc = oracle_mod.Oracle(p, threads=1)
    rep = orc.prepare(oracle_mod.default_options())
    assert not rep["prepared"] and rep["solve_status"] == 4
