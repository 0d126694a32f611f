"""Known-answer fixtures that pin the oracle where only its own restatement covered it before:

  * intrinsic group initialisation (initializeIntrinsicGroups, BundleAdjustPlaMatrixModel.cpp:84-171): medians over
    differing cameras with odd and even counts, every slot's bound, values clamped at the upper and the lower
    bound, the low-order distortion scale > 1, and all 13 prior sigmas. Every expected number below is written out
    by hand from the rules, for the inputs chosen here.
  * the intrinsic step clamp (applyIntrinsicStep, BundleAdjustPlaMatrixModel.cpp:248-269): the objective after a
    step that overshoots a bound equals the objective of the problem whose parameter sits on that bound.
  * the gauge scale camera (Assembly.cpp:69-128, 204-206, 295-306): tangent bases worked out by hand, the
    sphere-preserving step and its derivative, and the analytic directional decrease of a solve with a scale camera
    against central differences of the objective along the step.

A planted slip in a bound, a sigma or a basis index fails one of these checks.
"""
import numpy as np
import pytest

from plabundle_amd import options as PO
from plabundle_amd import synthetic as S
from plabundle_amd.problem import GaugePolicy

LOG = np.log


def _slots(f, aspect, cx, cy, k1=0.0, k2=0.0, k3=0.0, p1=0.0, p2=0.0, b2=0.0, k4=0.0, p3=0.0, p4=0.0):
    return [f, LOG(aspect), cx, cy, k1, k2, k3, p1, p2, b2, k4, p3, p4]


# Seven Metashape-convention cameras: group 0 = cameras 0, 2, 4 (odd count), group 1 = cameras 1, 3, 5, 6 (even).
CAM = {
    0: _slots(3010, 1.01, 2010, 1900, k1=0.1, k2=0.5, k3=-0.2, p1=0.06, p2=-0.001, b2=1.0, k4=0.01, p3=0.4, p4=-0.4),
    2: _slots(3290, 0.99, 2300, 1760, k1=0.5, k2=0.6, k3=0.1, p1=0.07, p2=0.002, b2=500.0, k4=0.02, p3=-0.5,
              p4=-0.5),
    4: _slots(2950, 1.20, 1990, 1750, k1=0.3, k2=0.4, k3=0.0, p1=0.01, p2=0.0, b2=-2.0, k4=0.03, p3=0.1, p4=-0.36),
    1: _slots(500, 1.00, 1000, 800, k1=0.9),
    3: _slots(600, 1.10, 1020, 800, k1=1.0),
    5: _slots(900, 0.90, 1040, 800, k1=-0.1),
    6: _slots(1000, 1.05, 1100, 800, k1=0.2),
}
REF = {
    0: _slots(3000, 1.0, 2000, 1500, k1=0.0, p1=0.001),
    2: _slots(3300, 1.0, 2000, 1500, k1=0.2, p1=0.002),
    4: _slots(2900, 1.0, 2000, 1500, k1=0.1, p1=0.003),
    1: _slots(2000, 1.0, 1000, 800),
    3: _slots(2100, 1.0, 1000, 800),
    5: _slots(2400, 1.0, 1000, 800),
    6: _slots(2200, 1.0, 1000, 800),
}
LOW_ORDER_SCALE = 2.5

# group 0: reference focal = median(3000, 3300, 2900) = 3000, reference aspect 1
G0_PARAMS = [3010.0,            # median(3010, 3290, 2950)
             LOG(1.01),         # log median(1.01, 0.99, 1.2)
             10.0,              # median(10, 300, -10) of cx - reference cx
             240.0,             # median(400, 260, 250) = 260 clamped to 0.08 * 3000
             0.3,               # median(0.1, 0.5, 0.3)
             0.35,              # median(0.5, 0.6, 0.4) = 0.5 clamped to |k2| <= 0.35 (not scaled)
             0.0,               # median(-0.2, 0.1, 0.0)
             0.05,              # median(0.06, 0.07, 0.01) = 0.06 clamped to 0.02 * 2.5
             0.0,               # median(-0.001, 0.002, 0.0)
             1.0,               # median(1, 500, -2); skew bound 0.10 * 3000 = 300
             0.02,              # median(0.01, 0.02, 0.03)
             0.1,               # median(0.4, -0.5, 0.1)
             -0.35]             # median(-0.4, -0.5, -0.36) = -0.4 clamped to -0.35
G0_PRIOR = [3000.0, 0.0, 0.0, 0.0, 0.1, 0.0, 0.0, 0.002, 0.0, 0.0, 0.0, 0.0, 0.0]
G0_UPPER = [12000.0, LOG(1.18), 240.0, 240.0, 0.875, 0.35, 0.35, 0.05, 0.05, 300.0, 0.35, 0.35, 0.35]
G0_LOWER = [1500.0, LOG(0.85)] + [-u for u in G0_UPPER[2:]]
G0_SIGMA = [1050.0, 0.08, 120.0, 120.0, 0.375, 0.15, 0.15, 0.025, 0.025, 60.0, 0.15, 0.15, 0.15]

# group 1: reference focal = median(2000, 2100, 2400, 2200) = (2100 + 2200) / 2 = 2150 (even count: mean of the
# two middle values, the oracle's reading of finiteMedian)
G1_PARAMS = [1075.0,            # median(500, 600, 900, 1000) = 750 clamped to 0.5 * 2150
             LOG(1.025),        # log median(1.0, 1.1, 0.9, 1.05) = (1.0 + 1.05) / 2
             30.0,              # median(0, 20, 40, 100) = (20 + 40) / 2
             0.0,
             0.55,              # median(0.9, 1.0, -0.1, 0.2) = (0.2 + 0.9) / 2
             0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0]
G1_PRIOR = [2150.0] + [0.0] * 12
G1_UPPER = [8600.0, LOG(1.18), 172.0, 172.0, 0.875, 0.35, 0.35, 0.05, 0.05, 215.0, 0.35, 0.35, 0.35]
G1_LOWER = [1075.0, LOG(0.85)] + [-u for u in G1_UPPER[2:]]
G1_SIGMA = [752.5, 0.08, 86.0, 86.0, 0.375, 0.15, 0.15, 0.025, 0.025, 43.0, 0.15, 0.15, 0.15]


def _group_problem():
    p = S.ring_problem(seed=21, num_cameras=7, num_points=120, max_views=5, metashape=True)
    p.cam_intrinsics = np.array([CAM[c] for c in range(7)], dtype=np.float64)
    p.ref_intrinsics = np.array([REF[c] for c in range(7)], dtype=np.float64)
    p.cam_group = np.array([0, 1, 0, 1, 0, 1, 1], dtype=np.int32)
    return p


def _all_slot_options():
    o = S.ring_options(PO.SolveOptions(), max_iterations=5)
    c = o.calibration
    c.refineSharedFocalLength = c.refineSharedFocalAspectRatio = c.refineSharedPrincipalPoint = True
    c.refineSharedRadialDistortion = c.refineSharedHighOrderDistortion = c.refineSharedMetashapeParameters = True
    c.sharedLowOrderDistortionScale = LOW_ORDER_SCALE
    o.solver.maxInitialTrackRms = 1e12  # the fixture's intrinsics are not meant to fit the observations
    return o


def test_group_initialisation_known_answers(oracle_mod):
    p = _group_problem()
    copts = PO.to_c(_all_slot_options())
    assert copts.shared_intrinsic_mask == 0x1FFF
    orc = oracle_mod.Oracle(p, threads=1)
    prep = orc.prepare(copts)
    assert prep["prepared"] and prep["intrinsic_block_count"] == 2
    g = orc.groups()
    assert list(g["enabled"]) == [0x1FFF, 0x1FFF]
    for gi, (par, pri, low, up, sig) in enumerate(((G0_PARAMS, G0_PRIOR, G0_LOWER, G0_UPPER, G0_SIGMA),
                                                   (G1_PARAMS, G1_PRIOR, G1_LOWER, G1_UPPER, G1_SIGMA))):
        np.testing.assert_allclose(g["params"][gi], par, rtol=1e-13, atol=1e-15, err_msg=f"group {gi} params")
        np.testing.assert_allclose(g["prior"][gi], pri, rtol=1e-13, atol=1e-15, err_msg=f"group {gi} prior")
        np.testing.assert_allclose(g["lower"][gi], low, rtol=1e-13, err_msg=f"group {gi} lower")
        np.testing.assert_allclose(g["upper"][gi], up, rtol=1e-13, err_msg=f"group {gi} upper")
        np.testing.assert_allclose(g["inv_sigma"][gi], 1.0 / np.array(sig), rtol=1e-13,
                                   err_msg=f"group {gi} inverse sigma")


def test_low_order_scale_below_one_is_one(oracle_mod):
    # std::max(1.0, sharedLowOrderDistortionScale): a scale below one leaves the k1, p1, p2 bounds and sigmas at
    # their option values
    p = _group_problem()
    o = _all_slot_options()
    o.calibration.sharedLowOrderDistortionScale = 0.25
    orc = oracle_mod.Oracle(p, threads=1)
    assert orc.prepare(PO.to_c(o))["prepared"]
    g = orc.groups()
    np.testing.assert_allclose(g["upper"][0][[4, 7, 8]], [0.35, 0.02, 0.02], rtol=1e-14)
    np.testing.assert_allclose(1.0 / g["inv_sigma"][0][[4, 7, 8]], [0.15, 0.01, 0.01], rtol=1e-14)
    # p1 median 0.06 now clamps to 0.02, k1 0.3 stays
    np.testing.assert_allclose(g["params"][0][[4, 7]], [0.3, 0.02], rtol=1e-14)


def test_disabled_slots_have_no_prior_weight(oracle_mod):
    # only the focal length refined: every other slot's inverse sigma is zero, its value still the clamped median
    p = _group_problem()
    o = S.ring_options(PO.SolveOptions(), max_iterations=5)
    o.calibration.refineSharedFocalLength = True
    o.solver.maxInitialTrackRms = 1e12
    orc = oracle_mod.Oracle(p, threads=1)
    assert orc.prepare(PO.to_c(o))["prepared"]
    g = orc.groups()
    assert list(g["enabled"]) == [1, 1]
    np.testing.assert_allclose(g["inv_sigma"][:, 0], [1 / 1050.0, 1 / 752.5], rtol=1e-14)
    assert np.all(g["inv_sigma"][:, 1:] == 0.0)
    np.testing.assert_allclose(g["params"][:, 0], [3010.0, 1075.0], rtol=1e-14)


# --- intrinsic step clamp ---------------------------------------------------------------------------------------
def _focal_only(p):
    """Only the shared focal length is refined (poses and points fixed): one primary block, slot 0."""
    q = p.copy()
    q.fixed_tracks = np.arange(q.num_tracks, dtype=np.int32)
    o = S.ring_options(PO.SolveOptions(), max_iterations=5)
    o.calibration.refineCameraPose = False
    o.calibration.refineSharedFocalLength = True
    o.solver.maxInitialTrackRms = 1e12  # the same tracks stay active at every focal length compared
    return q, PO.to_c(o)


def _cost_at_focal(oracle_mod, p, copts, focal):
    q = p.copy()
    q.ref_intrinsics = p.cam_intrinsics.copy()  # same reference: same prior, bounds and sigma
    q.cam_intrinsics = p.cam_intrinsics.copy()
    q.cam_intrinsics[:, 0] = focal
    orc = oracle_mod.Oracle(q, threads=1)
    assert orc.prepare(copts)["prepared"]
    assert orc.groups()["params"][0, 0] == pytest.approx(focal, rel=1e-15)
    return orc.initial_cost()


def test_intrinsic_step_is_clamped_to_the_group_bounds(oracle_mod):
    base = S.ring_problem(seed=22, num_cameras=8, num_points=80, max_views=5)
    p, copts = _focal_only(base)
    orc = oracle_mod.Oracle(p, threads=1)
    prep = orc.prepare(copts)
    assert prep["primary_block_count"] == 1 and prep["camera_block_count"] == 0 and prep["track_block_count"] == 0
    g = orc.groups()
    f0, lo, up = g["params"][0, 0], g["lower"][0, 0], g["upper"][0, 0]
    assert (lo, up) == (0.5 * f0, 4.0 * f0)  # the cameras' own focal is the reference: [0.5 f, 4 f]
    elim = np.zeros(0)
    for delta, landing in ((0.01 * f0, 1.01 * f0), (5.0 * f0, up), (-0.9 * f0, lo)):
        step = np.zeros(13)
        step[0] = delta
        got = orc.cost_after_step(step, elim, 1.0)
        want = _cost_at_focal(oracle_mod, p, copts, landing)
        assert got == pytest.approx(want, rel=1e-12), (delta, landing)


# --- gauge scale camera ------------------------------------------------------------------------------------------
@pytest.mark.parametrize("anchor,scale,first,second", [
    # scale camera along +x: target-frame unit (1, 0, 0), helper axis y, second = y x u = (0, 0, -1),
    # first = u x second = (0, 1, 0); back to the world frame (y and z negated)
    ((0.0, 0.0, 0.0), (10.0, 0.0, 0.0), (0.0, -1.0, 0.0), (0.0, 0.0, 1.0)),
    # scale camera straight below the anchor: target-frame unit (0, 0, 1), helper axis x, second = x x u =
    # (0, -1, 0), first = u x second = (1, 0, 0)
    ((1.0, 2.0, 3.0), (1.0, 2.0, -2.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0)),
])
def test_gauge_tangent_basis_known_answers(oracle_mod, anchor, scale, first, second):
    b = oracle_mod.gauge_basis(anchor, scale)
    np.testing.assert_allclose(b[0], first, atol=1e-15)
    np.testing.assert_allclose(b[1], second, atol=1e-15)


@pytest.mark.parametrize("scale", [(3.0, 4.0, 12.0), (-7.0, 0.5, 2.0), (0.1, -9.0, 0.3)])
def test_gauge_tangent_basis_and_step_properties(oracle_mod, scale):
    anchor = np.array([0.5, -1.0, 2.0])
    c = anchor + np.array(scale)
    u = (c - anchor) / np.linalg.norm(c - anchor)
    b = oracle_mod.gauge_basis(anchor, c)
    # orthonormal, tangent to the sphere around the anchor, and first = u x second, so first x second = -u
    np.testing.assert_allclose(b @ b.T, np.eye(2), atol=1e-14)
    np.testing.assert_allclose(b @ u, [0.0, 0.0], atol=1e-14)
    np.testing.assert_allclose(np.cross(b[0], b[1]), -u, atol=1e-14)
    baseline = 7.5
    # the step keeps the baseline; the zero step projects the centre onto the sphere
    c0 = oracle_mod.gauge_step(anchor, c, baseline, 0.0, 0.0)
    np.testing.assert_allclose(c0, anchor + baseline * u, atol=1e-13)
    for a, bb in ((0.3, -0.2), (-2.0, 1.0)):
        ca = oracle_mod.gauge_step(anchor, c, baseline, a, bb)
        assert np.linalg.norm(ca - anchor) == pytest.approx(baseline, rel=1e-14)
    # d centre / d (first, second) at zero is the basis scaled by baseline / |C - anchor| (central differences)
    h = 1e-6
    scale_factor = baseline / np.linalg.norm(c - anchor)
    for k in range(2):
        e = np.zeros(2)
        e[k] = h
        d = (oracle_mod.gauge_step(anchor, c, baseline, *e) - oracle_mod.gauge_step(anchor, c, baseline, *(-e))) / (2 * h)
        np.testing.assert_allclose(d, scale_factor * b[k], atol=1e-8)


def test_directional_decrease_with_a_gauge_scale_camera(oracle_mod):
    # AutoAnchor without an absolute scale fixes camera 0 and moves the farthest camera on the sphere of the
    # baseline (BundleAdjustValidation.cpp:869-893): its centre step enters through the two tangent columns. The
    # analytic decrease along a step must match central differences of the objective along it, for the solved
    # step and for steps of the scale camera's tangent components alone (which isolate those columns).
    p = S.ring_problem(seed=23, num_cameras=10, num_points=150, max_views=5)
    p.fixed_cameras = None
    p.gauge_policy = GaugePolicy.AutoAnchor
    copts = PO.to_c(S.ring_options(PO.SolveOptions(), max_iterations=5))
    orc = oracle_mod.Oracle(p, threads=2)
    prep = orc.prepare(copts)
    assert prep["gauge_anchor_camera"] == 0 and prep["gauge_scale_camera"] > 0 and prep["gauge_baseline"] > 0
    P, Ne = prep["primary_block_count"], prep["track_block_count"]
    primary, elim, dd, _, _ = orc.compute_step(P, Ne, 1e-3)
    sb = orc.masks()["camera_block"][prep["gauge_scale_camera"]]
    assert sb >= 0 and primary[sb * 13 + 5] == 0.0
    assert orc.decrease(primary, elim) == pytest.approx(dd, rel=1e-12)

    def fd(step, e, h):
        return -(orc.cost_after_step(step, e, h) - orc.cost_after_step(step, e, -h)) / (2 * h)

    assert fd(primary, elim, 1e-5) == pytest.approx(dd, rel=2e-5)
    zero = np.zeros_like(elim)
    for comp in ((1.0, 0.0), (0.0, 1.0), (0.6, -0.8)):
        step = np.zeros_like(primary)
        step[sb * 13 + 3], step[sb * 13 + 4] = 1e-3 * comp[0], 1e-3 * comp[1]
        want = orc.decrease(step, zero)
        assert abs(want) > 0
        assert fd(step, zero, 1e-2) == pytest.approx(want, rel=1e-6)
