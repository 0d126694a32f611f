"""CPU checks of the C ABI: the library loads, exports every function include/pb_hip.h declares,
and the ctypes structures match the C layouts (compiled here with gcc)."""
import os
import re
import subprocess
import tempfile

import ctypes as C
import pytest

from plabundle_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pb_hip.h")


def _declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:pb_status|void|int|const char\*)\s+(pb_\w+)\s*\(", text, re.M)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(_abi.LIBRARY_PATH):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "plabundle_amd", "csrc")], check=True)
    return _abi.load_library()


def test_header_functions_are_exported(lib):
    declared = _declared_functions()
    assert set(declared) == set(_abi.EXPORTED_SYMBOLS)
    out = subprocess.run(["nm", "-D", "--defined-only", _abi.LIBRARY_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\sT\s(pb_\w+)$", out, re.M))
    assert set(declared) <= exported
    for name in declared:
        assert hasattr(lib, name)


def test_struct_layouts_match_header():
    src = r"""
#include <stddef.h>
#include <stdio.h>
#include "pb_hip.h"
int main(void) {
  printf("%zu %zu %zu\n", sizeof(pb_problem_view), sizeof(pb_solve_options), sizeof(pb_solve_report));
  printf("%zu %zu %zu %zu %zu %zu\n", offsetof(pb_problem_view, gauge_baseline),
         offsetof(pb_solve_options, step_rms_tolerance), offsetof(pb_solve_report, message),
         offsetof(pb_problem_view, cp_matrix), offsetof(pb_problem_view, lp_weight),
         offsetof(pb_solve_options, control_point_huber_delta));
  printf("%zu %zu %zu %zu\n", sizeof(pb_batch_view), offsetof(pb_batch_view, windows), sizeof(pb_window_report),
         offsetof(pb_window_report, device_seconds));
  printf("%zu %zu %zu %zu\n", sizeof(pb_quality_options), offsetof(pb_quality_options, min_accepted_valid_track_ratio),
         sizeof(pb_quality_report), offsetof(pb_quality_report, message));
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "layout.c")
        exe = os.path.join(d, "layout")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        lines = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    sizes = [int(x) for x in lines[0].split()]
    offs = [int(x) for x in lines[1].split()]
    assert sizes == [C.sizeof(_abi.ProblemView), C.sizeof(_abi.SolveOptionsC), C.sizeof(_abi.SolveReportC)]
    assert offs == [_abi.ProblemView.gauge_baseline.offset, _abi.SolveOptionsC.step_rms_tolerance.offset,
                    _abi.SolveReportC.message.offset, _abi.ProblemView.cp_matrix.offset,
                    _abi.ProblemView.lp_weight.offset, _abi.SolveOptionsC.control_point_huber_delta.offset]
    batch = [int(x) for x in lines[2].split()]
    assert batch == [C.sizeof(_abi.BatchView), _abi.BatchView.windows.offset, C.sizeof(_abi.WindowReportC),
                     _abi.WindowReportC.device_seconds.offset]
    quality = [int(x) for x in lines[3].split()]
    assert quality == [C.sizeof(_abi.QualityOptionsC), _abi.QualityOptionsC.min_accepted_valid_track_ratio.offset,
                       C.sizeof(_abi.QualityReportC), _abi.QualityReportC.message.offset]


# Reference defaults written out from include/plabundle/options.h (SolverControlOptions :58-75, CalibrationOptions
# :77-120, ConstraintOptions :122-135, BackendOptions :137-164), keyed by the pb_solve_options field that carries
# them. A default changed on either side fails here.
REFERENCE_DEFAULTS = {
    "max_iterations": 20,                      # options.h:60 maxIterations
    "robust_loss": 0,                          # options.h:61 ImageRobustLoss::LeastSquares
    "robust_loss_scale": 1.0,                  # options.h:62 imageRobustLossScalePixels
    "armijo_coefficient": 1.0e-3,              # options.h:63 referenceArmijoCoefficient
    "line_search_steps": 34,                   # options.h:64 referenceLineSearchSteps
    "use_online_schur": 0,                     # options.h:65 useReferenceOnlineSchur = false
    "max_initial_track_rms": 100.0,            # options.h:69 maxInitialTrackRms
    "refine_camera_pose": 1,                   # options.h:79 refineCameraPose
    # options.h:80-86: only refineSharedHighOrderDistortion is on, which refines nothing without the radial
    # switch (src/backend.cpp:8-47), so no slot is refined by default
    "shared_intrinsic_mask": 0,
    "min_focal_scale": 0.5,                    # options.h:91 minSharedFocalScale
    "max_focal_scale": 4.0,                    # options.h:92 maxSharedFocalScale
    "min_aspect_scale": 0.85,                  # options.h:93
    "max_aspect_scale": 1.18,                  # options.h:94
    "max_principal_offset_fraction": 0.08,     # options.h:95
    "focal_prior_sigma": 0.35,                 # options.h:96
    "principal_prior_sigma_fraction": 0.04,    # options.h:97
    "aspect_prior_sigma": 0.08,                # options.h:98
    # options.h:99-107 maxShared{RadialK1,K2,K3,TangentialP1,P2}Abs, SkewFraction, RadialK4, TangentialP3, P4
    "max_abs": [0.35, 0.35, 0.35, 0.02, 0.02, 0.10, 0.35, 0.35, 0.35],
    # options.h:108-116 the matching prior sigmas
    "prior_sigma": [0.15, 0.15, 0.15, 0.01, 0.01, 0.02, 0.15, 0.15, 0.15],
    "low_order_distortion_scale": 1.0,         # options.h:117 sharedLowOrderDistortionScale
    "use_transition_prior": 0,                 # options.h:118 useReferenceCalibrationTransitionPrior
    "previous_intrinsic_mask": 0,              # options.h:119 referencePreviousIntrinsicParameterMask{}
    "laser_plane_weight": 1.0,                 # options.h:124
    "laser_huber_delta": 0.2,                  # options.h:125
    "control_point_weight": 1.0,               # options.h:128
    "control_point_huber_delta": 0.2,          # options.h:129
    "device": 0,                               # options.h:152 plaMatrixDevice
    "precision": 0,                            # options.h:153 enablePlaMatrixMixedPrecision = false
    "preconditioner_cluster_size": 1,          # options.h:164 plaMatrixPreconditionerClusterSize
}


def test_default_options_are_the_reference_defaults(lib):
    o = _abi.default_options(lib)
    for name, want in REFERENCE_DEFAULTS.items():
        got = getattr(o, name)
        if isinstance(want, list):
            assert list(got) == want, name
        else:
            assert got == want, name
    # the oracle's defaults (test infrastructure) carry the same reference values
    from oracle import oracle as O

    ref = O.default_options()
    for name, want in REFERENCE_DEFAULTS.items():
        got = getattr(ref, name)
        assert (list(got) if isinstance(want, list) else got) == want, name
    # the structured Python options flatten to the same values
    from plabundle_amd import options as PO

    flat = PO.to_c(PO.SolveOptions())
    for name, want in REFERENCE_DEFAULTS.items():
        got = getattr(flat, name)
        assert (list(got) if isinstance(want, list) else got) == pytest.approx(want), name


def test_invalid_calls_fail_cleanly_without_a_device(lib):
    # no context -> invalid; a missing device is reported, never a crash
    assert lib.pb_hip_upload(None, None) == _abi.PB_ERR_INVALID
    assert lib.pb_hip_solve_batch(None, None, None, None, None, None) == _abi.PB_ERR_INVALID
    assert lib.pb_hip_batch_download(None, 0, None, None, None, None, None) == _abi.PB_ERR_INVALID
    assert lib.pb_hip_destroy(None) == _abi.PB_OK
    ctx = C.c_void_p()
    if not lib.pb_hip_available(0):
        assert lib.pb_hip_create(0, C.byref(ctx)) == _abi.PB_ERR_DEVICE


def test_batch_factor_shapes_and_tuning_queries(lib):
    # the launch shapes of the dense window factorisation: every shape covers the largest window system
    # (threads x rows per thread = 512) with an even panel width no wider than a wave; the tuning query needs a
    # context and a precision index 0 or 1
    t, p = C.c_int32(), C.c_int32()
    shapes = []
    for v in range(_abi.PB_BATCH_FACTOR_SHAPES):
        assert lib.pb_hip_batch_factor_shape(v, C.byref(t), C.byref(p)) == _abi.PB_OK
        assert t.value in (128, 256, 512) and 512 % t.value == 0
        assert p.value % 2 == 0 and 2 <= p.value <= 64
        shapes.append((t.value, p.value))
    assert shapes[0] == (256, 16) and len(set(shapes)) == len(shapes)
    assert lib.pb_hip_batch_factor_shape(_abi.PB_BATCH_FACTOR_SHAPES, C.byref(t), C.byref(p)) == _abi.PB_ERR_INVALID
    assert lib.pb_hip_batch_factor_shape(-1, None, None) == _abi.PB_ERR_INVALID
    shape = C.c_int32()
    assert lib.pb_hip_batch_tuning(None, 1, C.byref(shape), None) == _abi.PB_ERR_INVALID


def test_problem_fingerprint_sees_in_place_changes():
    # a workspace re-uploads when the fingerprint changes: every array edited in place, and the gauge scalars, must
    # change it, while an equal copy must not
    import numpy as np

    from plabundle_amd import synthetic as S
    from plabundle_amd.solver import problem_fingerprint

    p = S.ring_problem(seed=3, num_cameras=8, num_points=100, max_views=4)
    base = problem_fingerprint(p)
    assert problem_fingerprint(p.copy()) == base
    for name in ("cam_center", "cam_rotation", "cam_intrinsics", "track_point", "obs_uv"):
        a = getattr(p, name)
        keep = a.flat[0]
        a.flat[0] = keep + 1e-9
        assert problem_fingerprint(p) != base, name
        a.flat[0] = keep
        assert problem_fingerprint(p) == base, name
    q = p.copy()
    q.cam_center = np.asfortranarray(q.cam_center)  # same values, other layout: same upload
    assert problem_fingerprint(q) == base
