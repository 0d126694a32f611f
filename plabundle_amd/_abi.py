"""ctypes mirror of include/pb_hip.h (the C ABI of the MI355X backend).

Only data layouts and the library loader live here. The structures must stay
field-for-field identical to the header; tests/test_abi.py checks the symbol
list and the struct sizes against the header.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

INTRINSIC_COUNT = 13

# pb_status
PB_OK = 0
PB_ERR_INVALID = 1
PB_ERR_DEVICE = 2
PB_ERR_OOM = 3
PB_ERR_NUMERIC = 4
PB_ERR_CANCELLED = 5
PB_ERR_UNSUPPORTED = 6
PB_NOALLOC_MAX_ROWS = 1024  # include/pb_hip.h

STATUS_NAMES = {
    PB_OK: "ok",
    PB_ERR_INVALID: "invalid",
    PB_ERR_DEVICE: "device",
    PB_ERR_OOM: "out_of_memory",
    PB_ERR_NUMERIC: "numeric",
    PB_ERR_CANCELLED: "cancelled",
    PB_ERR_UNSUPPORTED: "unsupported",
}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_lp = C.POINTER(C.c_int64)


class ProblemView(C.Structure):
    _fields_ = [
        ("num_cameras", C.c_int32),
        ("cam_rotation", _dp),
        ("cam_center", _dp),
        ("cam_intrinsics", _dp),
        ("cam_convention", _ip),
        ("cam_image_size", _ip),
        ("ref_intrinsics", _dp),
        ("cam_group", _ip),
        ("num_tracks", C.c_int32),
        ("track_point", _dp),
        ("track_offset", _lp),
        ("obs_camera", _ip),
        ("obs_uv", _dp),
        ("obs_weight", _dp),
        ("obs_scale", _dp),
        ("num_fixed_cameras", C.c_int32),
        ("fixed_cameras", _ip),
        ("num_fixed_tracks", C.c_int32),
        ("fixed_tracks", _ip),
        ("gauge_policy", C.c_int32),
        ("gauge_anchor_camera", C.c_int32),
        ("gauge_scale_camera", C.c_int32),
        ("gauge_baseline", C.c_double),
        ("num_control_points", C.c_int32),
        ("cp_track", _ip),
        ("cp_point", _dp),
        ("cp_sigma", _dp),
        ("cp_weight", _dp),
        ("cp_uncertainty", _ip),
        ("cp_matrix", _dp),
        ("num_laser_planes", C.c_int32),
        ("lp_track", _ip),
        ("lp_point", _dp),
        ("lp_normal", _dp),
        ("lp_weight", _dp),
        ("num_rig_captures", C.c_int32),
        ("rig_capture_rotation", _dp),
        ("rig_capture_center", _dp),
        ("rig_capture_fixed", _ip),
        ("num_rig_sensors", C.c_int32),
        ("rig_sensor_rotation", _dp),
        ("rig_sensor_center", _dp),
        ("rig_sensor_fixed", _ip),
        ("cam_rig_capture", _ip),
        ("cam_rig_sensor", _ip),
    ]


class SolveOptionsC(C.Structure):
    _fields_ = [
        ("max_iterations", C.c_int32),
        ("robust_loss", C.c_int32),
        ("robust_loss_scale", C.c_double),
        ("armijo_coefficient", C.c_double),
        ("line_search_steps", C.c_int32),
        ("use_online_schur", C.c_int32),
        ("max_initial_track_rms", C.c_double),
        ("refine_camera_pose", C.c_int32),
        ("shared_intrinsic_mask", C.c_uint32),
        ("min_focal_scale", C.c_double),
        ("max_focal_scale", C.c_double),
        ("min_aspect_scale", C.c_double),
        ("max_aspect_scale", C.c_double),
        ("max_principal_offset_fraction", C.c_double),
        ("focal_prior_sigma", C.c_double),
        ("principal_prior_sigma_fraction", C.c_double),
        ("aspect_prior_sigma", C.c_double),
        ("max_abs", C.c_double * 9),
        ("prior_sigma", C.c_double * 9),
        ("low_order_distortion_scale", C.c_double),
        ("device", C.c_int32),
        ("precision", C.c_int32),
        ("pcg_relative_tolerance", C.c_double),
        ("pcg_absolute_tolerance", C.c_double),
        ("pcg_max_iterations", C.c_int32),
        ("pcg_check_interval", C.c_int32),
        ("armijo_batch", C.c_int32),
        ("profile_kernels", C.c_int32),
        ("pcg_model_eta", C.c_double),
        ("step_rms_tolerance", C.c_double),
        ("laser_plane_weight", C.c_double),
        ("laser_huber_delta", C.c_double),
        ("control_point_weight", C.c_double),
        ("control_point_huber_delta", C.c_double),
        ("use_transition_prior", C.c_int32),
        ("previous_intrinsic_mask", C.c_uint32),
        ("batch_graphs", C.c_int32),
        ("batch_factor_shape", C.c_int32),
        ("pcg_graphs", C.c_int32),
        ("reduced_solver", C.c_int32),
        ("dense_max_rows", C.c_int32),
        ("preconditioner_cluster_size", C.c_int32),
        ("mixed_switch_decrease", C.c_double),
        ("batch_factor_tiles", C.c_int32),
    ]


class SolveReportC(C.Structure):
    _fields_ = [
        ("solve_status", C.c_int32),
        ("solution_usable", C.c_int32),
        ("converged", C.c_int32),
        ("iterations", C.c_int32),
        ("initial_cost", C.c_double),
        ("final_cost", C.c_double),
        ("accepted_steps", C.c_int32),
        ("rejected_steps", C.c_int32),
        ("linearizations", C.c_int32),
        ("objective_evaluations", C.c_int32),
        ("linear_iterations", C.c_int32),
        ("rejected_initial_tracks", C.c_int32),
        ("observation_count", C.c_int32),
        ("active_track_count", C.c_int32),
        ("camera_block_count", C.c_int32),
        ("intrinsic_block_count", C.c_int32),
        ("track_block_count", C.c_int32),
        ("primary_block_count", C.c_int32),
        ("online_schur_used", C.c_int32),
        ("mixed_precision_used", C.c_int32),
        ("gauge_anchor_camera", C.c_int32),
        ("gauge_scale_camera", C.c_int32),
        ("gauge_baseline", C.c_double),
        ("setup_seconds", C.c_double),
        ("solve_seconds", C.c_double),
        ("assembly_seconds", C.c_double),
        ("linear_solve_seconds", C.c_double),
        ("back_substitution_seconds", C.c_double),
        ("objective_seconds", C.c_double),
        ("linear_tolerance", C.c_double),
        ("device_name", C.c_char * 64),
        ("message", C.c_char * 256),
        ("linear_solver_name", C.c_char * 32),
        ("preconditioner_name", C.c_char * 32),
        ("dense_fallbacks", C.c_int32),
        ("mixed_seed_steps", C.c_int32),
        ("schur_pattern_builds", C.c_int32),
        ("schur_pattern_reuses", C.c_int32),
        ("schur_assembly_on_device", C.c_int32),
        ("linear_tolerance_minimum", C.c_double),
        ("linear_tolerance_maximum", C.c_double),
        ("trial_state_seconds", C.c_double),
        ("cholesky_factorization_seconds", C.c_double),
        ("triangular_solve_seconds", C.c_double),
        ("residual_check_seconds", C.c_double),
    ]

    def as_dict(self) -> dict:
        out = {}
        for name, _ in self._fields_:
            value = getattr(self, name)
            if isinstance(value, bytes):
                value = value.decode("utf-8", "replace")
            out[name] = value
        return out


class IterationSummaryC(C.Structure):
    _fields_ = [
        ("current_iteration", C.c_int32),
        ("max_iterations", C.c_int32),
        ("average_rms_pixels", C.c_double),
        ("valid_point_count", C.c_int32),
    ]


PROGRESS_FN = C.CFUNCTYPE(C.c_int32, C.POINTER(IterationSummaryC), C.c_void_p)


class BatchView(C.Structure):
    _fields_ = [
        ("num_windows", C.c_int32),
        ("windows", C.POINTER(ProblemView)),
    ]


class WindowReportC(C.Structure):
    _fields_ = [
        ("solve_status", C.c_int32),
        ("solution_usable", C.c_int32),
        ("converged", C.c_int32),
        ("iterations", C.c_int32),
        ("accepted_steps", C.c_int32),
        ("rejected_steps", C.c_int32),
        ("linearizations", C.c_int32),
        ("objective_evaluations", C.c_int32),
        ("fp64_iterations", C.c_int32),
        ("observation_count", C.c_int32),
        ("active_track_count", C.c_int32),
        ("camera_block_count", C.c_int32),
        ("primary_block_count", C.c_int32),
        ("initial_cost", C.c_double),
        ("final_cost", C.c_double),
        ("device_seconds", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class QualityOptionsC(C.Structure):
    _fields_ = [
        ("enable_point_filter", C.c_int32),
        ("filter_max_reproj_error", C.c_double),
        ("filter_sigma_factor", C.c_double),
        ("enable_quality_gate", C.c_int32),
        ("max_accepted_rms_growth", C.c_double),
        ("min_accepted_valid_track_ratio", C.c_double),
    ]


class QualityReportC(C.Structure):
    _fields_ = [
        ("total_tracks", C.c_int32),
        ("optimized_tracks", C.c_int32),
        ("valid_track_ratio", C.c_double),
        ("mean_rms_before", C.c_double),
        ("mean_rms_after", C.c_double),
        ("median_rms_after", C.c_double),
        ("filter_threshold", C.c_double),
        ("solve_status", C.c_int32),
        ("solution_usable", C.c_int32),
        ("quality_gate_rejected", C.c_int32),
        ("message", C.c_char * 128),
    ]

    def as_dict(self) -> dict:
        out = {}
        for name, _ in self._fields_:
            value = getattr(self, name)
            out[name] = value.decode("utf-8", "replace") if isinstance(value, bytes) else value
        return out


class OutcomeView(C.Structure):
    _fields_ = [
        ("termination_status", C.c_int32),
        ("termination_reason", C.c_char_p),
        ("requested_backend", C.c_int32),
        ("used_backend", C.c_int32),
        ("solution_usable", C.c_int32),
        ("backend_fallback", C.c_int32),
        ("quality_gate_rejected", C.c_int32),
        ("mean_rms_before", C.c_double),
        ("mean_rms_after", C.c_double),
        ("num_cameras", C.c_int32),
        ("camera_ids", C.POINTER(C.c_char_p)),
        ("cam_rotation", _dp),
        ("cam_center", _dp),
        ("cam_intrinsics", _dp),
        ("cam_image_size", _ip),
        ("num_tracks", C.c_int32),
        ("track_ids", C.POINTER(C.c_char_p)),
        ("track_point", _dp),
        ("track_valid", _ip),
        ("track_rms_before", _dp),
        ("track_rms_after", _dp),
    ]


class ComparisonToleranceC(C.Structure):
    _fields_ = [
        ("rms", C.c_double),
        ("camera_center_meters", C.c_double),
        ("camera_rotation_degrees", C.c_double),
        ("point_meters", C.c_double),
        ("require_termination_status_match", C.c_int32),
        ("require_termination_reason_match", C.c_int32),
        ("require_requested_backend_match", C.c_int32),
        ("require_used_backend_match", C.c_int32),
    ]


class ComparisonC(C.Structure):
    _fields_ = [
        ("comparable", C.c_int32),
        ("equivalent", C.c_int32),
        ("termination_status_matches", C.c_int32),
        ("termination_reason_matches", C.c_int32),
        ("requested_backend_matches", C.c_int32),
        ("used_backend_matches", C.c_int32),
        ("rms_before_difference", C.c_double),
        ("rms_after_difference", C.c_double),
        ("maximum_camera_center_difference_meters", C.c_double),
        ("maximum_camera_rotation_difference_degrees", C.c_double),
        ("maximum_point_difference_meters", C.c_double),
        ("valid_track_mask_mismatches", C.c_int32),
        ("message", C.c_char * 128),
    ]

    def as_dict(self) -> dict:
        out = {}
        for name, _ in self._fields_:
            value = getattr(self, name)
            out[name] = value.decode("utf-8", "replace") if isinstance(value, bytes) else value
        return out


PB_BATCH_FACTOR_SHAPES = 6

# Exported symbols of the shared library, in header order.
EXPORTED_SYMBOLS = (
    "pb_default_options",
    "pb_hip_available",
    "pb_hip_create",
    "pb_hip_destroy",
    "pb_hip_upload",
    "pb_hip_solve",
    "pb_hip_download",
    "pb_hip_get_masks",
    "pb_hip_get_rig_masks",
    "pb_hip_download_rig",
    "pb_hip_kernel_stats",
    "pb_hip_reduced_system",
    "pb_hip_evaluate_cost",
    "pb_hip_backward",
    "pb_hip_backward_pose",
    "pb_hip_solve_batch",
    "pb_hip_batch_factor_shape",
    "pb_hip_batch_tuning",
    "pb_hip_batch_download",
    "pb_default_quality_options",
    "pb_hip_quality",
    "pb_default_comparison_tolerance",
    "pb_validate_outcome",
    "pb_compare_outcomes",
    "pb_hip_set_progress",
    "pb_hip_dense_solve",
    "pb_hip_direct_solve",
    "pb_hip_dense_solve_noalloc",
    "pb_hip_last_error",
)

PACKAGE_DIR = os.path.dirname(os.path.abspath(__file__))
LIBRARY_PATH = os.path.join(PACKAGE_DIR, "libpb_hip.so")

_lib = None


class BackendLibraryMissing(RuntimeError):
    """The HIP backend library has not been built (run __graft_entry__.build())."""


def load_library(path: str | None = None):
    """Loads libpb_hip.so from the package directory; fails loudly when it is absent."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    lib_path = path or os.environ.get("PLABUNDLE_AMD_LIBRARY") or LIBRARY_PATH
    if not os.path.exists(lib_path):
        raise BackendLibraryMissing(f"{lib_path} not found; build the HIP backend first")
    lib = C.CDLL(lib_path, mode=C.RTLD_LOCAL)
    lib.pb_default_options.argtypes = [C.POINTER(SolveOptionsC)]
    lib.pb_default_options.restype = None
    lib.pb_hip_available.argtypes = [C.c_int]
    lib.pb_hip_available.restype = C.c_int
    lib.pb_hip_create.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
    lib.pb_hip_create.restype = C.c_int
    lib.pb_hip_destroy.argtypes = [C.c_void_p]
    lib.pb_hip_destroy.restype = C.c_int
    lib.pb_hip_upload.argtypes = [C.c_void_p, C.POINTER(ProblemView)]
    lib.pb_hip_upload.restype = C.c_int
    lib.pb_hip_solve.argtypes = [C.c_void_p, C.POINTER(SolveOptionsC), C.POINTER(SolveReportC), _ip]
    lib.pb_hip_solve.restype = C.c_int
    lib.pb_hip_download.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _ip]
    lib.pb_hip_download.restype = C.c_int
    lib.pb_hip_get_masks.argtypes = [C.c_void_p, _ip, _ip, _ip, _ip, _ip]
    lib.pb_hip_get_masks.restype = C.c_int
    lib.pb_hip_get_rig_masks.argtypes = [C.c_void_p, _ip, _ip]
    lib.pb_hip_get_rig_masks.restype = C.c_int
    lib.pb_hip_download_rig.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
    lib.pb_hip_download_rig.restype = C.c_int
    lib.pb_hip_kernel_stats.argtypes = [C.c_void_p, C.c_char_p, _lp, _dp, _dp]
    lib.pb_hip_kernel_stats.restype = C.c_int
    lib.pb_hip_reduced_system.argtypes = [C.c_void_p, C.POINTER(SolveOptionsC), C.c_double, _dp, _dp, _dp, _ip]
    lib.pb_hip_reduced_system.restype = C.c_int
    lib.pb_hip_evaluate_cost.argtypes = [C.c_void_p, C.POINTER(SolveOptionsC), _dp]
    lib.pb_hip_evaluate_cost.restype = C.c_int
    lib.pb_hip_backward.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
    lib.pb_hip_backward.restype = C.c_int
    lib.pb_hip_backward_pose.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp]
    lib.pb_hip_backward_pose.restype = C.c_int
    lib.pb_hip_solve_batch.argtypes = [C.c_void_p, C.POINTER(BatchView), C.POINTER(SolveOptionsC),
                                       C.POINTER(WindowReportC), C.POINTER(SolveReportC), _ip]
    lib.pb_hip_solve_batch.restype = C.c_int
    lib.pb_hip_batch_factor_shape.argtypes = [C.c_int32, _ip, _ip]
    lib.pb_hip_batch_factor_shape.restype = C.c_int
    lib.pb_hip_batch_tuning.argtypes = [C.c_void_p, C.c_int32, _ip, _dp]
    lib.pb_hip_batch_tuning.restype = C.c_int
    lib.pb_hip_batch_download.argtypes = [C.c_void_p, C.c_int32, _dp, _dp, _dp, _dp, _ip]
    lib.pb_hip_batch_download.restype = C.c_int
    lib.pb_default_quality_options.argtypes = [C.POINTER(QualityOptionsC)]
    lib.pb_default_quality_options.restype = None
    lib.pb_hip_quality.argtypes = [C.c_void_p, C.POINTER(QualityOptionsC), C.POINTER(QualityReportC), _dp, _dp, _ip]
    lib.pb_hip_quality.restype = C.c_int
    lib.pb_default_comparison_tolerance.argtypes = [C.POINTER(ComparisonToleranceC)]
    lib.pb_default_comparison_tolerance.restype = None
    lib.pb_validate_outcome.argtypes = [C.POINTER(OutcomeView), C.c_char_p, C.c_int32]
    lib.pb_validate_outcome.restype = C.c_int
    lib.pb_compare_outcomes.argtypes = [C.POINTER(OutcomeView), C.POINTER(OutcomeView),
                                        C.POINTER(ComparisonToleranceC), C.POINTER(ComparisonC), _dp, _dp, _dp, _ip]
    lib.pb_compare_outcomes.restype = C.c_int
    lib.pb_hip_set_progress.argtypes = [C.c_void_p, PROGRESS_FN, C.c_void_p]
    lib.pb_hip_set_progress.restype = C.c_int
    lib.pb_hip_direct_solve.argtypes = [C.c_void_p, C.POINTER(SolveOptionsC), C.c_double, C.c_int32, _dp, _dp, _dp,
                                        _ip, _dp, C.c_int64]
    lib.pb_hip_direct_solve.restype = C.c_int
    lib.pb_hip_dense_solve.argtypes = [C.c_void_p, C.c_int32, C.c_int32, _dp, _dp, _dp, _ip, _dp]
    lib.pb_hip_dense_solve.restype = C.c_int
    lib.pb_hip_dense_solve_noalloc.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, _ip]
    lib.pb_hip_dense_solve_noalloc.restype = C.c_int
    lib.pb_hip_last_error.argtypes = [C.c_void_p]
    lib.pb_hip_last_error.restype = C.c_char_p
    if path is None:
        _lib = lib
    return lib


def default_options(lib=None) -> SolveOptionsC:
    lib = lib or load_library()
    opts = SolveOptionsC()
    lib.pb_default_options(C.byref(opts))
    return opts


def ptr(array, ctype):
    """Pointer to a contiguous numpy array (or NULL for None)."""
    if array is None:
        return C.cast(None, C.POINTER(ctype))
    return array.ctypes.data_as(C.POINTER(ctype))


def make_view(problem) -> tuple[ProblemView, list]:
    """Builds a pb_problem_view over the arrays of a plabundle_amd.problem.Problem.

    Returns the view and the list of arrays it points into (keep it alive)."""
    keep = []

    def arr(a, dtype):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=dtype)
        keep.append(a)
        return a

    rot = arr(problem.cam_rotation, np.float64)
    cen = arr(problem.cam_center, np.float64)
    intr = arr(problem.cam_intrinsics, np.float64)
    conv = arr(problem.cam_convention, np.int32)
    size = arr(problem.cam_image_size, np.int32)
    ref = arr(problem.ref_intrinsics, np.float64)
    grp = arr(problem.cam_group, np.int32)
    pts = arr(problem.track_point, np.float64)
    off = arr(problem.track_offset, np.int64)
    ocam = arr(problem.obs_camera, np.int32)
    ouv = arr(problem.obs_uv, np.float64)
    ow = arr(problem.obs_weight, np.float64)
    osc = arr(problem.obs_scale, np.float64)
    fc = arr(problem.fixed_cameras, np.int32)
    ft = arr(problem.fixed_tracks, np.int32)
    v = ProblemView()
    v.num_cameras = int(rot.shape[0])
    v.cam_rotation = ptr(rot, C.c_double)
    v.cam_center = ptr(cen, C.c_double)
    v.cam_intrinsics = ptr(intr, C.c_double)
    v.cam_convention = ptr(conv, C.c_int32)
    v.cam_image_size = ptr(size, C.c_int32)
    v.ref_intrinsics = ptr(ref, C.c_double)
    v.cam_group = ptr(grp, C.c_int32)
    v.num_tracks = int(pts.shape[0])
    v.track_point = ptr(pts, C.c_double)
    v.track_offset = ptr(off, C.c_int64)
    v.obs_camera = ptr(ocam, C.c_int32)
    v.obs_uv = ptr(ouv, C.c_double)
    v.obs_weight = ptr(ow, C.c_double)
    v.obs_scale = ptr(osc, C.c_double)
    v.num_fixed_cameras = 0 if fc is None else int(fc.size)
    v.fixed_cameras = ptr(fc, C.c_int32)
    v.num_fixed_tracks = 0 if ft is None else int(ft.size)
    v.fixed_tracks = ptr(ft, C.c_int32)
    v.gauge_policy = int(problem.gauge_policy)
    v.gauge_anchor_camera = int(problem.gauge_anchor_camera)
    v.gauge_scale_camera = int(problem.gauge_scale_camera)
    v.gauge_baseline = float(problem.gauge_baseline)
    cp_track = arr(getattr(problem, "cp_track", None), np.int32)
    v.num_control_points = 0 if cp_track is None else int(cp_track.size)
    v.cp_track = ptr(cp_track, C.c_int32)
    v.cp_point = ptr(arr(getattr(problem, "cp_point", None), np.float64), C.c_double)
    v.cp_sigma = ptr(arr(getattr(problem, "cp_sigma", None), np.float64), C.c_double)
    v.cp_weight = ptr(arr(getattr(problem, "cp_weight", None), np.float64), C.c_double)
    v.cp_uncertainty = ptr(arr(getattr(problem, "cp_uncertainty", None), np.int32), C.c_int32)
    v.cp_matrix = ptr(arr(getattr(problem, "cp_matrix", None), np.float64), C.c_double)
    lp_track = arr(getattr(problem, "lp_track", None), np.int32)
    v.num_laser_planes = 0 if lp_track is None else int(lp_track.size)
    v.lp_track = ptr(lp_track, C.c_int32)
    v.lp_point = ptr(arr(getattr(problem, "lp_point", None), np.float64), C.c_double)
    v.lp_normal = ptr(arr(getattr(problem, "lp_normal", None), np.float64), C.c_double)
    v.lp_weight = ptr(arr(getattr(problem, "lp_weight", None), np.float64), C.c_double)
    kr = arr(getattr(problem, "rig_capture_rotation", None), np.float64)
    sr = arr(getattr(problem, "rig_sensor_rotation", None), np.float64)
    v.num_rig_captures = 0 if kr is None else int(kr.size // 9)
    v.rig_capture_rotation = ptr(kr, C.c_double)
    v.rig_capture_center = ptr(arr(getattr(problem, "rig_capture_center", None), np.float64), C.c_double)
    v.rig_capture_fixed = ptr(arr(getattr(problem, "rig_capture_fixed", None), np.int32), C.c_int32)
    v.num_rig_sensors = 0 if sr is None else int(sr.size // 9)
    v.rig_sensor_rotation = ptr(sr, C.c_double)
    v.rig_sensor_center = ptr(arr(getattr(problem, "rig_sensor_center", None), np.float64), C.c_double)
    v.rig_sensor_fixed = ptr(arr(getattr(problem, "rig_sensor_fixed", None), np.int32), C.c_int32)
    v.cam_rig_capture = ptr(arr(getattr(problem, "cam_rig_capture", None), np.int32), C.c_int32)
    v.cam_rig_sensor = ptr(arr(getattr(problem, "cam_rig_sensor", None), np.int32), C.c_int32)
    return v, keep
