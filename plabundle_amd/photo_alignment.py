"""Id-keyed problems and outcomes for legacy/new comparisons, and their dump files.

Mirrors include/plabundle/photo_alignment.h (PhotoAlignmentProblem, PhotoAlignmentOutcome,
PhotoAlignmentComparisonTolerance, PhotoAlignmentComparison and the make/validate/compare functions of
src/photo_alignment.cpp:134-318). Validation and comparison run in the C ABI (pb_validate_outcome,
pb_compare_outcomes); this module builds the views and reads/writes the dump format.

Dump format (numpy .npz, loaded with allow_pickle=False): a "format" entry naming the kind and version
("plabundle_amd.problem/1" or "plabundle_amd.outcome/1"), the SoA arrays of pb_problem_view / pb_outcome_view
under their C field names, camera and track ids as unicode arrays, scalar fields as 0-d arrays.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field, fields

import numpy as np

from . import _abi
from .problem import Problem

PROBLEM_FORMAT = "plabundle_amd.problem/1"
OUTCOME_FORMAT = "plabundle_amd.outcome/1"


@dataclass
class PhotoAlignmentProblem:
    problem: Problem
    cameraIds: list[str]
    trackIds: list[str]


@dataclass
class PhotoAlignmentOutcome:
    terminationStatus: int = 0
    terminationReason: str = ""
    requestedBackend: int = 0
    usedBackend: int = 0
    backendSelectionReason: str = ""
    solutionUsable: bool = False
    backendFallback: bool = False
    qualityGateRejected: bool = False
    meanRmsBefore: float = 0.0
    meanRmsAfter: float = 0.0
    camera_ids: list[str] = field(default_factory=list)
    cam_rotation: np.ndarray = field(default_factory=lambda: np.zeros((0, 9)))
    cam_center: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    cam_intrinsics: np.ndarray = field(default_factory=lambda: np.zeros((0, 13)))
    cam_image_size: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), dtype=np.int32))
    track_ids: list[str] = field(default_factory=list)
    track_point: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    track_valid: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.int32))
    track_rms_before: np.ndarray = field(default_factory=lambda: np.zeros(0))
    track_rms_after: np.ndarray = field(default_factory=lambda: np.zeros(0))


@dataclass
class PhotoAlignmentComparisonTolerance:
    rms: float = 1.0e-8
    cameraCenterMeters: float = 1.0e-8
    cameraRotationDegrees: float = 1.0e-8
    pointMeters: float = 1.0e-8
    requireTerminationStatusMatch: bool = True
    requireTerminationReasonMatch: bool = True
    requireRequestedBackendMatch: bool = True
    requireUsedBackendMatch: bool = True


def validatePhotoAlignmentProblem(p: PhotoAlignmentProblem) -> tuple[bool, str]:
    """validatePhotoAlignmentProblem (src/photo_alignment.cpp:116-132): ids align with the problem, are non-empty
    and unique."""
    if len(p.cameraIds) != p.problem.num_cameras:
        return False, "photo-alignment camera ids must align with problem cameras"
    if len(p.trackIds) != p.problem.num_tracks:
        return False, "photo-alignment track ids must align with problem tracks"
    for label, ids in (("camera", p.cameraIds), ("track", p.trackIds)):
        seen = set()
        for i in ids:
            if not i:
                return False, f"{label} ids must not be empty"
            if i in seen:
                return False, f"{label} ids must be unique: {i}"
            seen.add(i)
    return True, ""


def makePhotoAlignmentOutcome(p: PhotoAlignmentProblem, result) -> PhotoAlignmentOutcome:
    """makePhotoAlignmentOutcome (src/photo_alignment.cpp:134-183) from a Solver result that carries the
    post-solve quality (Solver.solve(..., quality=True))."""
    ok, err = validatePhotoAlignmentProblem(p)
    if not ok:
        raise ValueError(err)
    if result.cam_rotation is None or result.quality is None:
        raise ValueError("the result must carry the refined state and the post-solve quality")
    if result.cam_rotation.shape[0] != len(p.cameraIds) or result.track_point.shape[0] != len(p.trackIds):
        raise ValueError("solver result camera/track counts do not match the photo-alignment problem")
    q = result.quality
    size = p.problem.cam_image_size
    return PhotoAlignmentOutcome(
        terminationStatus=int(result.status), terminationReason=result.backendMessage,
        requestedBackend=int(result.requestedBackend), usedBackend=int(result.usedBackend),
        solutionUsable=bool(result.solutionUsable), qualityGateRejected=bool(result.qualityGateRejected),
        meanRmsBefore=float(q["mean_rms_before"]), meanRmsAfter=float(q["mean_rms_after"]),
        camera_ids=list(p.cameraIds), cam_rotation=np.asarray(result.cam_rotation, dtype=np.float64).reshape(-1, 9),
        cam_center=np.asarray(result.cam_center, dtype=np.float64),
        cam_intrinsics=np.asarray(result.cam_intrinsics, dtype=np.float64),
        cam_image_size=(np.asarray(size, dtype=np.int32) if size is not None
                        else np.zeros((len(p.cameraIds), 2), dtype=np.int32)),
        track_ids=list(p.trackIds), track_point=np.asarray(result.track_point, dtype=np.float64),
        track_valid=np.asarray(q["track_valid"], dtype=np.int32),
        track_rms_before=np.asarray(q["track_rms_before"], dtype=np.float64),
        track_rms_after=np.asarray(q["track_rms_after"], dtype=np.float64))


def _view(o: PhotoAlignmentOutcome):
    keep = []

    def arr(a, dtype, cols=None):
        a = np.ascontiguousarray(a, dtype=dtype)
        keep.append(a)
        return a

    def ids(values):
        buf = (C.c_char_p * max(1, len(values)))(*[v.encode() for v in values])
        keep.append(buf)
        return C.cast(buf, C.POINTER(C.c_char_p))

    v = _abi.OutcomeView()
    v.termination_status = int(o.terminationStatus)
    reason = o.terminationReason.encode()
    keep.append(reason)
    v.termination_reason = reason
    v.requested_backend = int(o.requestedBackend)
    v.used_backend = int(o.usedBackend)
    v.solution_usable = int(bool(o.solutionUsable))
    v.backend_fallback = int(bool(o.backendFallback))
    v.quality_gate_rejected = int(bool(o.qualityGateRejected))
    v.mean_rms_before = float(o.meanRmsBefore)
    v.mean_rms_after = float(o.meanRmsAfter)
    v.num_cameras = len(o.camera_ids)
    v.camera_ids = ids(o.camera_ids)
    v.cam_rotation = _abi.ptr(arr(o.cam_rotation, np.float64), C.c_double)
    v.cam_center = _abi.ptr(arr(o.cam_center, np.float64), C.c_double)
    v.cam_intrinsics = _abi.ptr(arr(o.cam_intrinsics, np.float64), C.c_double)
    v.cam_image_size = _abi.ptr(arr(o.cam_image_size, np.int32), C.c_int32)
    v.num_tracks = len(o.track_ids)
    v.track_ids = ids(o.track_ids)
    v.track_point = _abi.ptr(arr(o.track_point, np.float64), C.c_double)
    v.track_valid = _abi.ptr(arr(o.track_valid, np.int32), C.c_int32)
    v.track_rms_before = _abi.ptr(arr(o.track_rms_before, np.float64), C.c_double)
    v.track_rms_after = _abi.ptr(arr(o.track_rms_after, np.float64), C.c_double)
    return v, keep


def validatePhotoAlignmentOutcome(o: PhotoAlignmentOutcome) -> tuple[bool, str]:
    """validatePhotoAlignmentOutcome (src/photo_alignment.cpp:185-223), through pb_validate_outcome."""
    lib = _abi.load_library()
    v, keep = _view(o)
    err = C.create_string_buffer(256)
    st = lib.pb_validate_outcome(C.byref(v), err, 256)
    return st == _abi.PB_OK, err.value.decode()


def comparePhotoAlignmentOutcomes(reference: PhotoAlignmentOutcome, candidate: PhotoAlignmentOutcome,
                                  tolerance: PhotoAlignmentComparisonTolerance | None = None) -> dict:
    """comparePhotoAlignmentOutcomes (src/photo_alignment.cpp:225-318), through pb_compare_outcomes. Returns the
    comparison fields plus per-id arrays in the reference's order; "status" is the ABI status (invalid outcomes
    or differing id sets give PB_ERR_INVALID with comparable = 0)."""
    lib = _abi.load_library()
    t = tolerance or PhotoAlignmentComparisonTolerance()
    tc = _abi.ComparisonToleranceC(t.rms, t.cameraCenterMeters, t.cameraRotationDegrees, t.pointMeters,
                                   int(t.requireTerminationStatusMatch), int(t.requireTerminationReasonMatch),
                                   int(t.requireRequestedBackendMatch), int(t.requireUsedBackendMatch))
    rv, k1 = _view(reference)
    cv, k2 = _view(candidate)
    nc, nt = len(reference.camera_ids), len(reference.track_ids)
    center, rotation = np.zeros(nc), np.zeros(nc)
    point, mismatch = np.zeros(nt), np.zeros(nt, dtype=np.int32)
    out = _abi.ComparisonC()
    st = lib.pb_compare_outcomes(C.byref(rv), C.byref(cv), C.byref(tc), C.byref(out), _abi.ptr(center, C.c_double),
                                 _abi.ptr(rotation, C.c_double), _abi.ptr(point, C.c_double),
                                 _abi.ptr(mismatch, C.c_int32))
    res = out.as_dict()
    res.update(status=st, camera_center_difference=center, camera_rotation_difference_degrees=rotation,
               point_difference=point, valid_mismatch=mismatch,
               validTrackMaskMismatchIds=[reference.track_ids[i] for i in np.nonzero(mismatch)[0]])
    return res


# --- dump files --------------------------------------------------------------------------------
_PROBLEM_ARRAYS = ("cam_rotation", "cam_center", "cam_intrinsics", "cam_convention", "cam_image_size",
                   "ref_intrinsics", "cam_group", "track_point", "track_offset", "obs_camera", "obs_uv", "obs_weight",
                   "obs_scale", "fixed_cameras", "fixed_tracks")


def save_problem(path: str, p: PhotoAlignmentProblem) -> None:
    ok, err = validatePhotoAlignmentProblem(p)
    if not ok:
        raise ValueError(err)
    data = {"format": np.array(PROBLEM_FORMAT), "camera_ids": np.array(p.cameraIds, dtype=str),
            "track_ids": np.array(p.trackIds, dtype=str)}
    for name in _PROBLEM_ARRAYS:
        value = getattr(p.problem, name)
        if value is not None:
            data[name] = np.asarray(value)
    for name in ("gauge_policy", "gauge_anchor_camera", "gauge_scale_camera", "gauge_baseline"):
        data[name] = np.array(getattr(p.problem, name))
    np.savez(path, **data)


def load_problem(path: str) -> PhotoAlignmentProblem:
    with np.load(path, allow_pickle=False) as z:
        if str(z["format"]) != PROBLEM_FORMAT:
            raise ValueError(f"not a {PROBLEM_FORMAT} file")
        kwargs = {name: z[name].copy() for name in _PROBLEM_ARRAYS if name in z.files}
        kwargs["gauge_policy"] = int(z["gauge_policy"])
        kwargs["gauge_anchor_camera"] = int(z["gauge_anchor_camera"])
        kwargs["gauge_scale_camera"] = int(z["gauge_scale_camera"])
        kwargs["gauge_baseline"] = float(z["gauge_baseline"])
        return PhotoAlignmentProblem(Problem(**kwargs), [str(s) for s in z["camera_ids"]],
                                     [str(s) for s in z["track_ids"]])


def save_outcome(path: str, o: PhotoAlignmentOutcome) -> None:
    data = {"format": np.array(OUTCOME_FORMAT)}
    for f in fields(o):
        value = getattr(o, f.name)
        data[f.name] = np.array(value, dtype=str) if f.name in ("camera_ids", "track_ids") else np.asarray(value)
    np.savez(path, **data)


def load_outcome(path: str) -> PhotoAlignmentOutcome:
    with np.load(path, allow_pickle=False) as z:
        if str(z["format"]) != OUTCOME_FORMAT:
            raise ValueError(f"not a {OUTCOME_FORMAT} file")
        kwargs = {}
        for f in fields(PhotoAlignmentOutcome):
            value = z[f.name]
            if f.name in ("camera_ids", "track_ids"):
                kwargs[f.name] = [str(s) for s in value]
            elif value.ndim == 0:
                kwargs[f.name] = value.item()
            else:
                kwargs[f.name] = value.copy()
        return PhotoAlignmentOutcome(**kwargs)
