"""Python host mirror of plabundle::Solver for the MI355X backend.

Solver.solve(problem, options, workspace) follows include/plabundle/solver.h:34-54:
the problem is immutable, the result is returned by value and owns the refined
state, the workspace is caller-owned and must not be shared by concurrent
solves, and errors travel in Result.status plus a message (only a missing HIP
library raises, because there is no silent CPU path here).
"""
from __future__ import annotations

import ctypes as C
import enum
import hashlib
from dataclasses import dataclass, field

import numpy as np

from . import _abi
from .options import Backend, Precision, SolveOptions, to_c
from .problem import Problem


class SolveStatus(enum.IntEnum):
    """plabundle::SolveStatus (include/plabundle/backend.h:17-27)."""

    NotRun = 0
    Success = 1
    NoConvergence = 2
    Cancelled = 3
    InvalidInput = 4
    UnsupportedConfiguration = 5
    BackendUnavailable = 6
    NumericalFailure = 7


@dataclass
class Result:
    requestedBackend: Backend = Backend.Mi355x
    usedBackend: Backend = Backend.Mi355x
    status: SolveStatus = SolveStatus.NotRun
    solutionUsable: bool = False
    usedGpu: bool = False
    backendMessage: str = ""
    observationCount: int = 0
    report: dict = field(default_factory=dict)
    cam_rotation: np.ndarray | None = None
    cam_center: np.ndarray | None = None
    cam_intrinsics: np.ndarray | None = None
    track_point: np.ndarray | None = None
    track_valid: np.ndarray | None = None
    abi_status: int = _abi.PB_OK
    quality: dict | None = None
    qualityGateRejected: bool = False
    qualityGateMessage: str = ""
    # Result::refinedRig (result.h:149): capture rotations / centres and sensor rotations / centres, or None
    refinedRig: dict | None = None

    def usable(self) -> bool:
        return self.solutionUsable and self.status in (SolveStatus.Success, SolveStatus.NoConvergence)

    @property
    def initialCost(self) -> float:
        return self.report.get("initial_cost", float("nan"))

    @property
    def finalCost(self) -> float:
        return self.report.get("final_cost", float("nan"))


@dataclass
class IterationSummary:
    """plabundle::IterationSummary (include/plabundle/options.h:50-56)."""

    currentIteration: int = 0
    maxIterations: int = 0
    averageRmsPixels: float = 0.0
    validPointCount: int = 0


def problem_fingerprint(problem: Problem) -> bytes:
    """Digest of everything pb_hip_upload copies from a problem: every array (bytes, dtype, shape) and the gauge
    scalars. A workspace re-uploads whenever it changes, so arrays mutated in place are never solved from a stale
    device copy."""
    try:
        import xxhash

        h = xxhash.xxh3_128()
    except ImportError:  # pragma: no cover - xxhash ships with the image
        h = hashlib.blake2b(digest_size=16)
    for name in problem.__dataclass_fields__:
        if name == "meta":
            continue
        v = getattr(problem, name)
        if isinstance(v, np.ndarray):
            a = np.ascontiguousarray(v)
            h.update(f"{name}:{a.dtype.str}:{a.shape}".encode())
            h.update(a.view(np.uint8).reshape(-1) if a.size else b"")
        else:
            h.update(f"{name}={v!r}".encode())
    return h.digest()


def factor_shape(shape: int):
    """(threads, panel width) of a launch shape of the batched dense window factorisation."""
    t, p = C.c_int32(), C.c_int32()
    if _abi.load_library().pb_hip_batch_factor_shape(int(shape), C.byref(t), C.byref(p)) != _abi.PB_OK:
        raise ValueError(f"no factorisation shape {shape}")
    return t.value, p.value


class SolverWorkspace:
    """Reusable device workspace (the SolverWorkspace analogue, include/plabundle/solver.h:13-32)."""

    def __init__(self, device: int = 0):
        self.device = device
        self._lib = _abi.load_library()
        self._ctx = C.c_void_p()
        self._uploaded = None  # fingerprint of the problem on the context
        self._counts = None    # (cameras, tracks, observations, control points) of that problem
        self._num_tracks = 0
        st = self._lib.pb_hip_create(device, C.byref(self._ctx))
        if st != _abi.PB_OK:
            self._ctx = None
            raise RuntimeError(f"pb_hip_create failed: {_abi.STATUS_NAMES.get(st, st)}")

    @property
    def handle(self):
        return self._ctx

    def last_error(self) -> str:
        return self._lib.pb_hip_last_error(self._ctx).decode("utf-8", "replace") if self._ctx else "closed"

    def upload(self, problem: Problem, fingerprint: bytes | None = None) -> int:
        view, keep = _abi.make_view(problem)
        st = self._lib.pb_hip_upload(self._ctx, C.byref(view))
        del keep
        if st == _abi.PB_OK:
            self._uploaded = fingerprint or problem_fingerprint(problem)
            self._counts = (problem.num_cameras, problem.num_tracks, problem.num_observations,
                            problem.num_control_points)
            self._num_tracks = problem.num_tracks
        else:
            self._uploaded = self._counts = None
        return st

    def ensure_uploaded(self, problem: Problem) -> int:
        """Uploads the problem unless the context already holds one with the same content."""
        fp = problem_fingerprint(problem)
        if self._uploaded == fp:
            return _abi.PB_OK
        return self.upload(problem, fp)

    def clear(self):
        self._uploaded = self._counts = None

    def close(self):
        if self._ctx:
            self._lib.pb_hip_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- inspection entry points used by parity tests -------------------------------------
    def reduced_system(self, copts, damping: float, primary_blocks: int):
        n = primary_blocks * 13
        A = np.zeros((n, n))
        b = np.zeros(n)
        cost = C.c_double()
        P = C.c_int32()
        st = self._lib.pb_hip_reduced_system(self._ctx, C.byref(copts), float(damping), _abi.ptr(A, C.c_double),
                                             _abi.ptr(b, C.c_double), C.byref(cost), C.byref(P))
        if st != _abi.PB_OK:
            raise RuntimeError(f"pb_hip_reduced_system: {self.last_error()}")
        return A, b, cost.value

    def batch_tuning(self) -> dict:
        """Factorisation-shape timing of the batches solved on this workspace (pb_hip_batch_tuning): per
        assembly precision the chosen shape (threads, panel) and the measured milliseconds of every shape."""
        out = {}
        for name, fp64 in (("fp32", 0), ("fp64", 1)):
            shape = C.c_int32()
            ms = np.zeros(_abi.PB_BATCH_FACTOR_SHAPES)
            if self._lib.pb_hip_batch_tuning(self._ctx, fp64, C.byref(shape), _abi.ptr(ms, C.c_double)) != _abi.PB_OK:
                continue
            if shape.value < 0:
                out[name] = None
                continue
            out[name] = {"shape": shape.value, "threads_panel": factor_shape(shape.value),
                         "ms": {f"{t}x{p}": round(float(m), 4) for (t, p), m in
                                zip((factor_shape(v) for v in range(_abi.PB_BATCH_FACTOR_SHAPES)), ms) if m >= 0}}
        return out

    def evaluate_cost(self, copts) -> float:
        cost = C.c_double()
        st = self._lib.pb_hip_evaluate_cost(self._ctx, C.byref(copts), C.byref(cost))
        if st != _abi.PB_OK:
            raise RuntimeError(f"pb_hip_evaluate_cost: {self.last_error()}")
        return cost.value

    def masks(self, problem: Problem) -> dict:
        nc, nt = problem.num_cameras, problem.num_tracks
        out = {k: np.zeros(n, dtype=np.int32) for k, n in (
            ("camera_block", nc), ("intrinsic_block_by_camera", nc), ("calibration_group_by_camera", nc),
            ("track_block", nt), ("active_track", nt))}
        st = self._lib.pb_hip_get_masks(self._ctx, *(_abi.ptr(out[k], C.c_int32) for k in (
            "camera_block", "intrinsic_block_by_camera", "calibration_group_by_camera", "track_block",
            "active_track")))
        if st != _abi.PB_OK:
            raise RuntimeError(f"pb_hip_get_masks: {self.last_error()}")
        return out

    def rig_masks(self, problem: Problem) -> dict:
        """ActiveProblem::rigCaptureBlockByCamera and rigSensorBlockByCamera of the last solve."""
        nc = problem.num_cameras
        out = {"capture_block": np.zeros(nc, dtype=np.int32), "sensor_block": np.zeros(nc, dtype=np.int32)}
        st = self._lib.pb_hip_get_rig_masks(self._ctx, _abi.ptr(out["capture_block"], C.c_int32),
                                            _abi.ptr(out["sensor_block"], C.c_int32))
        if st != _abi.PB_OK:
            raise RuntimeError(f"pb_hip_get_rig_masks: {self.last_error()}")
        return out

    def rig_state(self, problem: Problem) -> dict:
        """The refined rig of the last solve (the input rig when nothing was published)."""
        out = {"capture_rotation": np.zeros((problem.num_rig_captures, 9)),
               "capture_center": np.zeros((problem.num_rig_captures, 3)),
               "sensor_rotation": np.zeros((problem.num_rig_sensors, 9)),
               "sensor_center": np.zeros((problem.num_rig_sensors, 3))}
        st = self._lib.pb_hip_download_rig(self._ctx, *(_abi.ptr(out[k], C.c_double) for k in (
            "capture_rotation", "capture_center", "sensor_rotation", "sensor_center")))
        if st != _abi.PB_OK:
            raise RuntimeError(f"pb_hip_download_rig: {self.last_error()}")
        return out

    def backward(self, problem: Problem, dl_dcenter=None, dl_dpoint=None, with_control: bool = False,
                 dl_drotation=None):
        """Adjoint of the last solve on this workspace: dl/d(u, v) per observation ([num_observations, 2])
        for a loss with gradients dl/dC ([num_cameras, 3]), dl/dR ([num_cameras, 3, 3] or [num_cameras, 9],
        camera-to-world rotations as returned) and dl/dX ([num_tracks, 3]) at the refined state.
        with_control=True also returns dl/dP of every control point's surveyed position ([num_control_points, 3])."""
        # pb_hip_backward reads and writes arrays sized by the problem on the context: refuse any other problem
        if self._uploaded is None or self._uploaded != problem_fingerprint(problem):
            raise ValueError("backward() needs the problem of the last solve on this workspace")
        nc, nt, no, ncp = self._counts
        dc = None if dl_dcenter is None else np.ascontiguousarray(dl_dcenter, dtype=np.float64).reshape(-1, 3)
        dp = None if dl_dpoint is None else np.ascontiguousarray(dl_dpoint, dtype=np.float64).reshape(-1, 3)
        if dc is not None and dc.shape[0] != nc:
            raise ValueError("dl_dcenter must have one row per camera")
        if dp is not None and dp.shape[0] != nt:
            raise ValueError("dl_dpoint must have one row per track")
        dr = None if dl_drotation is None else np.ascontiguousarray(dl_drotation, dtype=np.float64).reshape(-1, 9)
        if dr is not None and dr.shape[0] != nc:
            raise ValueError("dl_drotation must have one 3x3 matrix per camera")
        out = np.zeros((no, 2))
        ctl = np.zeros((max(1, ncp), 3)) if with_control else None
        st = self._lib.pb_hip_backward_pose(self._ctx, _abi.ptr(dc, C.c_double), _abi.ptr(dr, C.c_double),
                                            _abi.ptr(dp, C.c_double), _abi.ptr(out, C.c_double),
                                            _abi.ptr(ctl, C.c_double))
        if st != _abi.PB_OK:
            raise RuntimeError(f"pb_hip_backward_pose: {_abi.STATUS_NAMES.get(st, st)}: {self.last_error()}")
        return (out, ctl[:ncp]) if with_control else out

    def quality(self, qopts) -> dict:
        """Post-solve quality of the last solve on this workspace (pb_hip_quality): the report fields plus
        per-track strict RMS before/after and the filter's validity, in input track order."""
        n = self._num_tracks
        before, after = np.zeros(n), np.zeros(n)
        valid = np.zeros(n, dtype=np.int32)
        rep = _abi.QualityReportC()
        st = self._lib.pb_hip_quality(self._ctx, C.byref(qopts), C.byref(rep), _abi.ptr(before, C.c_double),
                                      _abi.ptr(after, C.c_double), _abi.ptr(valid, C.c_int32))
        if st != _abi.PB_OK:
            raise RuntimeError(f"pb_hip_quality: {_abi.STATUS_NAMES.get(st, st)}: {self.last_error()}")
        out = rep.as_dict()
        out.update(track_rms_before=before, track_rms_after=after, track_valid=valid)
        return out

    def kernel_stats(self, name: str) -> dict:
        launches = C.c_int64()
        ms = C.c_double()
        nbytes = C.c_double()
        st = self._lib.pb_hip_kernel_stats(self._ctx, name.encode(), C.byref(launches), C.byref(ms), C.byref(nbytes))
        if st != _abi.PB_OK:
            raise RuntimeError(self.last_error())
        return {"launches": launches.value, "total_ms": ms.value, "bytes_per_launch": nbytes.value}


class Solver:
    """plabundle::Solver restricted to the MI355X backend."""

    @staticmethod
    def isBackendAvailable(backend: Backend = Backend.Mi355x, deviceIndex: int = 0) -> bool:
        if backend != Backend.Mi355x:
            return False
        try:
            lib = _abi.load_library()
        except _abi.BackendLibraryMissing:
            return False
        return bool(lib.pb_hip_available(int(deviceIndex)))

    def solve(self, problem: Problem, options: SolveOptions | None = None, workspace: SolverWorkspace | None = None,
              precision: Precision | None = None, cancel=None, quality: bool = False) -> Result:
        """With quality=True the post-solve quality runs on the device after the solve (finalizeBundleAdjustResult
        and the image quality gate of src/solver.cpp:123-171): Result.quality holds its report, a solve that keeps
        no track is downgraded to NumericalFailure, and a gate rejection sets qualityGateRejected and makes the
        result unusable (the reference then falls back or publishes nothing)."""
        options = options or SolveOptions()
        result = Result(requestedBackend=options.backend.requested, usedBackend=Backend.Mi355x)
        if options.backend.requested not in (Backend.Mi355x, Backend.Auto):
            result.status = SolveStatus.UnsupportedConfiguration
            result.backendMessage = "this package provides the MI355X backend only"
            return result
        lib = _abi.load_library()  # raises when the HIP library is missing: no silent CPU fallback
        if not lib.pb_hip_available(int(options.backend.plaMatrixDevice)):
            result.status = SolveStatus.BackendUnavailable
            result.backendMessage = f"no gfx950 device at index {options.backend.plaMatrixDevice}"
            return result
        ws = workspace or SolverWorkspace(options.backend.plaMatrixDevice)
        st = ws.ensure_uploaded(problem)
        if st != _abi.PB_OK:
            result.status = SolveStatus.InvalidInput
            result.backendMessage = ws.last_error()
            result.abi_status = st
            return result
        copts = to_c(options, precision)
        rep = _abi.SolveReportC()
        flag = None
        if cancel is not None:
            flag = np.array([1 if cancel else 0], dtype=np.int32)
        callback = None
        user_cb = options.solver.progressCallback
        if user_cb is not None:
            def _progress(summary_ptr, _user):
                sm = summary_ptr.contents
                try:
                    go = user_cb(IterationSummary(sm.current_iteration, sm.max_iterations, sm.average_rms_pixels,
                                                  sm.valid_point_count))
                except Exception:  # a raising callback cancels, like a false return
                    return 0
                return 0 if go is False else 1

            callback = _abi.PROGRESS_FN(_progress)
            lib.pb_hip_set_progress(ws.handle, callback, None)
        try:
            st = lib.pb_hip_solve(ws.handle, C.byref(copts), C.byref(rep), _abi.ptr(flag, C.c_int32))
        finally:
            if callback is not None:
                lib.pb_hip_set_progress(ws.handle, _abi.PROGRESS_FN(), None)
        result.abi_status = st
        result.report = rep.as_dict()
        result.status = SolveStatus(rep.solve_status)
        result.solutionUsable = bool(rep.solution_usable)
        result.backendMessage = result.report["message"]
        result.observationCount = rep.observation_count
        result.usedGpu = st in (_abi.PB_OK,)
        nc, nt = problem.num_cameras, problem.num_tracks
        R = np.zeros((nc, 9))
        Cc = np.zeros((nc, 3))
        I = np.zeros((nc, 13))
        X = np.zeros((nt, 3))
        V = np.zeros(nt, dtype=np.int32)
        lib.pb_hip_download(ws.handle, _abi.ptr(R, C.c_double), _abi.ptr(Cc, C.c_double), _abi.ptr(I, C.c_double),
                            _abi.ptr(X, C.c_double), _abi.ptr(V, C.c_int32))
        result.cam_rotation, result.cam_center, result.cam_intrinsics = R, Cc, I
        result.track_point, result.track_valid = X, V
        if problem.has_rig:
            result.refinedRig = ws.rig_state(problem)
        if quality and st == _abi.PB_OK and result.usable():
            from .options import quality_to_c

            q = ws.quality(quality_to_c(options))
            result.quality = q
            result.status = SolveStatus(q["solve_status"])
            result.solutionUsable = bool(q["solution_usable"]) and not q["quality_gate_rejected"]
            result.qualityGateRejected = bool(q["quality_gate_rejected"])
            result.qualityGateMessage = q["message"]
        return result

    def solve_batch(self, problems: list[Problem], options: SolveOptions | None = None,
                    workspace: SolverWorkspace | None = None, precision: Precision | None = None,
                    cancel=None) -> tuple[list[Result], dict]:
        """Solves independent windows in one call (pb_hip_solve_batch).

        Each window is solved as solve() would solve it alone with the same options, except that every LM step
        solves the window's reduced camera system exactly (dense Cholesky) instead of by PCG. Returns one
        Result per window (report = the window's pb_window_report) and the batch summary (pb_solve_report).
        The workspace then holds the batch; the next solve() uploads its problem again."""
        options = options or SolveOptions()
        lib = _abi.load_library()  # raises when the HIP library is missing: no silent CPU fallback
        results = [Result(requestedBackend=options.backend.requested) for _ in problems]
        if not problems:
            return results, {"solve_status": int(SolveStatus.InvalidInput), "message": "batch has no windows"}
        if not lib.pb_hip_available(int(options.backend.plaMatrixDevice)):
            for r in results:
                r.status = SolveStatus.BackendUnavailable
                r.backendMessage = f"no gfx950 device at index {options.backend.plaMatrixDevice}"
            return results, {"solve_status": int(SolveStatus.BackendUnavailable)}
        ws = workspace or SolverWorkspace(options.backend.plaMatrixDevice)
        ws.clear()
        views = (_abi.ProblemView * len(problems))()
        keep = []
        for i, p in enumerate(problems):
            v, k = _abi.make_view(p)
            views[i] = v
            keep.append(k)
        batch = _abi.BatchView(len(problems), C.cast(views, C.POINTER(_abi.ProblemView)))
        copts = to_c(options, precision)
        reports = (_abi.WindowReportC * len(problems))()
        summary = _abi.SolveReportC()
        flag = None if cancel is None else np.array([1 if cancel else 0], dtype=np.int32)
        st = lib.pb_hip_solve_batch(ws.handle, C.byref(batch), C.byref(copts), reports, C.byref(summary),
                                    _abi.ptr(flag, C.c_int32))
        del keep
        out = summary.as_dict()
        out["abi_status"] = st
        out["factor_tuning"] = ws.batch_tuning()
        for w, (p, r) in enumerate(zip(problems, results)):
            rep = reports[w]
            r.abi_status = st
            r.report = rep.as_dict()
            r.status = SolveStatus(rep.solve_status) if st == _abi.PB_OK else SolveStatus(summary.solve_status)
            r.solutionUsable = st == _abi.PB_OK and bool(rep.solution_usable)
            r.usedGpu = st == _abi.PB_OK
            r.backendMessage = out["message"] if st == _abi.PB_OK else ws.last_error()
            r.observationCount = rep.observation_count
            if st != _abi.PB_OK:
                continue
            nc, nt = p.num_cameras, p.num_tracks
            R, Cc, I = np.zeros((nc, 9)), np.zeros((nc, 3)), np.zeros((nc, 13))
            X, V = np.zeros((nt, 3)), np.zeros(nt, dtype=np.int32)
            lib.pb_hip_batch_download(ws.handle, w, _abi.ptr(R, C.c_double), _abi.ptr(Cc, C.c_double),
                                      _abi.ptr(I, C.c_double), _abi.ptr(X, C.c_double), _abi.ptr(V, C.c_int32))
            r.cam_rotation, r.cam_center, r.cam_intrinsics = R, Cc, I
            r.track_point, r.track_valid = X, V
        return results, out
