"""ctypes wrapper of the FP64 CPU oracle (oracle/pb_oracle.cpp).

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's CPU
baseline leg load this module, and only as the checker or as the timed CPU
baseline. The shipped backend never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from plabundle_amd import _abi

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(ORACLE_DIR, "libpb_oracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
    return ORACLE_LIB


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_LIB):
        build()
    L = C.CDLL(ORACLE_LIB)
    L.oracle_create.argtypes = [C.POINTER(_abi.ProblemView)]
    L.oracle_create.restype = C.c_void_p
    L.oracle_destroy.argtypes = [C.c_void_p]
    L.oracle_set_threads.argtypes = [C.c_void_p, C.c_int]
    L.oracle_max_threads.restype = C.c_int
    L.oracle_solve.argtypes = [C.c_void_p, C.POINTER(_abi.SolveOptionsC), C.POINTER(_abi.SolveReportC), _ip]
    L.oracle_solve.restype = C.c_int
    L.oracle_download.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _ip]
    L.oracle_masks.argtypes = [C.c_void_p, _ip, _ip, _ip, _ip, _ip]
    L.oracle_rig_masks.argtypes = [C.c_void_p, _ip, _ip]
    L.oracle_rig_masks.restype = None
    L.oracle_download_rig.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
    L.oracle_download_rig.restype = None
    L.oracle_prepare.argtypes = [C.c_void_p, C.POINTER(_abi.SolveOptionsC), C.POINTER(_abi.SolveReportC)]
    L.oracle_prepare.restype = C.c_int
    L.oracle_reduced_system.argtypes = [C.c_void_p, C.c_double, _dp, _dp, _dp, _ip]
    L.oracle_reduced_system.restype = C.c_int
    L.oracle_compute_step.argtypes = [C.c_void_p, C.c_double, _dp, _dp, _dp, _dp, _dp]
    L.oracle_compute_step.restype = C.c_int
    L.oracle_cost_after_step.argtypes = [C.c_void_p, _dp, _dp, C.c_double]
    L.oracle_cost_after_step.restype = C.c_double
    L.oracle_decrease.argtypes = [C.c_void_p, _dp, _dp]
    L.oracle_decrease.restype = C.c_double
    L.oracle_default_quality_options.argtypes = [C.POINTER(_abi.QualityOptionsC)]
    L.oracle_default_quality_options.restype = None
    L.oracle_quality.argtypes = [C.c_void_p, C.POINTER(_abi.QualityOptionsC), C.POINTER(_abi.QualityReportC), _dp,
                                 _dp, _ip]
    L.oracle_quality.restype = C.c_int
    L.oracle_initial_cost.argtypes = [C.c_void_p]
    L.oracle_initial_cost.restype = C.c_double
    L.oracle_project.argtypes = [_dp, _dp, _dp, C.c_int, _dp, _dp, _dp, _dp]
    L.oracle_project.restype = C.c_int
    L.oracle_robust_loss.argtypes = [C.c_double, C.c_int, C.c_double, _dp, _dp]
    L.oracle_robust_loss.restype = C.c_int
    L.oracle_point_scale.argtypes = [_dp]
    L.oracle_point_scale.restype = C.c_double
    L.oracle_retract.argtypes = [_dp, _dp, _dp, C.c_int]
    L.oracle_gauge_basis.argtypes = [_dp, _dp, _dp]
    L.oracle_gauge_basis.restype = None
    L.oracle_gauge_step.argtypes = [_dp, _dp, C.c_double, C.c_double, C.c_double, _dp]
    L.oracle_gauge_step.restype = None
    L.oracle_linearize.argtypes = [_dp, _dp, _dp, C.c_int, _dp, C.c_double, C.c_double, C.c_double, C.c_double,
                                   C.c_int, C.c_double, C.c_int, C.c_uint32, _dp, _dp, _dp, _dp, _dp, _dp]
    L.oracle_linearize.restype = C.c_int
    L.oracle_groups.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, C.POINTER(C.c_uint32)]
    L.oracle_groups.restype = C.c_int
    L.oracle_control_whitening.argtypes = [C.c_int, C.c_double, _dp, _dp]
    L.oracle_control_whitening.restype = C.c_int
    _lib = L
    return L


def _p(a, t=C.c_double):
    return _abi.ptr(a, t)


def default_quality_options() -> _abi.QualityOptionsC:
    """Reference quality defaults (options.h:66-68, 168-177) without loading the HIP library."""
    q = _abi.QualityOptionsC()
    lib().oracle_default_quality_options(C.byref(q))
    return q


def default_options() -> _abi.SolveOptionsC:
    """Reference defaults (include/plabundle/options.h:189-297) without loading the HIP library."""
    o = _abi.SolveOptionsC()
    o.max_iterations = 20
    o.robust_loss = 0
    o.robust_loss_scale = 1.0
    o.armijo_coefficient = 1e-3
    o.line_search_steps = 34
    o.use_online_schur = 0
    o.max_initial_track_rms = 100.0
    o.refine_camera_pose = 1
    o.shared_intrinsic_mask = 0
    o.min_focal_scale, o.max_focal_scale = 0.5, 4.0
    o.min_aspect_scale, o.max_aspect_scale = 0.85, 1.18
    o.max_principal_offset_fraction = 0.08
    o.focal_prior_sigma = 0.35
    o.principal_prior_sigma_fraction = 0.04
    o.aspect_prior_sigma = 0.08
    for i, v in enumerate((0.35, 0.35, 0.35, 0.02, 0.02, 0.10, 0.35, 0.35, 0.35)):
        o.max_abs[i] = v
    for i, v in enumerate((0.15, 0.15, 0.15, 0.01, 0.01, 0.02, 0.15, 0.15, 0.15)):
        o.prior_sigma[i] = v
    o.low_order_distortion_scale = 1.0
    o.pcg_model_eta = -1.0
    o.step_rms_tolerance = 0.0
    o.laser_plane_weight = 1.0
    o.laser_huber_delta = 0.2
    o.control_point_weight = 1.0
    o.control_point_huber_delta = 0.2
    o.use_transition_prior = 0
    o.previous_intrinsic_mask = 0
    o.batch_graphs = 1
    o.batch_factor_shape = -1
    o.pcg_graphs = 1
    o.reduced_solver = 0
    o.dense_max_rows = 0
    o.preconditioner_cluster_size = 1
    o.batch_factor_tiles = 1
    return o


class Oracle:
    """One oracle problem instance (mirrors the C ABI's context lifecycle)."""

    def __init__(self, problem, threads: int | None = None):
        self.problem = problem
        view, self._keep = _abi.make_view(problem)
        self._view = view
        self.ctx = lib().oracle_create(C.byref(view))
        lib().oracle_set_threads(self.ctx, threads or lib().oracle_max_threads())

    def close(self):
        if self.ctx:
            lib().oracle_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def prepare(self, options) -> dict:
        rep = _abi.SolveReportC()
        ok = lib().oracle_prepare(self.ctx, C.byref(options), C.byref(rep))
        d = rep.as_dict()
        d["prepared"] = bool(ok)
        return d

    def solve(self, options, cancel=None) -> dict:
        rep = _abi.SolveReportC()
        flag = None
        if cancel is not None:
            flag = np.array([int(cancel)], dtype=np.int32)
        lib().oracle_solve(self.ctx, C.byref(options), C.byref(rep), _p(flag, C.c_int32))
        return rep.as_dict()

    def state(self):
        p = self.problem
        R = np.zeros((p.num_cameras, 9))
        Cc = np.zeros((p.num_cameras, 3))
        I = np.zeros((p.num_cameras, 13))
        X = np.zeros((p.num_tracks, 3))
        V = np.zeros(p.num_tracks, dtype=np.int32)
        lib().oracle_download(self.ctx, _p(R), _p(Cc), _p(I), _p(X), _p(V, C.c_int32))
        return {"cam_rotation": R, "cam_center": Cc, "cam_intrinsics": I, "track_point": X, "track_valid": V}

    def masks(self):
        p = self.problem
        cb = np.zeros(p.num_cameras, dtype=np.int32)
        ib = np.zeros(p.num_cameras, dtype=np.int32)
        gb = np.zeros(p.num_cameras, dtype=np.int32)
        tb = np.zeros(p.num_tracks, dtype=np.int32)
        at = np.zeros(p.num_tracks, dtype=np.int32)
        lib().oracle_masks(self.ctx, *(_p(a, C.c_int32) for a in (cb, ib, gb, tb, at)))
        return {"camera_block": cb, "intrinsic_block_by_camera": ib, "calibration_group_by_camera": gb,
                "track_block": tb, "active_track": at}

    def rig_masks(self):
        n = self.problem.num_cameras
        kb = np.zeros(n, dtype=np.int32)
        sb = np.zeros(n, dtype=np.int32)
        lib().oracle_rig_masks(self.ctx, _p(kb, C.c_int32), _p(sb, C.c_int32))
        return {"capture_block": kb, "sensor_block": sb}

    def rig_state(self):
        p = self.problem
        out = {"capture_rotation": np.zeros((p.num_rig_captures, 9)), "capture_center": np.zeros((p.num_rig_captures, 3)),
               "sensor_rotation": np.zeros((p.num_rig_sensors, 9)), "sensor_center": np.zeros((p.num_rig_sensors, 3))}
        lib().oracle_download_rig(self.ctx, *(_p(out[k]) for k in ("capture_rotation", "capture_center",
                                                                    "sensor_rotation", "sensor_center")))
        return out

    def reduced_system(self, primary_blocks: int, damping: float = 0.0, dense: bool = True):
        n = primary_blocks * 13
        A = np.zeros((n, n)) if dense else None
        b = np.zeros(n)
        cost = C.c_double()
        sing = C.c_int32()
        P = lib().oracle_reduced_system(self.ctx, damping, _p(A), _p(b), C.byref(cost), C.byref(sing))
        if P < 0:
            raise RuntimeError("oracle reduced system failed")
        return A, b, cost.value, bool(sing.value)

    def compute_step(self, primary_blocks: int, track_blocks: int, damping: float = 0.0):
        p = np.zeros(primary_blocks * 13)
        e = np.zeros(track_blocks * 3)
        dd, rg, rp = C.c_double(), C.c_double(), C.c_double()
        rc = lib().oracle_compute_step(self.ctx, damping, _p(p), _p(e), C.byref(dd), C.byref(rg), C.byref(rp))
        if rc != 0:
            raise RuntimeError(f"oracle step failed ({rc})")
        return p, e, dd.value, rg.value, rp.value

    def cost_after_step(self, primary, elim, alpha: float) -> float:
        primary = np.ascontiguousarray(primary, dtype=np.float64)
        elim = np.ascontiguousarray(elim, dtype=np.float64)
        return lib().oracle_cost_after_step(self.ctx, _p(primary), _p(elim), alpha)

    def decrease(self, primary, elim) -> float:
        """Analytic decrease of the objective at the initial state along (primary, elim) (after prepare())."""
        primary = np.ascontiguousarray(primary, dtype=np.float64)
        elim = np.ascontiguousarray(elim, dtype=np.float64)
        return lib().oracle_decrease(self.ctx, _p(primary), _p(elim))

    def quality(self, qopts=None) -> dict:
        """finalizeBundleAdjustResult + the image quality gate on the last solve (per-track arrays included)."""
        if qopts is None:
            qopts = default_quality_options()
        n = self.problem.num_tracks
        before, after = np.zeros(n), np.zeros(n)
        valid = np.zeros(n, dtype=np.int32)
        rep = _abi.QualityReportC()
        lib().oracle_quality(self.ctx, C.byref(qopts), C.byref(rep), _p(before), _p(after), _p(valid, C.c_int32))
        out = rep.as_dict()
        out.update(track_rms_before=before, track_rms_after=after, track_valid=valid)
        return out

    def groups(self) -> dict:
        """Intrinsic groups after prepare(): params, prior, inv_sigma (transition weights when that prior is on),
        lower, upper ([G, 13]) and the enabled slot masks ([G])."""
        n = lib().oracle_groups(self.ctx, None, None, None, None, None, None)
        if n < 0:
            raise RuntimeError("oracle not prepared")
        out = {k: np.zeros((n, 13)) for k in ("params", "prior", "inv_sigma", "lower", "upper")}
        en = np.zeros(n, dtype=np.uint32)
        lib().oracle_groups(self.ctx, *(_p(out[k]) for k in ("params", "prior", "inv_sigma", "lower", "upper")),
                            _p(en, C.c_uint32))
        out["enabled"] = en
        return out

    def initial_cost(self) -> float:
        return lib().oracle_initial_cost(self.ctx)


def project(R, Cc, slots, conv, X):
    R = np.ascontiguousarray(R, dtype=np.float64).reshape(9)
    Cc = np.ascontiguousarray(Cc, dtype=np.float64).reshape(3)
    slots = np.ascontiguousarray(slots, dtype=np.float64).reshape(13)
    X = np.ascontiguousarray(X, dtype=np.float64).reshape(3)
    uv = np.zeros(2)
    Jc = np.zeros(6)
    Js = np.zeros(26)
    ok = lib().oracle_project(_p(R), _p(Cc), _p(slots), int(conv), _p(X), _p(uv), _p(Jc), _p(Js))
    if not ok:
        return None
    return uv, Jc.reshape(2, 3), Js.reshape(2, 13)


def robust_loss(s, loss, scale):
    c, w = C.c_double(), C.c_double()
    ok = lib().oracle_robust_loss(float(s), int(loss), float(scale), C.byref(c), C.byref(w))
    return (c.value, w.value) if ok else None


def point_scale(X):
    X = np.ascontiguousarray(X, dtype=np.float64).reshape(3)
    return lib().oracle_point_scale(_p(X))


def retract(R, Cc, delta, reference_param: bool):
    R = np.ascontiguousarray(R, dtype=np.float64).reshape(9).copy()
    Cc = np.ascontiguousarray(Cc, dtype=np.float64).reshape(3).copy()
    delta = np.ascontiguousarray(delta, dtype=np.float64).reshape(6)
    lib().oracle_retract(_p(R), _p(Cc), _p(delta), 1 if reference_param else 0)
    return R, Cc


def linearize(R, Cc, slots, conv, X, u, v, weight=1.0, scale=1.0, loss=0, loss_scale=1.0,
              reference_param=True, active_mask=0):
    R = np.ascontiguousarray(R, dtype=np.float64).reshape(9)
    Cc = np.ascontiguousarray(Cc, dtype=np.float64).reshape(3)
    slots = np.ascontiguousarray(slots, dtype=np.float64).reshape(13)
    X = np.ascontiguousarray(X, dtype=np.float64).reshape(3)
    r = np.zeros(2)
    Jp = np.zeros(6)
    Jpose = np.zeros(12)
    Jint = np.zeros(26)
    w, c = C.c_double(), C.c_double()
    ok = lib().oracle_linearize(_p(R), _p(Cc), _p(slots), int(conv), _p(X), float(u), float(v), float(weight),
                                float(scale), int(loss), float(loss_scale), 1 if reference_param else 0,
                                int(active_mask), _p(r), _p(Jp), _p(Jpose), _p(Jint), C.byref(w), C.byref(c))
    if not ok:
        return None
    return {"residual": r, "Jp": Jp.reshape(2, 3), "Jpose": Jpose.reshape(2, 6), "Jint": Jint.reshape(2, 13),
            "weight": w.value, "cost": c.value}


def control_whitening(kind, sigma=1.0, matrix=None):
    """Whitening matrix A of one control point (residual sqrt(w) A (X - P)), or None when it is rejected."""
    m = None if matrix is None else np.ascontiguousarray(matrix, dtype=np.float64).reshape(9)
    A = np.zeros(9)
    ok = lib().oracle_control_whitening(int(kind), float(sigma), _p(m), _p(A))
    return A.reshape(3, 3) if ok else None


def gauge_basis(anchor, scale):
    """World tangent basis (2 x 3) of the gauge scale camera at `scale` around the anchor at `anchor`."""
    a = np.ascontiguousarray(anchor, dtype=np.float64)
    c = np.ascontiguousarray(scale, dtype=np.float64)
    out = np.zeros(6)
    lib().oracle_gauge_basis(_p(a), _p(c), _p(out))
    return out.reshape(2, 3)


def gauge_step(anchor, scale, baseline, first, second):
    """World centre of the gauge scale camera after the tangent step (first, second)."""
    a = np.ascontiguousarray(anchor, dtype=np.float64)
    c = np.ascontiguousarray(scale, dtype=np.float64)
    out = np.zeros(3)
    lib().oracle_gauge_step(_p(a), _p(c), float(baseline), float(first), float(second), _p(out))
    return out
