"""Frame bundle-adjustment problem in structure-of-arrays form.

Mirrors plabundle::Problem (include/plabundle/problem.h:37-54) for frame
pinhole cameras: cameras (pose + 13-slot intrinsic block), tracks (initial
point + observations), fixed cameras/tracks, calibration groups and the gauge.
Observations are stored grouped by track (track_offset), which is the layout
the C ABI expects.
"""
from __future__ import annotations

import enum
from dataclasses import dataclass, field

import numpy as np


class GaugePolicy(enum.IntEnum):
    AutoAnchor = 0
    RequireExplicitGauge = 1
    CallerManaged = 2


class BrownConvention(enum.IntEnum):
    OpenCv = 0
    Metashape = 1


class ControlPointUncertainty(enum.IntEnum):
    """plabundle::ControlPointConstraint::Uncertainty (include/plabundle/constraints.h:81-87)."""
    IsotropicSigma = 0
    Covariance = 1
    SqrtInformation = 2


@dataclass
class Problem:
    cam_rotation: np.ndarray            # (Nc, 9) camera-to-world, row-major
    cam_center: np.ndarray              # (Nc, 3)
    cam_intrinsics: np.ndarray          # (Nc, 13) parameter blocks
    track_point: np.ndarray             # (Nt, 3)
    track_offset: np.ndarray            # (Nt + 1,) int64
    obs_camera: np.ndarray              # (No,) int32
    obs_uv: np.ndarray                  # (No, 2)
    obs_weight: np.ndarray | None = None
    obs_scale: np.ndarray | None = None
    cam_convention: np.ndarray | None = None
    cam_image_size: np.ndarray | None = None
    ref_intrinsics: np.ndarray | None = None
    cam_group: np.ndarray | None = None
    fixed_cameras: np.ndarray | None = None
    fixed_tracks: np.ndarray | None = None
    gauge_policy: int = GaugePolicy.AutoAnchor
    gauge_anchor_camera: int = -1
    gauge_scale_camera: int = -1
    gauge_baseline: float = 0.0
    # Survey constraints of plabundle::Track (constraints.h:39-46, 73-88, 136-141) as flat arrays keyed by track.
    cp_track: np.ndarray | None = None        # (Ncp,) int32 track index
    cp_point: np.ndarray | None = None        # (Ncp, 3) surveyed position
    cp_sigma: np.ndarray | None = None        # (Ncp,) sigmaMeters (isotropic)
    cp_weight: np.ndarray | None = None       # (Ncp,)
    cp_uncertainty: np.ndarray | None = None  # (Ncp,) ControlPointUncertainty
    cp_matrix: np.ndarray | None = None       # (Ncp, 9) covariance or square-root information, row-major
    lp_track: np.ndarray | None = None        # (Nlp,) int32 track index
    lp_point: np.ndarray | None = None        # (Nlp, 3) point on the plane
    lp_normal: np.ndarray | None = None       # (Nlp, 3) plane normal
    lp_weight: np.ndarray | None = None       # (Nlp,)
    # Camera rig (Problem::rig, problem.h:52; placamera::RigTopology): capture poses (rig-to-world rotation, rig
    # centre), sensor poses (camera-to-rig rotation, camera centre in the rig) and each camera's binding. A bound
    # camera's pose is R_rig R_sensor, C_rig + R_rig c_sensor; the solve refines captures and sensors.
    rig_capture_rotation: np.ndarray | None = None  # (Nk, 9) row-major
    rig_capture_center: np.ndarray | None = None    # (Nk, 3)
    rig_capture_fixed: np.ndarray | None = None     # (Nk,) int32, RigCapture::fixedPose
    rig_sensor_rotation: np.ndarray | None = None   # (Ns, 9) row-major
    rig_sensor_center: np.ndarray | None = None     # (Ns, 3)
    rig_sensor_fixed: np.ndarray | None = None      # (Ns,) int32, RigSensor::fixedExtrinsic
    cam_rig_capture: np.ndarray | None = None       # (Nc,) int32 capture of each camera's binding
    cam_rig_sensor: np.ndarray | None = None        # (Nc,) int32 sensor of each camera's binding
    meta: dict = field(default_factory=dict)

    @property
    def num_rig_captures(self) -> int:
        return 0 if self.rig_capture_rotation is None else int(np.asarray(self.rig_capture_rotation).shape[0])

    @property
    def num_rig_sensors(self) -> int:
        return 0 if self.rig_sensor_rotation is None else int(np.asarray(self.rig_sensor_rotation).shape[0])

    @property
    def has_rig(self) -> bool:
        return self.num_rig_captures > 0 or self.num_rig_sensors > 0

    @property
    def num_control_points(self) -> int:
        return 0 if self.cp_track is None else int(np.asarray(self.cp_track).size)

    @property
    def num_laser_planes(self) -> int:
        return 0 if self.lp_track is None else int(np.asarray(self.lp_track).size)

    @property
    def num_cameras(self) -> int:
        return int(self.cam_rotation.shape[0])

    @property
    def num_tracks(self) -> int:
        return int(self.track_point.shape[0])

    @property
    def num_observations(self) -> int:
        return int(self.obs_camera.shape[0])

    def copy(self) -> "Problem":
        kwargs = {}
        for name in self.__dataclass_fields__:
            value = getattr(self, name)
            if isinstance(value, np.ndarray):
                value = value.copy()
            elif isinstance(value, dict):
                value = dict(value)
            kwargs[name] = value
        return Problem(**kwargs)

    @staticmethod
    def from_tracks(cameras, tracks, **kwargs) -> "Problem":
        """cameras: list of (R(9), C(3), intrinsics(13)); tracks: list of (X(3), [(cam, u, v, w, ms), ...])."""
        rot = np.array([np.asarray(c[0], dtype=np.float64).reshape(9) for c in cameras])
        cen = np.array([np.asarray(c[1], dtype=np.float64).reshape(3) for c in cameras])
        intr = np.array([np.asarray(c[2], dtype=np.float64).reshape(13) for c in cameras])
        pts = np.array([np.asarray(t[0], dtype=np.float64).reshape(3) for t in tracks]).reshape(-1, 3)
        off = [0]
        cam, uv, w, ms = [], [], [], []
        for t in tracks:
            for o in t[1]:
                cam.append(o[0])
                uv.append((o[1], o[2]))
                w.append(o[3] if len(o) > 3 else 1.0)
                ms.append(o[4] if len(o) > 4 else 1.0)
            off.append(len(cam))
        return Problem(
            cam_rotation=rot,
            cam_center=cen,
            cam_intrinsics=intr,
            track_point=pts,
            track_offset=np.array(off, dtype=np.int64),
            obs_camera=np.array(cam, dtype=np.int32),
            obs_uv=np.array(uv, dtype=np.float64).reshape(-1, 2),
            obs_weight=np.array(w, dtype=np.float64),
            obs_scale=np.array(ms, dtype=np.float64),
            **kwargs,
        )


def intrinsic_block(fx, fy, cx, cy, k1=0.0, k2=0.0, k3=0.0, p1=0.0, p2=0.0, b2=0.0, k4=0.0, p3=0.0, p4=0.0):
    """13-slot parameter block (CameraState.cpp:164-185) from plain intrinsics."""
    return np.array([fx, np.log(fy / fx), cx, cy, k1, k2, k3, p1, p2, b2, k4, p3, p4], dtype=np.float64)
