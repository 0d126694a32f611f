"""Solve options of the MI355X backend, mirroring plabundle::SolveOptions.

Field names follow include/plabundle/options.h:58-186 so that code written
against PlaBundle's structured options reads the same here. Only the fields the
frame-BA path consumes are carried across the C ABI (pb_solve_options).
"""
from __future__ import annotations

import enum
from dataclasses import dataclass, field

from . import _abi


class Backend(enum.IntEnum):
    """plabundle::Backend (include/plabundle/backend.h:8-15) plus the MI355X backend."""

    Auto = 0
    PlaMatrixCpu = 1
    PlaMatrixCuda = 2
    PlaMatrixVulkan = 3
    PlaMatrixOpenCl = 4
    Mi355x = 5


class ImageRobustLoss(enum.IntEnum):
    LeastSquares = 0
    Huber = 1
    Cauchy = 2


class IntrinsicParameter(enum.IntEnum):
    """Slot order of the 13-entry parameter block (options.h:24-45)."""

    FocalLength = 0
    FocalAspectRatio = 1
    PrincipalPointX = 2
    PrincipalPointY = 3
    RadialK1 = 4
    RadialK2 = 5
    RadialK3 = 6
    TangentialP1 = 7
    TangentialP2 = 8
    SkewB2 = 9
    RadialK4 = 10
    TangentialP3 = 11
    TangentialP4 = 12


class Precision(enum.IntEnum):
    """pb_precision: Fp64; Mixed (FP32 Jacobians, Schur and reduced solve, FP64 cost and acceptance); Fp32Seed
    (enablePlaMatrixMixedPrecision, options.h:272-273: FP64 system, guarded FP32 seed of the reduced solve)."""

    Fp64 = 0
    Mixed = 1
    Fp32Seed = 2


class ReducedSolver(enum.IntEnum):
    """pb_reduced_solver: Auto (band when the camera graph fits it, else dense up to denseMaxRows rows, else PCG),
    Pcg (block-Jacobi PCG), Dense (Cholesky), Band (block-tridiagonal cyclic reduction)."""

    Auto = 0
    Pcg = 1
    Dense = 2
    Band = 3


@dataclass
class SolverControlOptions:
    maxIterations: int = 20
    imageRobustLoss: ImageRobustLoss = ImageRobustLoss.LeastSquares
    imageRobustLossScalePixels: float = 1.0
    referenceArmijoCoefficient: float = 1.0e-3
    referenceLineSearchSteps: int = 34
    useReferenceOnlineSchur: bool = False
    enablePointFilter: bool = True
    filterMaxReprojError: float = 2.5
    filterSigmaFactor: float = 3.0
    maxInitialTrackRms: float = 100.0
    numThreads: int = 0
    logIterationProgress: bool = True
    # SolverControlOptions::progressCallback (options.h:74): called with an IterationSummary after every trial
    # step; returning False cancels the solve
    progressCallback: object = None


@dataclass
class CalibrationOptions:
    refineCameraPose: bool = True
    refineSharedFocalLength: bool = False
    refineSharedFocalAspectRatio: bool = False
    refineSharedPrincipalPoint: bool = False
    refineSharedModelCoefficients: bool = False
    refineSharedRadialDistortion: bool = False
    refineSharedHighOrderDistortion: bool = True
    refineSharedMetashapeParameters: bool = False
    useSharedIntrinsicParameterMask: bool = False
    sharedIntrinsicParameterMask: list = field(default_factory=lambda: [True] * 13)
    minSharedFocalScale: float = 0.5
    maxSharedFocalScale: float = 4.0
    minSharedFocalAspectScale: float = 0.85
    maxSharedFocalAspectScale: float = 1.18
    maxSharedPrincipalPointOffsetFraction: float = 0.08
    sharedFocalPriorSigma: float = 0.35
    sharedPrincipalPointPriorSigmaFraction: float = 0.04
    sharedFocalAspectPriorSigma: float = 0.08
    maxSharedRadialK1Abs: float = 0.35
    maxSharedRadialK2Abs: float = 0.35
    maxSharedRadialK3Abs: float = 0.35
    maxSharedTangentialP1Abs: float = 0.02
    maxSharedTangentialP2Abs: float = 0.02
    maxSharedSkewFraction: float = 0.10
    maxSharedRadialK4Abs: float = 0.35
    maxSharedTangentialP3Abs: float = 0.35
    maxSharedTangentialP4Abs: float = 0.35
    sharedRadialK1PriorSigma: float = 0.15
    sharedRadialK2PriorSigma: float = 0.15
    sharedRadialK3PriorSigma: float = 0.15
    sharedTangentialP1PriorSigma: float = 0.01
    sharedTangentialP2PriorSigma: float = 0.01
    sharedSkewPriorSigmaFraction: float = 0.02
    sharedRadialK4PriorSigma: float = 0.15
    sharedTangentialP3PriorSigma: float = 0.15
    sharedTangentialP4PriorSigma: float = 0.15
    sharedLowOrderDistortionScale: float = 1.0
    useReferenceCalibrationTransitionPrior: bool = False
    referencePreviousIntrinsicParameterMask: list = field(default_factory=lambda: [False] * 13)


@dataclass
class ConstraintOptions:
    """plabundle::ConstraintOptions (include/plabundle/options.h:122-135), the kinds this backend implements."""

    laserPlaneWeight: float = 1.0
    laserHuberDeltaMeters: float = 0.2
    controlPointWeight: float = 1.0
    controlPointHuberDeltaMeters: float = 0.2


@dataclass
class BackendOptions:
    requested: Backend = Backend.Mi355x
    plaMatrixDevice: int = 0
    enablePlaMatrixMixedPrecision: bool = False
    plaMatrixPreconditionerClusterSize: int = 1
    allowFallback: bool = False
    # MI355X controls (no reference counterpart)
    pcgRelativeTolerance: float = 0.0
    pcgAbsoluteTolerance: float = 0.0
    pcgMaxIterations: int = 0
    pcgCheckInterval: int = 0
    armijoBatch: int = 0
    profileKernels: bool = False
    pcgModelEta: float = -1.0  # < 0: path default (off for FP64, 0.1 for mixed)
    stepRmsTolerance: float = 0.0  # 0: the reference's 1e-6
    batchGraphs: bool = True  # batched windows: replay each round's launches as captured HIP graphs
    batchFactorShape: int = -1  # batched windows: panel factorisation launch shape (-1 = timed on the first round)
    batchFactorTiles: bool = True  # batched windows: MFMA tile factorisation (False: the panel factorisation)
    pcgGraphs: bool = True  # replay PCG iterations from a captured HIP graph
    reducedSolver: ReducedSolver = ReducedSolver.Auto
    denseMaxRows: int = 0  # 0: 16384 rows
    mixedSwitchDecrease: float = 0.0  # Mixed: FP32 -> FP64 assembly below this relative decrease (0: 1e-6, < 0: never)


@dataclass
class QualityGateOptions:
    """plabundle::QualityGateOptions (include/plabundle/options.h:168-177), image checks."""

    enabled: bool = True
    maxAcceptedRmsGrowth: float = 1.25
    minAcceptedValidTrackRatio: float = 0.60


@dataclass
class SolveOptions:
    solver: SolverControlOptions = field(default_factory=SolverControlOptions)
    calibration: CalibrationOptions = field(default_factory=CalibrationOptions)
    backend: BackendOptions = field(default_factory=BackendOptions)
    quality: QualityGateOptions = field(default_factory=QualityGateOptions)
    constraints: ConstraintOptions = field(default_factory=ConstraintOptions)


def shared_intrinsic_parameter_enabled(calibration: CalibrationOptions, parameter: int) -> bool:
    """sharedIntrinsicParameterEnabled (src/backend.cpp:8-47)."""
    c = calibration
    p = IntrinsicParameter(parameter)
    radial = c.refineSharedModelCoefficients or c.refineSharedRadialDistortion
    if p == IntrinsicParameter.FocalLength:
        enabled = c.refineSharedFocalLength
    elif p == IntrinsicParameter.FocalAspectRatio:
        enabled = c.refineSharedFocalAspectRatio
    elif p in (IntrinsicParameter.PrincipalPointX, IntrinsicParameter.PrincipalPointY):
        enabled = c.refineSharedPrincipalPoint
    elif p == IntrinsicParameter.RadialK1:
        enabled = radial
    elif p in (IntrinsicParameter.RadialK2, IntrinsicParameter.RadialK3,
               IntrinsicParameter.TangentialP1, IntrinsicParameter.TangentialP2):
        enabled = radial and c.refineSharedHighOrderDistortion
    else:
        enabled = c.refineSharedMetashapeParameters
    return bool(enabled and (not c.useSharedIntrinsicParameterMask or c.sharedIntrinsicParameterMask[int(p)]))


def shared_intrinsic_mask(calibration: CalibrationOptions) -> int:
    mask = 0
    for i in range(13):
        if shared_intrinsic_parameter_enabled(calibration, i):
            mask |= 1 << i
    return mask


def to_c(options: SolveOptions, precision: Precision | None = None) -> _abi.SolveOptionsC:
    """Flattens SolveOptions into pb_solve_options (the makeCompatibilityOptions step)."""
    s, c, b = options.solver, options.calibration, options.backend
    o = _abi.SolveOptionsC()
    o.max_iterations = int(s.maxIterations)
    o.robust_loss = int(s.imageRobustLoss)
    o.robust_loss_scale = float(s.imageRobustLossScalePixels)
    o.armijo_coefficient = float(s.referenceArmijoCoefficient)
    o.line_search_steps = int(s.referenceLineSearchSteps)
    o.use_online_schur = 1 if s.useReferenceOnlineSchur else 0
    o.max_initial_track_rms = float(s.maxInitialTrackRms)
    o.refine_camera_pose = 1 if c.refineCameraPose else 0
    o.shared_intrinsic_mask = shared_intrinsic_mask(c)
    o.min_focal_scale = c.minSharedFocalScale
    o.max_focal_scale = c.maxSharedFocalScale
    o.min_aspect_scale = c.minSharedFocalAspectScale
    o.max_aspect_scale = c.maxSharedFocalAspectScale
    o.max_principal_offset_fraction = c.maxSharedPrincipalPointOffsetFraction
    o.focal_prior_sigma = c.sharedFocalPriorSigma
    o.principal_prior_sigma_fraction = c.sharedPrincipalPointPriorSigmaFraction
    o.aspect_prior_sigma = c.sharedFocalAspectPriorSigma
    for i, v in enumerate((c.maxSharedRadialK1Abs, c.maxSharedRadialK2Abs, c.maxSharedRadialK3Abs,
                           c.maxSharedTangentialP1Abs, c.maxSharedTangentialP2Abs, c.maxSharedSkewFraction,
                           c.maxSharedRadialK4Abs, c.maxSharedTangentialP3Abs, c.maxSharedTangentialP4Abs)):
        o.max_abs[i] = v
    for i, v in enumerate((c.sharedRadialK1PriorSigma, c.sharedRadialK2PriorSigma, c.sharedRadialK3PriorSigma,
                           c.sharedTangentialP1PriorSigma, c.sharedTangentialP2PriorSigma,
                           c.sharedSkewPriorSigmaFraction, c.sharedRadialK4PriorSigma,
                           c.sharedTangentialP3PriorSigma, c.sharedTangentialP4PriorSigma)):
        o.prior_sigma[i] = v
    o.low_order_distortion_scale = c.sharedLowOrderDistortionScale
    o.use_transition_prior = 1 if c.useReferenceCalibrationTransitionPrior else 0
    o.previous_intrinsic_mask = sum(1 << i for i, v in enumerate(c.referencePreviousIntrinsicParameterMask) if v)
    o.device = int(b.plaMatrixDevice)
    if precision is None:
        precision = Precision.Fp32Seed if b.enablePlaMatrixMixedPrecision else Precision.Fp64
    o.precision = int(precision)
    o.pcg_relative_tolerance = float(b.pcgRelativeTolerance)
    o.pcg_absolute_tolerance = float(b.pcgAbsoluteTolerance)
    o.pcg_max_iterations = int(b.pcgMaxIterations)
    o.pcg_check_interval = int(b.pcgCheckInterval)
    o.armijo_batch = int(b.armijoBatch)
    o.profile_kernels = 1 if b.profileKernels else 0
    o.pcg_model_eta = float(b.pcgModelEta)
    o.step_rms_tolerance = float(b.stepRmsTolerance)
    o.batch_graphs = 1 if b.batchGraphs else 0
    o.batch_factor_shape = int(b.batchFactorShape)
    o.pcg_graphs = 1 if b.pcgGraphs else 0
    o.reduced_solver = int(b.reducedSolver)
    o.dense_max_rows = int(b.denseMaxRows)
    o.preconditioner_cluster_size = int(b.plaMatrixPreconditionerClusterSize)
    o.mixed_switch_decrease = float(b.mixedSwitchDecrease)
    o.batch_factor_tiles = 1 if b.batchFactorTiles else 0
    k = options.constraints
    o.laser_plane_weight = float(k.laserPlaneWeight)
    o.laser_huber_delta = float(k.laserHuberDeltaMeters)
    o.control_point_weight = float(k.controlPointWeight)
    o.control_point_huber_delta = float(k.controlPointHuberDeltaMeters)
    return o


def quality_to_c(options: SolveOptions) -> _abi.QualityOptionsC:
    """The post-solve quality fields of SolveOptions as pb_quality_options."""
    q = _abi.QualityOptionsC()
    q.enable_point_filter = 1 if options.solver.enablePointFilter else 0
    q.filter_max_reproj_error = float(options.solver.filterMaxReprojError)
    q.filter_sigma_factor = float(options.solver.filterSigmaFactor)
    q.enable_quality_gate = 1 if options.quality.enabled else 0
    q.max_accepted_rms_growth = float(options.quality.maxAcceptedRmsGrowth)
    q.min_accepted_valid_track_ratio = float(options.quality.minAcceptedValidTrackRatio)
    return q
