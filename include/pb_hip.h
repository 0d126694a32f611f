/*
 * pb_hip.h - C ABI of the MI355X (gfx950) bundle-adjustment backend.
 *
 * The backend runs the frame-BA nonlinear solve of PlaBundle's online point
 * Schur path on one GPU: per-observation residual/Jacobian/robust weight,
 * damped 3x3 point-block inversion, Schur reduction into camera-camera
 * blocks, block-Jacobi PCG, point back-substitution and the Armijo
 * acceptance loop, on a device-resident workspace that is reused across
 * iterations and solves.
 *
 * Reference interfaces replaced (paths relative to the PlaBundle tree):
 *   - internal::optimizePointsWithPlaMatrix(cameras, tracks, BAOptions, workspace)
 *       src/internal/BundleAdjustPlaMatrix.h:27-30, src/internal/BundleAdjustPlaMatrix.cpp:162-631
 *     -> pb_hip_upload + pb_hip_solve + pb_hip_download
 *   - plamatrix::internal::SchurComplementSolverWorkspace<double> (the reusable
 *     workspace handed through Solver::solve(problem, options, workspace))
 *       include/plabundle/solver.h:13-32, src/solver.cpp:22-38
 *     -> pb_ctx (created by pb_hip_create, owned by the caller)
 *   - isPlaMatrixBackendAvailable(backend, device)
 *       src/internal/BundleAdjustPlaMatrixRuntime.cpp:58-153
 *     -> pb_hip_available
 *   - validateAndNormalizeBundleAdjustOptions (gauge normalisation) and
 *     prepareActiveProblem (block numbering)
 *       src/internal/BundleAdjustValidation.cpp:371-896,
 *       src/internal/BundleAdjustPlaMatrixProblem.cpp:29-278
 *     -> run inside pb_hip_solve; results exposed by pb_hip_get_masks
 *
 * Conventions
 *   - All functions return pb_status; no exception crosses the ABI.
 *   - Host buffers passed in are borrowed for the duration of the call and
 *     copied; the context owns every device allocation.
 *   - One context per host thread; a context is not re-entrant. Each context
 *     owns one HIP stream on the device it was created for.
 *   - Camera rotations are camera-to-world, row-major 3x3. A world point X
 *     maps to the camera frame as R^T (X - C); the camera looks along +Z.
 *   - Per-camera intrinsics use the 13-slot parameter block of
 *     CameraState::parameterBlock (src/internal/CameraState.cpp:164-185):
 *       [fx, log(fy/fx), cx, cy, k1, k2, k3, p1, p2, b2, k4, p3, p4]
 *     with pixel = (fx*xd + b2*yd + cx, fy*yd + cy).
 */
#ifndef PB_HIP_H
#define PB_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PB_ABI_VERSION 4
#define PB_INTRINSIC_COUNT 13

typedef enum pb_status {
    PB_OK = 0,
    PB_ERR_INVALID = 1,     /* bad arguments or problem data */
    PB_ERR_DEVICE = 2,      /* HIP runtime error or no usable gfx950 device */
    PB_ERR_OOM = 3,         /* device or host allocation failed */
    PB_ERR_NUMERIC = 4,     /* numerical failure (non-finite state, invalid projection in a build) */
    PB_ERR_CANCELLED = 5,   /* cancel flag observed */
    PB_ERR_UNSUPPORTED = 6  /* configuration outside what this backend implements */
} pb_status;

/* Mirrors plabundle::SolveStatus (include/plabundle/backend.h:17-27). */
typedef enum pb_solve_status {
    PB_SOLVE_NOT_RUN = 0,
    PB_SOLVE_SUCCESS = 1,
    PB_SOLVE_NO_CONVERGENCE = 2,
    PB_SOLVE_CANCELLED = 3,
    PB_SOLVE_INVALID_INPUT = 4,
    PB_SOLVE_UNSUPPORTED = 5,
    PB_SOLVE_BACKEND_UNAVAILABLE = 6,
    PB_SOLVE_NUMERICAL_FAILURE = 7
} pb_solve_status;

/* Mirrors plabundle::ImageRobustLoss (include/plabundle/options.h:17-22). */
typedef enum pb_robust_loss {
    PB_LOSS_LEAST_SQUARES = 0,
    PB_LOSS_HUBER = 1,
    PB_LOSS_CAUCHY = 2
} pb_robust_loss;

/* Mirrors plabundle::GaugePolicy (include/plabundle/problem.h:15-20). */
typedef enum pb_gauge_policy {
    PB_GAUGE_AUTO_ANCHOR = 0,
    PB_GAUGE_REQUIRE_EXPLICIT = 1,
    PB_GAUGE_CALLER_MANAGED = 2
} pb_gauge_policy;

/* Brown tangential convention of a frame camera (placamera::BrownTangentialConvention). */
typedef enum pb_brown_convention {
    PB_BROWN_OPENCV = 0,
    PB_BROWN_METASHAPE = 1
} pb_brown_convention;

typedef enum pb_precision {
    PB_PRECISION_FP64 = 0,      /* every stage in FP64 */
    PB_PRECISION_MIXED = 1,     /* FP32 Jacobians/Schur and FP32 reduced solve (dense: FP32 factorisation refined in
                                   FP64 against the FP32 system; PCG: FP32 PCG); FP64 cost, acceptance, point
                                   inverses; a step that fails its line search is repeated in FP64 */
    PB_PRECISION_FP32_SEED = 2  /* enablePlaMatrixMixedPrecision (options.h:272-273): FP64 system, guarded FP32 seed
                                   of the reduced solve (dense route: FP32 factorisation refined in FP64 to the
                                   FP64 tolerance, else the step is solved with the FP64 factorisation) */
} pb_precision;

/* Reduced-system solver of each LM step (no reference option: the reference picks PlaMatrix's dense CPU or
 * PCG GPU route by backend, BundleAdjustPlaMatrixRuntime.cpp:18-56). */
typedef enum pb_reduced_solver {
    PB_REDUCED_AUTO = 0,   /* band when the pattern fits it (at least three super-blocks), else dense when
                              6 * units <= dense_max_rows, else PCG */
    PB_REDUCED_PCG = 1,    /* block-Jacobi preconditioned CG (fused iteration) */
    PB_REDUCED_DENSE = 2,  /* dense Cholesky of the reduced system (64 x 64 tiles, MFMA trailing updates) */
    PB_REDUCED_BAND = 3    /* exact block-tridiagonal solve: pose units in reverse Cuthill-McKee order cut into
                              super-blocks of up to 128 rows, intrinsic-group units as a dense border, odd-even
                              cyclic reduction; PB_ERR_UNSUPPORTED when the camera graph is too wide (bandwidth
                              above 21 pose units) or the border above 256 rows */
} pb_reduced_solver;

/* Uncertainty form of a control point (plabundle::ControlPointConstraint::Uncertainty,
 * include/plabundle/constraints.h:81-87). */
typedef enum pb_control_uncertainty {
    PB_CONTROL_ISOTROPIC_SIGMA = 0,  /* sigma (m) on every axis */
    PB_CONTROL_COVARIANCE = 1,       /* row-major 3x3 covariance (m^2) */
    PB_CONTROL_SQRT_INFORMATION = 2  /* row-major 3x3 square-root information (1/m) */
} pb_control_uncertainty;

/* Problem, structure-of-arrays, observations grouped by track.
 * Mirrors plabundle::Problem (include/plabundle/problem.h:37-54) restricted to
 * frame-pinhole cameras, with the per-track survey constraints of
 * plabundle::Track (control points and LiDAR point-to-plane constraints,
 * include/plabundle/constraints.h:136-141) as flat arrays keyed by track and the
 * camera rig as capture / sensor poses with one binding per camera. */
typedef struct pb_problem_view {
    int32_t num_cameras;
    const double* cam_rotation;     /* [num_cameras*9] camera-to-world, row-major */
    const double* cam_center;       /* [num_cameras*3] world camera centres */
    const double* cam_intrinsics;   /* [num_cameras*13] parameter blocks (slot order above) */
    const int32_t* cam_convention;  /* [num_cameras] pb_brown_convention, or NULL = all OpenCV */
    const int32_t* cam_image_size;  /* [num_cameras*2] width, height in pixels, or NULL */
    const double* ref_intrinsics;   /* [num_cameras*13] shared-intrinsic reference cameras, or NULL = cam_intrinsics */
    const int32_t* cam_group;       /* [num_cameras] calibration group ids (>= 0), or NULL = one group */

    int32_t num_tracks;
    const double* track_point;      /* [num_tracks*3] initial world points */
    const int64_t* track_offset;    /* [num_tracks+1] observation ranges per track */
    const int32_t* obs_camera;      /* [num_obs] camera index */
    const double* obs_uv;           /* [num_obs*2] measured pixel (u, v) */
    const double* obs_weight;       /* [num_obs] or NULL = 1 */
    const double* obs_scale;        /* [num_obs] measurement scale, or NULL = 1 */

    int32_t num_fixed_cameras;
    const int32_t* fixed_cameras;
    int32_t num_fixed_tracks;
    const int32_t* fixed_tracks;

    int32_t gauge_policy;           /* pb_gauge_policy */
    int32_t gauge_anchor_camera;    /* reference gauge anchor (-1 = none) */
    int32_t gauge_scale_camera;     /* reference gauge scale camera (-1 = none) */
    double gauge_baseline;          /* reference gauge baseline (0 = none) */

    /* Control points (ControlPointConstraint, constraints.h:73-88). A track may carry several; within a track
     * they are taken in array order. Any of the optional arrays may be NULL. */
    int32_t num_control_points;
    const int32_t* cp_track;        /* [n] track index */
    const double* cp_point;         /* [n*3] surveyed position (solver frame) */
    const double* cp_sigma;         /* [n] sigmaMeters of an isotropic constraint, or NULL = 1 */
    const double* cp_weight;        /* [n] weight (> 0), or NULL = 1 */
    const int32_t* cp_uncertainty;  /* [n] pb_control_uncertainty, or NULL = all isotropic */
    const double* cp_matrix;        /* [n*9] covariance or square-root information, or NULL when all isotropic */

    /* LiDAR point-to-plane constraints (LaserPlaneConstraint, constraints.h:39-46): the residual is
     * sqrt(w) * normal . (X - point), the normal taken as given (not normalised). */
    int32_t num_laser_planes;
    const int32_t* lp_track;        /* [n] track index */
    const double* lp_point;         /* [n*3] point on the plane */
    const double* lp_normal;        /* [n*3] plane normal (non-zero) */
    const double* lp_weight;        /* [n] weight (> 0), or NULL = 1 */

    /* Camera rig (Problem::rig, problem.h:52; placamera::RigTopology): captures (rig poses), sensors (camera
     * poses in the rig) and one binding (capture, sensor) per bound camera. A bound camera's pose is the
     * composition of its capture and its sensor: camera-to-world R = R_rig R_sensor, centre C = C_rig + R_rig
     * c_sensor (runConcreteBackend composes the rig before the solve, solver.cpp:282); cam_rotation /
     * cam_center of bound cameras are ignored. The solve refines capture and sensor poses instead of camera
     * poses (BundleAdjustPlaMatrixProblem.cpp:157-199, Assembly.cpp:207-284, 1152-1208) on the assembled Schur
     * path. With no captures and no sensors the problem has no rig. */
    int32_t num_rig_captures;
    const double* rig_capture_rotation;  /* [n*9] rig-to-world, row-major (RigCapture::rigToWorldRotation) */
    const double* rig_capture_center;    /* [n*3] rig centre in the world (RigCapture::rigCenterInWorld) */
    const int32_t* rig_capture_fixed;    /* [n] RigCapture::fixedPose, or NULL = none fixed */
    int32_t num_rig_sensors;
    const double* rig_sensor_rotation;   /* [n*9] camera-to-rig, row-major (RigSensor::cameraToRigRotation) */
    const double* rig_sensor_center;     /* [n*3] camera centre in the rig frame (RigSensor::cameraCenterInRig) */
    const int32_t* rig_sensor_fixed;     /* [n] RigSensor::fixedExtrinsic, or NULL = none fixed */
    const int32_t* cam_rig_capture;      /* [num_cameras] capture of the camera's binding, -1 = unbound */
    const int32_t* cam_rig_sensor;       /* [num_cameras] sensor of the camera's binding, -1 = unbound */
} pb_problem_view;

/* Solve options. Mirrors the fields of plabundle::Options
 * (include/plabundle/options.h:189-297) that the frame-BA path reads.
 * pb_default_options fills the reference defaults. */
typedef struct pb_solve_options {
    int32_t max_iterations;              /* maxIterations = 20 */
    int32_t robust_loss;                 /* imageRobustLoss = LeastSquares */
    double robust_loss_scale;            /* imageRobustLossScalePixels = 1 */
    double armijo_coefficient;           /* referenceArmijoCoefficient = 1e-3 */
    int32_t line_search_steps;           /* referenceLineSearchSteps = 34 */
    int32_t use_online_schur;            /* useReferenceOnlineSchur: 1 = online point Schur (path B) when the
                                            problem allows it (no laser planes, canUseReferenceOnlineSchur,
                                            ReferenceSchur.cpp:667-673); 0 = assembled damped Schur (path A) */
    double max_initial_track_rms;        /* maxInitialTrackRms = 100 */
    int32_t refine_camera_pose;          /* refineCameraPose = true */
    uint32_t shared_intrinsic_mask;      /* bit i: slot i refined per group (sharedIntrinsicParameterEnabled) */

    /* CalibrationOptions bounds and priors (options.h:79-128). */
    double min_focal_scale;              /* 0.5 */
    double max_focal_scale;              /* 4.0 */
    double min_aspect_scale;             /* 0.85 */
    double max_aspect_scale;             /* 1.18 */
    double max_principal_offset_fraction;/* 0.08 */
    double focal_prior_sigma;            /* 0.35 (fraction of the focal reference) */
    double principal_prior_sigma_fraction;/* 0.04 */
    double aspect_prior_sigma;           /* 0.08 */
    double max_abs[9];                   /* |k1| |k2| |k3| |p1| |p2| skew-fraction |k4| |p3| |p4| */
    double prior_sigma[9];               /* k1 k2 k3 p1 p2 skew-fraction k4 p3 p4 */
    double low_order_distortion_scale;   /* sharedLowOrderDistortionScale = 1 */

    /* Backend controls. */
    int32_t device;                      /* plaMatrixDevice */
    int32_t precision;                   /* pb_precision */
    double pcg_relative_tolerance;       /* 0 = path default (2e-10 FP64, 1e-3 mixed) */
    double pcg_absolute_tolerance;       /* 0 = 1e-12 */
    int32_t pcg_max_iterations;          /* 0 = max(200, 30 * primary blocks) */
    int32_t pcg_check_interval;          /* host polls PCG convergence every N iterations (0 = 8) */
    int32_t armijo_batch;                /* trial step lengths evaluated per cost pass (0 = 4) */
    int32_t profile_kernels;             /* 1 = time kernels with HIP events (pb_hip_kernel_stats) */
    double pcg_model_eta;                /* inexact-Newton stop: end PCG when i*(Q_i - Q_{i-1})/Q_i < eta for the
                                            quadratic model Q (0 = off; negative = path default: off for FP64,
                                            0.1 for mixed) */
    double step_rms_tolerance;           /* convergence when both step RMS values are at or below it
                                            (0 = 1e-6, BundleAdjustPlaMatrix.cpp:446-466) */

    /* ConstraintOptions (options.h:122-135) of the constraints this backend implements. A constraint kind is
     * enabled when the problem carries at least one (Conversion.cpp:103-106). */
    double laser_plane_weight;           /* laserPlaneWeight = 1 */
    double laser_huber_delta;            /* laserHuberDeltaMeters = 0.2 (scaled by sqrt(weight) per plane) */
    double control_point_weight;         /* controlPointWeight = 1 */
    double control_point_huber_delta;    /* controlPointHuberDeltaMeters = 0.2 (on the whitened residual) */

    /* CalibrationOptions::useReferenceCalibrationTransitionPrior and referencePreviousIntrinsicParameterMask
     * (options.h:118-119; BundleAdjustPlaMatrixModel.cpp:173-225): every refined slot but the aspect replaces its
     * prior sigma by a transition weight, non-zero only when the refined set differs from the previous mask. */
    int32_t use_transition_prior;        /* false */
    uint32_t previous_intrinsic_mask;    /* bit i: slot i was refined in the previous stage */

    /* Backend control of pb_hip_solve_batch: 1 (default) replays each round's fixed launch sequences (build,
     * factorisation, back-substitution, trial passes, controller) as captured HIP graphs, 0 launches them directly.
     * Both give bitwise the same results. */
    int32_t batch_graphs;
    /* Backend control of pb_hip_solve_batch: launch shape of the dense window factorisation (threads of the
     * window's workgroup and column-panel width, pb_hip_batch_factor_shape). -1 (default): every shape that fits
     * the LDS is timed on the first round of a batch and the fastest is kept for the rest of the batch and for
     * later batches of the same class on the context (window count rounded up to a power of two, largest window
     * system rounded up to a multiple of 64 unknowns; pb_hip_batch_tuning reports the times);
     * 0 .. PB_BATCH_FACTOR_SHAPES-1: that shape. Every shape gives bitwise the same results. */
    int32_t batch_factor_shape;
    /* Backend control of pb_hip_solve and pb_hip_backward: 1 (default) replays the PCG iterations from a captured
     * HIP graph of 16 launches (captured once per problem and stop parameters), 0 launches them one by one. Both
     * give bitwise the same results. */
    int32_t pcg_graphs;
    /* pb_reduced_solver (default PB_REDUCED_AUTO). */
    int32_t reduced_solver;
    /* Largest reduced system (rows = 6 * units) solved densely under PB_REDUCED_AUTO (0 = 16384). */
    int32_t dense_max_rows;
    /* plaMatrixPreconditionerClusterSize (options.h:164,285; valid 1..16, src/options.cpp:393-394): consecutive
     * primary units whose diagonal blocks form one block of the PCG preconditioner (1 = block-Jacobi). */
    int32_t preconditioner_cluster_size;
    /* PB_PRECISION_MIXED: an accepted FP32 step that lowers the objective by less than this fraction, or that
     * meets the step-RMS stop, hands the solve to FP64 assembly for the rest (0 = 1e-6, the FP32 rounding floor of
     * the assembled model; negative = FP32 to the end, stopping only when a line search fails). */
    double mixed_switch_decrease;
    /* Batched windows: 1 (default) factors each window's system with the MFMA tile Cholesky (64 x 64 tiles, one
     * fixed launch); 0 uses the left-looking panel factorisation whose launch shape batch_factor_shape fixes or
     * the autotuner times. The two order their sums differently and agree to rounding. */
    int32_t batch_factor_tiles;
} pb_solve_options;

/* Per-solve report. Mirrors plabundle::PlaMatrixDiagnostics (include/plabundle/result.h:91-128). */
typedef struct pb_solve_report {
    int32_t solve_status;                /* pb_solve_status */
    int32_t solution_usable;
    int32_t converged;
    int32_t iterations;
    double initial_cost;
    double final_cost;
    int32_t accepted_steps;
    int32_t rejected_steps;
    int32_t linearizations;
    int32_t objective_evaluations;
    int32_t linear_iterations;
    int32_t rejected_initial_tracks;
    int32_t observation_count;
    int32_t active_track_count;
    int32_t camera_block_count;
    int32_t intrinsic_block_count;
    int32_t track_block_count;
    int32_t primary_block_count;
    int32_t online_schur_used;           /* 1: path B (online point Schur); 0: path A (assembled damped Schur) */
    int32_t mixed_precision_used;
    int32_t gauge_anchor_camera;         /* after normalisation */
    int32_t gauge_scale_camera;
    double gauge_baseline;
    double setup_seconds;                /* host preparation + upload of derived structures */
    double solve_seconds;                /* first device launch of the solve to controller stop */
    double assembly_seconds;
    double linear_solve_seconds;
    double back_substitution_seconds;
    double objective_seconds;
    double linear_tolerance;
    char device_name[64];
    char message[256];
    /* further PlaMatrixDiagnostics fields (result.h:91-128) */
    char linear_solver_name[32];         /* "dense-cholesky-fp64", "dense-cholesky-fp32-refined", "pcg" */
    char preconditioner_name[32];        /* "block-jacobi-6", "cluster-jacobi", "none" */
    int32_t dense_fallbacks;             /* FP32 factorisations rejected and solved again in FP64 */
    int32_t mixed_seed_steps;            /* reduced solves whose FP32 seed was accepted */
    int32_t schur_pattern_builds;        /* reduced-system patterns built for this solve (0 when reused) */
    int32_t schur_pattern_reuses;
    int32_t schur_assembly_on_device;    /* 1 */
    double linear_tolerance_minimum;     /* smallest / largest linear stopping tolerance used */
    double linear_tolerance_maximum;
    double trial_state_seconds;          /* share of objective_seconds spent forming trial states */
    double cholesky_factorization_seconds;
    double triangular_solve_seconds;     /* dense solves incl. refinement */
    double residual_check_seconds;
} pb_solve_report;

/* plabundle::IterationSummary (options.h:50-56), passed to the progress callback after every trial step (the
 * calls of BundleAdjustPlaMatrix.cpp:455-464,544-554). */
typedef struct pb_iteration_summary {
    int32_t current_iteration;
    int32_t max_iterations;
    double average_rms_pixels;           /* sqrt(max(0, cost) / max(1, observation count)) */
    int32_t valid_point_count;           /* active tracks */
} pb_iteration_summary;

/* Returns non-zero to continue, 0 to cancel (SolverControlOptions::progressCallback, options.h:72-74). */
typedef int32_t (*pb_progress_fn)(const pb_iteration_summary* summary, void* user);

typedef struct pb_ctx pb_ctx;

/* Fills *options with the reference defaults (options.h:189-297) and the backend defaults. */
void pb_default_options(pb_solve_options* options);

/* 1 if a gfx950 device with the given index is usable, else 0 (never throws). */
int pb_hip_available(int device);

pb_status pb_hip_create(int device, pb_ctx** ctx);
pb_status pb_hip_destroy(pb_ctx* ctx);

/* Copies the problem into the context (host side). Device structures are built
 * by the next pb_hip_solve; uploading again discards the previous problem. */
pb_status pb_hip_upload(pb_ctx* ctx, const pb_problem_view* problem);

/* Runs the nonlinear solve on the uploaded problem, starting from its initial
 * state. cancel (may be NULL) is polled between trials; when it reads non-zero
 * the solve stops and publishes nothing (state stays at the input). */
pb_status pb_hip_solve(pb_ctx* ctx, const pb_solve_options* options, pb_solve_report* report,
                       const volatile int32_t* cancel);

/* Installs the progress callback of later solves on this context (NULL removes it). It is called on the
 * calling thread of pb_hip_solve after every trial step; returning 0 cancels the solve like the cancel flag. */
pb_status pb_hip_set_progress(pb_ctx* ctx, pb_progress_fn fn, void* user);

/* Refined state of the last solve. Any pointer may be NULL.
 *   cam_rotation [num_cameras*9], cam_center [num_cameras*3],
 *   cam_intrinsics [num_cameras*13] (published per publishIntrinsics),
 *   track_point [num_tracks*3], track_valid [num_tracks]. */
pb_status pb_hip_download(pb_ctx* ctx, double* cam_rotation, double* cam_center, double* cam_intrinsics,
                          double* track_point, int32_t* track_valid);

/* Block numbering of the last solve (bit-exact with prepareActiveProblem).
 * Arrays sized [num_cameras] / [num_tracks]; any pointer may be NULL. */
pb_status pb_hip_get_masks(pb_ctx* ctx, int32_t* camera_block, int32_t* intrinsic_block_by_camera,
                           int32_t* calibration_group_by_camera, int32_t* track_block, int32_t* active_track);

/* Rig block numbering of the last solve: ActiveProblem::rigCaptureBlockByCamera and rigSensorBlockByCamera
 * (BundleAdjustPlaMatrixProblem.h:13-39, numbered by BundleAdjustPlaMatrixProblem.cpp:157-199), arrays sized
 * [num_cameras], -1 where a camera is unbound or its capture / sensor is not refined; any pointer may be NULL.
 * camera_block of pb_hip_get_masks is -1 for every camera of a rig problem, as in the reference. */
pb_status pb_hip_get_rig_masks(pb_ctx* ctx, int32_t* capture_block, int32_t* sensor_block);

/* Refined rig of the last solve (BAResult::refinedRig, BundleAdjustPlaMatrix.cpp:601): capture rotations
 * [num_rig_captures*9] and centres [*3], sensor rotations [num_rig_sensors*9] and centres [*3], in the layout
 * of pb_problem_view. Any pointer may be NULL. Without a solve (or when the solve published nothing) the input
 * rig is returned, as the reference does (:172, :574). */
pb_status pb_hip_download_rig(pb_ctx* ctx, double* capture_rotation, double* capture_center,
                              double* sensor_rotation, double* sensor_center);

/* Kernel timing collected when options.profile_kernels was set.
 * name: kernel family ("linearize_points", "camera_direct", "schur_blocks", "pcg_iteration", "cost",
 * "back_substitution", "dense_factor", "dense_solve"). Returns launch count, total milliseconds and the algorithmic
 * HBM bytes of one launch; for the two dense families a "launch" is one factorisation (gather included) or one
 * pair of triangular sweeps, and bytes_per_launch receives their algorithmic floating-point operations
 * (n^3 / 3 and 2 n^2 for the reduced order n). */
pb_status pb_hip_kernel_stats(pb_ctx* ctx, const char* name, int64_t* launches, double* total_ms,
                              double* bytes_per_launch);

/* Reduced camera system at the uploaded initial state and the given damping
 * (one linearisation, no solve), in the reference's 13-padded block layout:
 * rhs [P*13] and the dense reduced matrix [P*13][P*13] (row-major) after
 * finalizePrimaryDiagonal, plus the objective at that state. Intended for
 * parity checks on small problems (P*13 <= 4096). */
pb_status pb_hip_reduced_system(pb_ctx* ctx, const pb_solve_options* options, double damping,
                                double* matrix, double* rhs, double* cost, int32_t* primary_blocks);

/* Objective (image robust cost + intrinsic priors) at the uploaded initial state. */
pb_status pb_hip_evaluate_cost(pb_ctx* ctx, const pb_solve_options* options, double* cost);

/* Adjoint at the converged state of the last pb_hip_solve on this context (no reference counterpart in
 * PlaBundle; the backward entry of the MI355X backend). For a scalar loss l of the refined state, given
 * dl/dC of the camera centres (dl_dcenter [num_cameras*3] or NULL) and dl/dX of the points
 * (dl_dpoint [num_tracks*3] or NULL), writes dl/d(u, v) of every observation (dl_dobs [num_obs*2], input
 * order; unusable or inactive observations get 0). Implicit-function theorem on the optimality condition:
 * H lambda = dl/dtheta is solved in FP64 through the online point Schur, with H = sum J^T M J and
 * M = d(w_n r)/dr the observation's exact weight matrix (its eigenvalue along r clamped at zero), then
 * dl/dy_i = M_i J_i lambda. Second derivatives of the projection are not included. dl_dcontrol (may be NULL)
 * receives dl/dP of every control point's surveyed position [num_control_points*3] (0 for a control point whose
 * track is not eliminated or whose constraint is skipped); laser planes enter H like control points. */
pb_status pb_hip_backward(pb_ctx* ctx, const double* dl_dcenter, const double* dl_dpoint, double* dl_dobs,
                          double* dl_dcontrol);

/* pb_hip_backward for a loss of the whole refined pose (no reference counterpart, like pb_hip_backward): also
 * takes dl/dR of every camera's refined rotation (dl_drotation [num_cameras*9] or NULL, row-major, for the
 * camera-to-world matrices that pb_hip_download returns). The rotation enters through the step's rotation
 * parameters (dR/dw = -D [e_k]x D R for the reference retraction, [e_k]x R for the default one), so a gradient
 * along a direction R cannot move (outside the tangent space of rotations) contributes nothing. */
pb_status pb_hip_backward_pose(pb_ctx* ctx, const double* dl_dcenter, const double* dl_drotation,
                               const double* dl_dpoint, double* dl_dobs, double* dl_dcontrol);

/* Batched local windows: independent problems solved in one call. Each window is a complete problem (its own
 * camera and track numbering, calibration groups, fixed blocks and gauge) and is solved as pb_hip_upload +
 * pb_hip_solve would solve it alone, with the same options, except that each LM step solves the window's
 * reduced camera system exactly (dense FP64 Cholesky, at most 85 primary units per window) instead of by PCG;
 * the PCG options are not used. No reference counterpart: PlaBundle solves windows one Solver::solve at a time
 * (src/solver.cpp:275-302 per window); this entry is the batched driver of the MI355X backend. */
typedef struct pb_batch_view {
    int32_t num_windows;
    const pb_problem_view* windows;  /* [num_windows] */
} pb_batch_view;

/* Outcome of one window (the per-window fields of pb_solve_report). */
typedef struct pb_window_report {
    int32_t solve_status;          /* pb_solve_status */
    int32_t solution_usable;
    int32_t converged;
    int32_t iterations;
    int32_t accepted_steps;
    int32_t rejected_steps;
    int32_t linearizations;
    int32_t objective_evaluations;
    int32_t fp64_iterations;       /* iterations assembled in FP64 (all of them in an FP64 batch) */
    int32_t observation_count;
    int32_t active_track_count;
    int32_t camera_block_count;
    int32_t primary_block_count;
    double initial_cost;
    double final_cost;
    double device_seconds;         /* device clock from the start of the batch to the window's last step */
} pb_window_report;

/* Solves every window of the batch. The windows share each device launch (one block-diagonal problem) and
 * keep separate LM controllers on the device (damping, line search, acceptance and stopping per window);
 * a window stops when it converges, exhausts its line search or reaches max_iterations, and later launches
 * skip it. PB_PRECISION_MIXED assembles a window's system in FP32 while its accepted steps still lower the
 * objective by at least 1e-2 (relative) and in FP64 afterwards; a window's stop is only taken in FP64. reports [num_windows] receives each window's outcome; summary receives the batch totals
 * (iterations = controller rounds, costs = sums over the windows, solve_seconds = host wall time of the
 * rounds, setup_seconds = host preparation and upload). A window whose input is invalid fails the whole call
 * with PB_ERR_INVALID / PB_ERR_UNSUPPORTED and a message naming the window. cancel is polled between rounds;
 * a cancelled batch publishes nothing. The context then holds the batch: pb_hip_batch_download reads a
 * window, pb_hip_download the concatenated state; pb_hip_solve and pb_hip_backward need a new pb_hip_upload. */
pb_status pb_hip_solve_batch(pb_ctx* ctx, const pb_batch_view* batch, const pb_solve_options* options,
                             pb_window_report* reports, pb_solve_report* summary, const volatile int32_t* cancel);

/* Launch shapes of the dense window factorisation of pb_hip_solve_batch (no reference counterpart). */
#define PB_BATCH_FACTOR_SHAPES 6

/* Threads of the window workgroup and column-panel width of factorisation shape `shape`. */
pb_status pb_hip_batch_factor_shape(int32_t shape, int32_t* threads, int32_t* panel);

/* Factorisation-shape timing of the batches solved on the context, for FP32 (fp64 = 0) or FP64 (fp64 = 1)
 * assembly: *shape receives the timed choice (-1 if no batch was timed), ms [PB_BATCH_FACTOR_SHAPES] (may be
 * NULL) the measured milliseconds of each shape's launch over the timed round (-1 for a shape not timed). */
pb_status pb_hip_batch_tuning(pb_ctx* ctx, int32_t fp64, int32_t* shape, double* ms);

/* Refined state of one window of the last batch, in the window's own numbering (as pb_hip_download). */
pb_status pb_hip_batch_download(pb_ctx* ctx, int32_t window, double* cam_rotation, double* cam_center,
                                double* cam_intrinsics, double* track_point, int32_t* track_valid);

/* Post-solve quality. Mirrors SolverControlOptions::enablePointFilter / filterMaxReprojError /
 * filterSigmaFactor (include/plabundle/options.h:66-68) and the image part of QualityGateOptions
 * (options.h:168-177). pb_default_quality_options fills the reference defaults. */
typedef struct pb_quality_options {
    int32_t enable_point_filter;            /* true */
    double filter_max_reproj_error;         /* 2.5 px: floor of the adaptive threshold */
    double filter_sigma_factor;             /* 3.0: threshold = max(floor, factor * median track RMS) */
    int32_t enable_quality_gate;            /* QualityGateOptions::enabled = true */
    double max_accepted_rms_growth;         /* 1.25 (0 disables only the RMS-growth check) */
    double min_accepted_valid_track_ratio;  /* 0.60 */
} pb_quality_options;

/* Mirrors the image fields of plabundle::QualitySummary (include/plabundle/result.h:50-56) and the outcome of
 * finalizeBundleAdjustResult + resultFailsQualityGate. */
typedef struct pb_quality_report {
    int32_t total_tracks;
    int32_t optimized_tracks;       /* tracks valid after the point filter */
    double valid_track_ratio;       /* optimized_tracks / total_tracks */
    double mean_rms_before;         /* strict reprojection RMS (px) over the optimized tracks, input state */
    double mean_rms_after;          /* the same at the refined state (inf when no track is optimized) */
    double median_rms_after;        /* median strict RMS of the filter candidates */
    double filter_threshold;        /* px (inf when the filter is off) */
    int32_t solve_status;           /* the solve's status after the zero-track downgrade (pb_solve_status) */
    int32_t solution_usable;
    int32_t quality_gate_rejected;  /* 1 when the image quality gate rejects the result */
    char message[128];
} pb_quality_report;

void pb_default_quality_options(pb_quality_options* options);

/* Quality of the last pb_hip_solve, computed on the device: the strict reprojection RMS of every track at the
 * input and at the refined state (a track with a projection behind a camera, an invalid weighted observation or
 * fewer than two distinct cameras has RMS inf), the adaptive point filter, the track statistics and the image
 * quality gate. Replaces finalizeBundleAdjustResult (src/internal/BundleAdjustQuality.cpp:302-502) and the
 * image checks of resultFailsQualityGate (src/solver.cpp:123-171). Per-track outputs (input track order, any
 * pointer may be NULL): track_rms_before / track_rms_after [num_tracks], track_valid [num_tracks] (1 when the
 * track is kept by the filter). */
pb_status pb_hip_quality(pb_ctx* ctx, const pb_quality_options* options, pb_quality_report* report,
                         double* track_rms_before, double* track_rms_after, int32_t* track_valid);

/* Solver-neutral outcome of one solve keyed by caller-owned ids, for legacy/new comparisons. Mirrors
 * plabundle::PhotoAlignmentOutcome (include/plabundle/photo_alignment.h:40-53); the arrays are borrowed. */
typedef struct pb_outcome_view {
    int32_t termination_status;       /* pb_solve_status */
    const char* termination_reason;   /* NULL = "" */
    int32_t requested_backend;        /* plabundle::Backend numbering */
    int32_t used_backend;
    int32_t solution_usable;
    int32_t backend_fallback;
    int32_t quality_gate_rejected;
    double mean_rms_before;
    double mean_rms_after;
    int32_t num_cameras;
    const char* const* camera_ids;    /* [num_cameras] non-empty, unique */
    const double* cam_rotation;       /* [num_cameras*9] camera-to-world, row-major */
    const double* cam_center;         /* [num_cameras*3] */
    const double* cam_intrinsics;     /* [num_cameras*13] */
    const int32_t* cam_image_size;    /* [num_cameras*2] width, height (> 0) */
    int32_t num_tracks;
    const char* const* track_ids;     /* [num_tracks] non-empty, unique */
    const double* track_point;        /* [num_tracks*3] */
    const int32_t* track_valid;       /* [num_tracks] */
    const double* track_rms_before;   /* [num_tracks] */
    const double* track_rms_after;    /* [num_tracks] */
} pb_outcome_view;

/* Mirrors plabundle::PhotoAlignmentComparisonTolerance (photo_alignment.h:61-71). */
typedef struct pb_comparison_tolerance {
    double rms;                        /* 1e-8 px */
    double camera_center_meters;       /* 1e-8 */
    double camera_rotation_degrees;    /* 1e-8 */
    double point_meters;               /* 1e-8 */
    int32_t require_termination_status_match;  /* 1 */
    int32_t require_termination_reason_match;  /* 1 */
    int32_t require_requested_backend_match;   /* 1 */
    int32_t require_used_backend_match;        /* 1 */
} pb_comparison_tolerance;

/* Mirrors plabundle::PhotoAlignmentComparison (photo_alignment.h:86-103); per-id lists are optional outputs of
 * pb_compare_outcomes. */
typedef struct pb_comparison {
    int32_t comparable;
    int32_t equivalent;
    int32_t termination_status_matches;
    int32_t termination_reason_matches;
    int32_t requested_backend_matches;
    int32_t used_backend_matches;
    double rms_before_difference;
    double rms_after_difference;
    double maximum_camera_center_difference_meters;
    double maximum_camera_rotation_difference_degrees;
    double maximum_point_difference_meters;
    int32_t valid_track_mask_mismatches;  /* number of ids whose valid flags differ */
    char message[128];
} pb_comparison;

void pb_default_comparison_tolerance(pb_comparison_tolerance* tolerance);

/* validatePhotoAlignmentOutcome (src/photo_alignment.cpp:193-223): finite mean RMS values, non-empty unique ids,
 * every camera with a valid image size, finite non-zero focal lengths and a finite centre, every track with a
 * finite point and finite RMS values. PB_ERR_INVALID with the reason in error (error_size bytes) otherwise. */
pb_status pb_validate_outcome(const pb_outcome_view* outcome, char* error, int32_t error_size);

/* comparePhotoAlignmentOutcomes (src/photo_alignment.cpp:225-318): matches cameras and tracks by id (the id sets
 * must be equal), the centre distance and rotation angle per camera, the point distance per track valid on both
 * sides, the valid-mask mismatches, and equivalence under the tolerance. Optional per-id outputs in the
 * reference outcome's order: camera_center_difference / camera_rotation_difference_degrees [num_cameras],
 * point_difference [num_tracks] (0 unless valid on both sides), valid_mismatch [num_tracks]. */
pb_status pb_compare_outcomes(const pb_outcome_view* reference, const pb_outcome_view* candidate,
                              const pb_comparison_tolerance* tolerance, pb_comparison* comparison,
                              double* camera_center_difference, double* camera_rotation_difference_degrees,
                              double* point_difference, int32_t* valid_mismatch);

/* Test hook of the direct reduced-system routes (no reference counterpart): the reduced system of the uploaded
 * initial state in FP64 (damping as given) solved once by route (PB_REDUCED_DENSE or PB_REDUCED_BAND; the band
 * route returns PB_ERR_UNSUPPORTED when the pattern does not fit it) into x [6 * units] (unit order). When not
 * NULL: matrix [(6 units)^2] (column-major, unit order) and rhs [6 units] receive the system; info [8] receives
 * the not-positive-definite flag, the relative residual in units of 1e-15 (clamped to 2^31 - 1), for the band
 * route its super-block rows, super-blocks, border rows, levels and bandwidth (pose units), and the number of
 * unknowns; arena (arena_capacity elements) receives the band route's arena after the solve. */
pb_status pb_hip_direct_solve(pb_ctx* ctx, const pb_solve_options* options, double damping, int32_t route,
                              double* x, double* matrix, double* rhs, int32_t* info, double* arena,
                              int64_t arena_capacity);

/* Test and measurement hook of the dense reduced solver (no reference counterpart): factors the symmetric
 * positive-definite n x n matrix A (column-major; the lower triangle is read) on the context's device in FP64
 * (fp32 = 0) or FP32 (fp32 = 1), the way a reduced system is factored per LM step, and solves A x = b (x in FP64;
 * no refinement). info (may be NULL) receives 0, or 1 when a pivot was not positive. seconds (may be NULL)
 * receives the device time of the factorisation [0] and of the two triangular sweeps [1]. */
pb_status pb_hip_dense_solve(pb_ctx* ctx, int32_t n, int32_t fp32, const double* A, const double* b, double* x,
                             int32_t* info, double* seconds);

/* Largest system and scratch size of pb_hip_dense_solve_noalloc. */
#define PB_NOALLOC_MAX_ROWS 1024
#define PB_NOALLOC_SCRATCH_BYTES (8u * 1024u * 1024u + 16384u)

/* No counterpart in the reference: for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). The FP64 factor-and-solve of pb_hip_dense_solve on caller
 * device memory: A (device, n x n column-major with leading dimension n; the lower triangle is read, A is not
 * modified) and b (device, n doubles, overwritten by x = A^-1 b), on the context's stream, with every
 * intermediate in the library's static per-device scratch array of PB_NOALLOC_SCRATCH_BYTES (held until the
 * process ends). n above PB_NOALLOC_MAX_ROWS is refused with PB_ERR_INVALID. Allocates nothing; returns once the
 * stream has finished. info (host, may be NULL) receives 0, or 1 when a pivot was not positive. Calls that overlap
 * on one GPU share its scratch array and race. */
pb_status pb_hip_dense_solve_noalloc(pb_ctx* ctx, int32_t n, const double* A_device, double* b_device,
                                     int32_t* info);

/* Last error message of the context (never NULL). */
const char* pb_hip_last_error(const pb_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* PB_HIP_H */
