// pb_kernels.h - device data layout of one uploaded problem and the kernel launchers.
//
// Layout (all arrays device-resident, allocated once per problem and reused across solves):
//   units      6-wide primary parameter groups: one per camera pose block (unit id = reference
//              camera block id), then ceil(K/6) per intrinsic group with K active slots
//   entries    per eliminated point, one 6x3 cross block W = sum w J_unit^T Jp per distinct unit
//   pairs      (entry, entry) pairs of one point, sorted by reduced-system block; the Schur
//              complement of a block is the sum over its pair run
//   S          reduced system as block-sparse rows of 6x6 blocks (both triangles stored)
#pragma once

#include "pb_device.h"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace pbk
{
    constexpr int kMaxAlpha = 8;             // step lengths evaluated per trial-cost pass
    constexpr int kGroupUnits = 3;           // 6-wide units of an intrinsic group with all 13 slots refined
    constexpr int kGroupPartialStride = 104;  // >= K(K+1)/2 + K for K <= 13
    // Pairs per Schur work item. Short items keep the pairs that the waves of one XCD hold at once within a few
    // block rows, whose cross blocks then fit that XCD's L2.
    constexpr int kPairChunk = 256;
    // blocks with more work items than this (intrinsic groups: one item per 256 points they see) sum their item
    // partials with the whole wave instead of one lane per element
    constexpr int kFinalizeSerialItems = 16;
    constexpr int kXcds = 8;                 // accelerator dies (XCDs) of an MI355X, each with its own L2
    constexpr int kWaveLanes = 64;           // lanes of a wavefront
    constexpr int kCgWaves = 8;              // waves per workgroup of the fused PCG iteration
    // SpMV steps per wave whose loads are issued before the scalars are known (more steps cost registers and
    // with them the occupancy that lets the whole grid be resident at once)
    template <typename T>
    constexpr int cgPrefetch() { return sizeof(T) == 4 ? 6 : 4; }
    constexpr int kCgAlways = 2;             // steps every wave loads (kCgSegment / 10)
    constexpr int kCgPartLoads = 4;          // unrolled row-partial loads per thread (rows up to 4 * 64 * kCgWaves)
    constexpr int kCgDirect = 60;            // column ids per slot stored by slot (the prefetched steps)
    constexpr int kCgSegment = 20;           // stored blocks per wave before a row gets another wave
    // PCG iterations before the truncated-Newton model test may stop (the first iterations of a badly scaled
    // system barely move the model; a mixed step that still fails its line search is repeated in exact FP64)
    constexpr int kCgModelMinIterations = 3;

    enum Flag
    {
        kFlagInvalid = 0,   // an observation failed to project during linearisation
        kFlagSingular = 1,  // a damped point block failed the 3x3 inverse
        kFlagNotSpd = 2,    // a 6x6 diagonal block of the reduced system is not SPD
        kFlagCount = 4
    };

    enum Iflag
    {
        kIflagDone0 = 0,  // PCG stopped, by launch parity (a launch reads one and writes the other)
        kIflagDone1 = 1,
        kIflagIters = 2,
        kIflagBreakdown = 3,
        kIflagModelStop = 4,  // PCG ended by the quadratic-model test
        kIflagItBase = 5,     // iteration counter of graph-replayed PCG launches (even at every replay)
        kIflagCount = 8
    };

    enum Scalar
    {
        kScalarCost = 0,
        kScalarBnorm,
        kScalarThreshold,
        kScalarRnorm,
        kScalarDd,
        kScalarRmsGlobal,
        kScalarRmsPoint,
        kScalarQ0,  // quadratic model value after `it` steps, double-buffered by iteration parity
        kScalarQ1,
        kScalarDq0,  // model decrease of the last step
        kScalarDq1,
        kScalarInvGamma0,  // 1 / (r, M r) of the last step
        kScalarInvGamma1,
        kScalarInvAlpha0,  // 1 / step length of the last step
        kScalarInvAlpha1,
        kScalarThreshold2,  // squared PCG residual threshold
        kScalarCount = 24
    };

    struct Problem
    {
        int nCam = 0, nGroup = 0, nObs = 0, nActive = 0, nElim = 0, nEntries = 0;
        int nUnits = 0, nBlocks = 0, nNnz = 0, nItems = 0;
        long nPairs = 0;
        int maxGroupK = 0;
        int loss = 0;
        double lossScale = 1.0;
        int gaugeAnchor = -1, gaugeScale = -1;
        double gaugeBaseline = 0.0;
        int globalDim = 0;
        int adjoint = 0;  // 1: linearisations use the adjoint weight root (pb_hip_backward builds)
        // 1: reference parameterisation (online Schur without survey constraints: target-frame Euler rotation,
        // point step scaled by s/atan s); 0: default (world-frame Exp rotation, unscaled point step)
        int refParam = 1;

        // Survey constraints on points (control points, laser planes), grouped by active track: residual
        // r = J (X - target) with 1 or 3 rows, Huber(|r|^2, delta). nCons = 0 when the problem has none.
        int nCons = 0;
        const int32_t* consOffset = nullptr;  // [nActive+1]
        const int32_t* consPoint = nullptr;   // [nCons] active track index
        const int32_t* consRows = nullptr;    // [nCons] 1 or 3
        const double* consJ = nullptr;        // [nCons*9] rows x 3, row-major (sqrt(weight) * whitening / normal)
        const double* consTarget = nullptr;   // [nCons*3]
        const double* consDelta = nullptr;    // [nCons] Huber threshold
        const int32_t* consIndex = nullptr;   // [nCons] caller's control-point index, -1 for a laser plane

        // cameras
        double* camR = nullptr;  // state, camera-to-world [nCam*9]
        double* camC = nullptr;  // state, centres [nCam*3]
        const double* camSlots = nullptr;
        const double* refSlots = nullptr;
        const int32_t* camConv = nullptr;
        const int32_t* camGroup = nullptr;  // group index when intrinsic blocks exist, else -1
        const int32_t* camUnit = nullptr;   // pose unit or -1
        pbdev::Camera* cams = nullptr;      // effective cameras of the current pass
        const int32_t* camObsOffset = nullptr;
        const int32_t* camObs = nullptr;
        const int32_t* camObsPoint = nullptr;  // [nObs] camera-major copies of obsPoint, obsUV, obsW
        const double* camObsUV = nullptr;
        const double* camObsW = nullptr;
        const int32_t* camDiagBlock = nullptr;
        const int32_t* camGroupBlock = nullptr;  // [nCam*kGroupUnits] (pose unit, group unit j) block ids

        // intrinsic groups
        double* groupParams = nullptr;  // [nGroup*13]
        const double* groupPrior = nullptr;
        const double* groupInvSigma = nullptr;
        const double* groupLower = nullptr;
        const double* groupUpper = nullptr;
        const uint32_t* groupEnabled = nullptr;
        const int8_t* groupActive = nullptr;  // [nGroup*12] compact active slot list
        const int32_t* groupK = nullptr;
        const int32_t* groupUnit = nullptr;
        const int32_t* groupCamOffset = nullptr;
        const int32_t* groupCams = nullptr;
        const int32_t* groupPairBlock = nullptr;  // [nGroup*kGroupUnits^2] (unit i, unit j) block ids, i <= j
        const int32_t* slotToCompact = nullptr;   // [nGroup*13]

        // observations (usable observations of active tracks, track order)
        const int32_t* obsCam = nullptr;
        const int32_t* obsPoint = nullptr;  // active track index
        const double* obsUV = nullptr;
        const double* obsW = nullptr;       // weight / scale^2
        const int32_t* obsPoseEntry = nullptr;
        const int32_t* obsGroupEntry = nullptr;

        // active tracks
        double* pointX = nullptr;  // [nActive*3]
        const int32_t* obsOffset = nullptr;  // [nActive+1]
        const int32_t* pointElim = nullptr;  // eliminated index or -1 (fixed track)
        const int32_t* elimTrack = nullptr;  // [nElim] active track index

        // cross-block entries
        const int32_t* entryOffset = nullptr;  // [nElim+1]
        const int32_t* entryUnit = nullptr;
        const int32_t* entryPoint = nullptr;  // eliminated index

        // point linearisation: lanes of the lane-per-observation waves, and the points left to one thread each
        int nLinWaves = 0, nLinFallback = 0;
        const int32_t* linLaneObs = nullptr;    // [nLinWaves*kWaveLanes] observation or -1
        const int32_t* linLanePoint = nullptr;  // [nLinWaves*kWaveLanes] eliminated point or -1
        const int32_t* linFallback = nullptr;   // [nLinFallback] eliminated points

        // reduced system structure
        const int32_t* blockRow = nullptr;  // upper blocks (row <= col)
        const int32_t* blockCol = nullptr;
        const int32_t* blockPosUp = nullptr;  // position of (row, col) in the block-sparse rows
        const int32_t* blockPosLo = nullptr;  // position of (col, row)
        const int32_t* blockItemOffset = nullptr;
        const int32_t* itemBlock = nullptr;
        const int32_t* itemBegin = nullptr;  // [nItems+1] pair ranges
        const int32_t* pairE1 = nullptr;
        const int32_t* pairE2 = nullptr;
        const int32_t* rowPtr = nullptr;
        const int32_t* colIdx = nullptr;
        const int32_t* diagPos = nullptr;
        const int32_t* unitWidth = nullptr;  // active columns of a unit (6, or the group remainder)
        // fused PCG: kCgWaves wave slots per workgroup; a slot walks [begin, end) of one block row, and the
        // first slot of a row (lead = number of slots of the row, all in the same workgroup) owns the row
        int cgBlocks = 0;
        const int4* cgSlot = nullptr;  // [cgBlocks*kCgWaves] {row (-1: empty), begin, end, lead}
        const int32_t* cgColDirect = nullptr;  // [cgBlocks*kCgWaves*kCgDirect] column of block begin+i, -1 past end

        // post-solve quality (pb_hip_quality)
        const pbdev::Camera* inputCams = nullptr;  // [nCam] input cameras with their own intrinsics
        const double* pointX0 = nullptr;           // [nActive*3] input points
        const int8_t* trackStrictBad = nullptr;    // [nActive] 1: a weighted observation has an invalid camera or pixel

        // Batched windows (pb_hip_solve_batch): independent problems solved together as one block-diagonal
        // problem whose cameras, groups, tracks and units are numbered window by window. nWin = 0 for a
        // single solve; then every pointer below is null and the scalar gauge/damping arguments apply.
        int nWin = 0;
        const int32_t* camWin = nullptr;          // [nCam]
        const int32_t* groupWin = nullptr;        // [nGroup]
        const int32_t* elimWin = nullptr;         // [nElim]
        const int32_t* unitWin = nullptr;         // [nUnits]
        const int32_t* camGaugeAnchor = nullptr;  // [nCam] anchor camera of a window's gauge scale camera, else -1
        const double* winBaseline = nullptr;      // [nWin] gauge baseline of the window
        const double* winDamping = nullptr;       // [nWin] LM damping of the current round
        int32_t* winFlags = nullptr;              // [nWin*kFlagCount] build flags per window
        const int32_t* winActive = nullptr;       // [nWin] 1: the window builds and solves this round
        const double* winAlpha = nullptr;         // [nWin*kMaxAlpha] trial step lengths of the current pass
        const int32_t* winNAlpha = nullptr;       // [nWin] trial step lengths of the window in the current pass
        const int32_t* nnzRow = nullptr;          // [nNnz] block row (unit) of each stored block

        // Camera rig (nCap + nSen > 0). Every camera is bound to a capture and a sensor and its pose is their
        // composition. The linearisation, Schur and back-substitution kernels work on camera-level pose units
        // (one per camera whose capture or sensor is refined); the reduced system is solved over the rig units
        // (refined captures, then refined sensors, then the intrinsic-group units): S_rig = T^T S_cam T and
        // rhs_rig = T^T rhs_cam with the chain maps T of each camera (pbdev::rigMaps), and the camera-level step
        // is T x_rig. The reduced-system fields above (nUnits, rowPtr, ...) stay camera-level; the rig system's
        // structure is a second Problem value whose reduced-system fields are the rig's (sysProblem on the host).
        int nCap = 0, nSen = 0;
        double* capR = nullptr;  // state: rig-to-world rotations [nCap*9], rig centres [nCap*3]
        double* capC = nullptr;
        double* senR = nullptr;  // state: camera-to-rig rotations [nSen*9], centres in the rig [nSen*3]
        double* senC = nullptr;
        const int32_t* camCapture = nullptr;  // [nCam]
        const int32_t* camSensor = nullptr;   // [nCam]
        const int32_t* capUnit = nullptr;     // [nCap] rig-level unit of a refined capture, else -1
        const int32_t* senUnit = nullptr;     // [nSen]
        double* rigMap = nullptr;             // [nCam*72] chain maps (capture, sensor) of the current state
        int nSysUnits = 0, nSysBlocks = 0;
        const int32_t* sysBlockRow = nullptr;   // [nSysBlocks] upper blocks of the rig system
        const int32_t* sysBlockCol = nullptr;
        const int32_t* sysPosUp = nullptr;      // stored position of (row, col) / (col, row) in the rig rows
        const int32_t* sysPosLo = nullptr;
        const int32_t* projOffset = nullptr;    // [nSysBlocks+1] contributions of each rig block
        const int32_t* projPos = nullptr;       // camera-level stored block (row a, column b)
        const int32_t* projMapL = nullptr;      // chain map of a (2 camera + 0 capture / 1 sensor), -1 identity
        const int32_t* projMapR = nullptr;      // chain map of b
        const int32_t* sysRhsOffset = nullptr;  // [nSysUnits+1] camera-level units of each rig unit
        const int32_t* sysRhsUnit = nullptr;
        const int32_t* sysRhsMap = nullptr;
        const int32_t* expUnit = nullptr;       // [nUnits*2] rig units of a camera-level unit (-1: none)
        const int32_t* expMap = nullptr;        // [nUnits*2] their chain maps
    };

    // Per-window LM controller state and window-local structure of a batched solve (device arrays).
    enum WinStatus
    {
        kWinRunning = 0,
        kWinConverged = 1,
        kWinNoConvergence = 2,  // iteration budget or line search exhausted
        kWinNumeric = 3         // invalid projection during a build
    };

    enum WinCounter
    {
        kWinIterations = 0,
        kWinAccepted,
        kWinRejected,
        kWinLinearizations,
        kWinEvaluations,
        kWinFp64Iterations,  // iterations built and solved with FP64 assembly
        kWinCounterCount = 8
    };

    constexpr int kMaxDense = 512;     // largest window system (6 * primary units)
    constexpr int kDenseLoads = 16;    // global loads a thread of the dense window gather issues together

    // Launch shapes of the dense window factorisation: workgroup threads and column-panel width. Every shape
    // computes bitwise the same factor and solution (each element is updated by one thread in a fixed order),
    // so the batch driver picks among them by measured time alone. Shape 0 is the fixed default.
    struct FactorShape
    {
        int threads, panel;
    };
    constexpr int kFactorShapes = 6;
    constexpr int kFactorTiles = kFactorShapes;  // BatchLauncher::factor: the MFMA tile factorisation instead
    constexpr FactorShape kFactorShape[kFactorShapes] = {{256, 16}, {128, 16}, {512, 16},
                                                         {256, 8},  {128, 8},  {512, 8}};
    constexpr int kFactorLdsLimit = 160 * 1024;  // LDS of one CU (gfx950)
    constexpr int kTrialChunk = 1024;  // observations per workgroup of a batched trial-cost pass

    struct Batch
    {
        int nWin = 0;
        int nChunks = 0;
        int maxDim = 0;  // largest window system dimension
        // structure
        const int32_t* camOffset = nullptr;        // [nWin+1]
        const int32_t* groupOffset = nullptr;      // [nWin+1]
        const int32_t* elimOffset = nullptr;       // [nWin+1]
        const int32_t* poseUnitOffset = nullptr;   // [nWin+1] pose units [a, b) of window w
        const int32_t* groupUnitOffset = nullptr;  // [nWin+1] intrinsic-group units
        const int32_t* globalDim = nullptr;        // [nWin] active primary parameters (step RMS)
        const int64_t* denseOffset = nullptr;      // [nWin+1] dense system storage (doubles)
        const int32_t* chunkWin = nullptr;         // [nChunks]
        const int32_t* chunkBegin = nullptr;       // [nChunks] observation range [chunkBegin[k], chunkEnd[k])
        const int32_t* chunkEnd = nullptr;         // [nChunks]
        const int32_t* winChunkOffset = nullptr;   // [nWin+1]
        // controls
        int maxIterations = 0, lineSearchSteps = 0;
        double armijo = 0.0, rmsTol = 0.0;
        // Mixed batches start every window with FP32 assembly (phase 0) and move a window to FP64 assembly
        // (phase 1) once an accepted FP32 step lowers its objective by less than switchDecrease (relative), or
        // would meet the step-RMS stop, or an FP32 step fails; the stop is only taken in FP64.
        double switchDecrease = 0.0;
        // state
        int32_t* status = nullptr;   // WinStatus
        int32_t* phase = nullptr;    // 0: FP32 assembly, 1: FP64 assembly
        int32_t* active = nullptr;   // Problem::winActive
        int32_t* reject = nullptr;   // step of this round rejected before the line search
        double* damping = nullptr;   // Problem::winDamping
        double* cost = nullptr;      // objective at the current state
        double* initialCost = nullptr;
        int32_t* counters = nullptr;  // [nWin*kWinCounterCount]
        double* dd = nullptr;
        double* rmsG = nullptr;
        double* rmsP = nullptr;
        int32_t* searching = nullptr;
        int32_t* lsCount = nullptr;
        double* alphaNext = nullptr;
        double* alpha = nullptr;    // Problem::winAlpha
        int32_t* nAlpha = nullptr;  // Problem::winNAlpha
        int32_t* acceptIdx = nullptr;
        double* acceptAlpha = nullptr;
        double* acceptCost = nullptr;
        double* dense = nullptr;
        double* trialPart = nullptr;   // [nChunks*kMaxAlpha]
        int32_t* trialBad = nullptr;
        unsigned long long* doneClock = nullptr;  // [nWin+1]; the last entry is the start of the solve
        int32_t* counts = nullptr;  // [4] running windows in phase 0 / phase 1 after a round, windows searching
    };

    template <typename T>
    struct Work
    {
        T* W = nullptr;       // [nEntries*18] cross blocks W, replaced by U = W L once the point is eliminated
        T* Hinv = nullptr;    // [nElim*9] first 6: lower factor L of Hpp^-1 = L L^T (L00, L10, L11, L20, L21, L22)
        double* bp = nullptr; // [nElim*3]
        double* y = nullptr;  // z = L^T bp
        double* dp = nullptr; // point step
        double* D = nullptr;  // direct part per upper block [nBlocks*36]
        double* directRhs = nullptr;  // [nUnits*6]
        double* rhs = nullptr;        // reduced right-hand side
        double* camCost = nullptr;
        double* camGroupPartial = nullptr;  // [nCam*kGroupPartialStride]
        double* groupPriorCost = nullptr;
        T* itemPartial = nullptr;       // [nItems*36]
        double* itemPartialRhs = nullptr;  // [nItems*6]
        T* S = nullptr;     // [nNnz*36]
        T* Minv = nullptr;  // [nUnits*36]
        T* x = nullptr;  // PCG solution
        T* p = nullptr;  // PCG direction
        T* cgR[2] = {nullptr, nullptr};   // fused PCG vectors by iteration parity: residual r
        T* cgU[2] = {nullptr, nullptr};   // u = M r
        T* cgW[2] = {nullptr, nullptr};   // w = S u
        T* cgMW[2] = {nullptr, nullptr};  // M w
        T* cgS[2] = {nullptr, nullptr};   // s = S p
        T* cgMS[2] = {nullptr, nullptr};  // M s
        double* cgPart[2] = {nullptr, nullptr};  // [nUnits*3] row partials of (r,Mr), (w,Mr), (r,r)
        double* partDd = nullptr;
        double* partSs = nullptr;
        double* scalars = nullptr;
        int32_t* flags = nullptr;
        int32_t* iflags = nullptr;
        double* candGroupParams = nullptr;  // [kMaxAlpha*nGroup*13]
        double* candPriorCost = nullptr;
        double* candR = nullptr;
        double* candC = nullptr;
        pbdev::Camera* candCams = nullptr;  // [kMaxAlpha*nCam]
        double* alphas = nullptr;           // [kMaxAlpha]
        double* trialPart = nullptr;        // [costBlocks*kMaxAlpha]
        int32_t* trialBad = nullptr;
        double* trialCost = nullptr;        // [kMaxAlpha]
        // camera rig: the rig system's solution and the candidate rig poses of a line-search pass
        const T* rigX = nullptr;     // [nSysUnits*6]
        double* candCapR = nullptr;  // [kMaxAlpha*nCap*9]
        double* candCapC = nullptr;
        double* candSenR = nullptr;  // [kMaxAlpha*nSen*9]
        double* candSenC = nullptr;
    };

    template <typename T>
    struct Launcher
    {
        static void prepareCameras(const Problem& P, hipStream_t s);
        static void linearizePoints(const Problem& P, const Work<T>& W, double damping, hipStream_t s);
        static void cameraDirect(const Problem& P, const Work<T>& W, hipStream_t s);
        static void schurBlocks(const Problem& P, const Work<T>& W, double damping, hipStream_t s);
        static void pcgInit(const Problem& P, const Work<T>& W, double relTol, double absTol, hipStream_t s);
        // counted: the absolute iteration is it + the device counter (launches captured into a graph)
        static void pcgIteration(const Problem& P, const Work<T>& W, int it, int maxIt, double relTol, double absTol,
                                 double modelEta, hipStream_t s, bool counted = false);
        static void pcgAdvance(const Problem& P, const Work<T>& W, int launches, hipStream_t s);
        static void backSubstitute(const Problem& P, const Work<T>& W, hipStream_t s);
        static void pointSteps(const Problem& P, const Work<T>& W, hipStream_t s);  // back-substitution only
        static void trialStates(const Problem& P, const Work<T>& W, const double* alphas, int nAlpha, hipStream_t s);
        // dl/dC [nCam*3], dl/dR [nCam*9] (row-major, camera-to-world) and dl/dX (elimination order) into the
        // adjoint right-hand side in step parameters
        static void adjointRhs(const Problem& P, const Work<T>& W, const double* dlCenter, const double* dlRot,
                               const double* dlPoint, hipStream_t s);
        static void adjointObservations(const Problem& P, const Work<T>& W, double* out, hipStream_t s);
        // d loss / d target of every control point (out [control points * 3], zero-filled by the caller)
        static void adjointConstraints(const Problem& P, const Work<T>& W, double* out, hipStream_t s);
        // dense column-major copy of S (A, zero-filled by the caller) and of the reduced right-hand side
        static void denseSystem(const Problem& P, const Work<T>& W, double* A, double* b, hipStream_t s);
        static void trialCosts(const Problem& P, const Work<T>& W, const double* alphas, int nAlpha, int costBlocks,
                               hipStream_t s);
        // the objective part of trialCosts (trial states already formed by trialStates)
        static void trialObjective(const Problem& P, const Work<T>& W, const double* alphas, int nAlpha,
                                   int costBlocks, hipStream_t s);
        static void accept(const Problem& P, const Work<T>& W, int ai, double alpha, hipStream_t s);
        // zeroes the flags, the direct blocks and the direct right-hand side before an assembly (in one launch);
        // with unitAlpha also sets alphas[0] = 1
        static void clearBuild(const Problem& P, const Work<T>& W, bool unitAlpha, hipStream_t s);
        // the host-visible results of a step enqueued with one host wait, written to host memory hostPin
        // (pinned, device-accessible): the scalars at pin[0] and pin[step], the flags at pin[flags] (int32), the
        // dense solver words (int32) and two norms at pin[dense] and pin[dense + 4], trialCost[0] at pin[trial]
        static void publishStep(const Work<T>& W, const int32_t* denseWords, const double* denseNorm, double* hostPin,
                                int flags, int dense, int step, int trial, hipStream_t s);
        // camera rig: chain maps of the current state; the rig system from the camera-level one (damped diagonal,
        // exact zeros -> 1) and its right-hand side; the camera-level step from the rig step; the step RMS over the
        // rig units (scalars[kScalarRmsGlobal], after backSubstitute)
        static void rigMaps(const Problem& P, hipStream_t s);
        static void rigProject(const Problem& P, const Problem& Pr, const Work<T>& W, const Work<T>& Wr,
                               double damping, hipStream_t s);
        static void rigExpand(const Problem& P, const Work<T>& Wr, const Work<T>& W, hipStream_t s);
        static void rigStepRms(const Problem& Pr, const Work<T>& Wr, const Work<T>& W, hipStream_t s);
    };

    // Dense reduced solve (pb_dense.hip): the reduced system gathered into a column-major matrix of order
    // npad = 64 * nt (identity padding), factored L L^T in place in 64 x 64 tiles, and solved in two
    // single-launch triangular sweeps.
    constexpr int kDenseTile = 64;
    constexpr int kDenseBlock = 4;  // tile columns per block of the factorisation (one trailing update per block)
    enum DenseWord
    {
        kDenseEpoch = 0,     // advanced by every factorisation and solve; the progress words carry it
        kDenseTicketLower,   // tile tickets of the two sweeps (dispatch order)
        kDenseTicketUpper,
        kDenseNotSpd,        // a pivot of the factorisation was not positive and finite
        kDenseTimeout,       // a sweep stopped waiting for a tile (bounded spin; not expected to happen)
        kDenseWordCount = 8
    };
    struct DenseState
    {
        int32_t* words = nullptr;               // [kDenseWordCount]
        unsigned long long* progress = nullptr;  // [2] (epoch << 32) | tiles done, of the lower and the upper sweep
    };

    template <typename TA>
    struct DenseLauncher
    {
        // lower tiles of the padded dense matrix (leading dimension 64 * nt) from the block-sparse rows of S
        template <typename TW>
        static void gather(const Problem& P, const TW* S, TA* A, int nt, hipStream_t s);
        // A = L L^T in place (lower triangle); a pivot that is not positive sets words[kDenseNotSpd]
        // with reset false the failure words are kept (a factorisation that continues an earlier one)
        static void factor(TA* A, int nt, const DenseState& st, hipStream_t s, bool reset = true);
        static void begin(const DenseState& st, hipStream_t s);  // new epoch, cleared failure words
        // x <- (L L^T)^-1 x (x has 64 * nt entries, the padded tail zero)
        static void solve(const TA* A, int nt, TA* x, const DenseState& st, hipStream_t s);
        // r = b - S x in FP64 over the block-sparse rows, norm[0] = |r|^2 (fixed-order sum of rowNorm)
        template <typename TW>
        static void residual(const Problem& P, const TW* S, const double* b, const double* x, double* r,
                             double* rowNorm, double* norm, hipStream_t s);
        static void norm2(const double* v, int n, double* out, hipStream_t s);  // out[0] = |v|^2, fixed order
        static void toDense(const double* src, int n, int npad, TA* dst, hipStream_t s);   // padded copy
        static void accumulate(double* x, const TA* c, int n, bool overwrite, hipStream_t s);  // x += c (or = c)
        template <typename TO>
        static void store(const double* src, int n, TO* dst, hipStream_t s);
    };

    // Block-tridiagonal reduced solve (pb_band.hip). The pose units, in a bandwidth-reducing order, are cut into
    // m super-blocks of sp = 64 t rows (t <= kBandMaxTiles) such that S couples only neighbouring super-blocks;
    // the intrinsic-group units form a dense border of nbp rows coupled to every super-block. Odd-even cyclic
    // reduction eliminates every second remaining super-block per level (all of a level in parallel): the
    // eliminated block e is factored, W_e = L_e^-T kept, its couplings X = W_e^T K to its two surviving
    // neighbours and to the border formed, and the survivors, the new coupling between them and the border
    // updated with -X^T X products. The last super-block and the border form a dense system of sp + nbp rows
    // solved by the dense route. Every matrix lives in one arena (offsets in elements).
    constexpr int kBandMaxTiles = 2;
    constexpr int kBandMaxRows = kBandMaxTiles * kDenseTile;
    struct BandElim
    {
        int32_t e, a, b;             // eliminated super-block, its surviving neighbours (b = -1 at the end)
        int32_t slotA, slotB;        // coupling slots (rows e, columns a / b); slotB = -1 when b = -1
        int32_t pendBegin, pendEnd;  // range in the pending list: slots whose X^T y the forward sweep subtracts
        int32_t pad;
    };
    // out = beta * out + (alphaNeg ? -1 : 1) * sum over terms of P^T Q (P, Q column-major with leading dimension
    // sp and sp rows; with kTri P is upper triangular, so row tile i of the result needs rows k < 64 (i + 1))
    struct BandTask
    {
        long long out, p0, q0, p1, q1;
        int32_t ldo, terms, alphaNeg, beta, kTri, pad;
    };
    struct BandTile
    {
        int32_t task, ti, tj, pad;
    };
    constexpr int kBandMaxBorder = 256;  // border rows (nbp) the solve kernels hold in LDS
    // Device-visible description of a band plan (passed to the kernels by value).
    struct BandDev
    {
        int sp = 0, m = 0, nb = 0, nbp = 0, nF = 0, su = 0, nPose = 0, nSlots = 0;
        int n = 0, vN = 0;  // unknowns of S and rows of the band vector (m sp + nbp)
        // arena offsets (elements): D, K, B, C first (cleared before every gather), then X, XB, Pc, F and the
        // solve vectors
        long long offD = 0, offK = 0, offB = 0, offC = 0, offX = 0, offXB = 0, offPc = 0, offF = 0;
        long long offV = 0, offFV = 0, offPV = 0, offPB = 0, size = 0;
        long long offY = 0;  // [m sp] a super-block's y_e between the two launches of a forward level, its t_e in
                             // the backward sweep
        int lastPendBegin = 0, lastPendEnd = 0;  // pending slots of super-block 0 (the last system)
        const int32_t* unitBlk = nullptr;  // [nUnits] super-block of a pose unit, -1 for a border unit
        const int32_t* unitOff = nullptr;  // [nUnits] first row in its super-block or in the border
        const int32_t* k0slot = nullptr;   // [m - 1] slot of the first-level coupling of super-blocks k, k + 1
        const int32_t* rowOf = nullptr;    // [n] band-vector row of unknown j
        const int32_t* srcOf = nullptr;    // [vN] unknown of band-vector row i, or -1
        const BandElim* elims = nullptr;   // every level's eliminations, level after level
        const int32_t* pend = nullptr;     // pending slots, grouped per eliminated block (and super-block 0)
        const BandTask* tasks = nullptr;
        const BandTile* tiles = nullptr;
    };
    struct BandView
    {
        BandDev d;
        int levels = 0;
        std::vector<int> elimStart;   // [levels + 1]
        std::vector<int> tileXStart;  // [levels + 1] tiles forming the couplings X of a level
        std::vector<int> tileUStart;  // [levels + 1] tiles of the updates of a level
        double factorFlops = 0.0;     // floating-point operations of one factorisation (as planned)
        double solveFlops = 0.0;      // of one forward and backward solve
    };

    template <typename TA>
    struct BandLauncher
    {
        template <typename TW>
        static void gather(const Problem& P, const TW* S, const BandView& B, TA* Z, hipStream_t s);
        // cyclic reduction and the dense factorisation of the last system; a pivot that is not positive sets
        // words[kDenseNotSpd]
        static void factor(const BandView& B, TA* Z, const DenseState& st, hipStream_t s);
        // the band vector Z + offV <- (S)^-1 of itself (padding rows zero)
        static void solve(const BandView& B, TA* Z, const DenseState& st, hipStream_t s);
        static void toBand(const BandView& B, const double* src, TA* Z, hipStream_t s);
        static void fromBand(const BandView& B, double* x, const TA* Z, bool overwrite, hipStream_t s);
    };

    // Cluster block-Jacobi PCG (plaMatrixPreconditionerClusterSize, options.h:164; used whenever a preconditioner
    // block spans more than one unit: clusters of several primary blocks, or one intrinsic group of more than six
    // refined slots). A cluster is a run of consecutive primary blocks (a pose unit, or every unit of one group),
    // so its units are a contiguous range; its d x d diagonal block of S (d = 6 units) is gathered, factored and
    // inverted once per build, and the PCG applies the inverse as a dense product. The iterations run as four
    // launches (S p, the step, the preconditioner, the direction), every dot product reduced from per-block
    // partials in a fixed order.
    struct ClusterDev
    {
        int n = 0;                          // clusters
        int maxDim = 0;
        const int32_t* unitBegin = nullptr;  // [n+1] unit range of cluster c
        const int32_t* unitCluster = nullptr;  // [nUnits]
        const long long* matOff = nullptr;   // [n+1] offset (doubles) of the cluster's d x d block
        double* inv = nullptr;               // the inverses (symmetric, row-major)
        double* work = nullptr;              // factorisation scratch, same layout
    };

    template <typename T>
    struct ClusterLauncher
    {
        // gather, Cholesky and inverse of every cluster's diagonal block of S (flags[kFlagNotSpd] on a pivot that
        // is not positive)
        static void factor(const Problem& P, const T* S, const ClusterDev& C, int32_t* flags, hipStream_t s);
        // x = 0, r = rhs, z = M r, p = z, the stop threshold max(relTol |rhs|, absTol)
        static void init(const Problem& P, const Work<T>& W, const ClusterDev& C, double relTol, double absTol,
                         hipStream_t s);
        // iteration it (stops at maxIt, on the residual threshold, or on the truncated-Newton model test when
        // modelEta > 0); launches after the stop return at once
        static void iteration(const Problem& P, const Work<T>& W, const ClusterDev& C, int it, int maxIt,
                              double modelEta, hipStream_t s);
    };

    // Post-solve quality buffers (device): strict track RMS at the input and refined state, sort keys, filter.
    struct Quality
    {
        double* before = nullptr;  // [nActive]
        double* after = nullptr;   // [nActive]
        double* key = nullptr;     // [nActive] refined RMS of filter candidates, inf otherwise
        double* sorted = nullptr;  // [nActive]
        int32_t* valid = nullptr;  // [nActive] kept by the filter
        double* out = nullptr;     // [8] candidates, median, threshold, kept, sum before, sum after
        void* temp = nullptr;      // radix-sort scratch
        size_t tempBytes = 0;
    };

    struct QualityLauncher
    {
        static size_t sortScratchBytes(int n);
        static void run(const Problem& P, const Quality& Q, int filter, double floorPx, double factor, hipStream_t s);
    };

    // Batched windows: one LM iteration of every running window of one precision phase per call sequence
    //   roundBegin, build (shared kernels), afterBuild, factor, backSubstitute (shared), stepStats,
    //   { trialPass } while windows search, finishRound (control, apply), count.
    template <typename T>
    struct BatchLauncher
    {
        // kernel attributes (outside any stream capture); usable[v] = factorisation shape v can be launched
        static void configure(const Batch& B, bool (&usable)[kFactorShapes]);
        static void start(const Batch& B, hipStream_t s);
        static void count(const Batch& B, hipStream_t s);  // running windows per phase -> counts[0], counts[1]
        static void roundBegin(const Problem& P, const Work<T>& W, const Batch& B, int phase, hipStream_t s);
        static void build(const Problem& P, const Work<T>& W, hipStream_t s);
        static void afterBuild(const Problem& P, const Work<T>& W, const Batch& B, hipStream_t s);
        static void factor(const Problem& P, const Work<T>& W, const Batch& B, int shape, hipStream_t s);
        static void backSubstitute(const Problem& P, const Work<T>& W, const Batch& B, hipStream_t s);
        static void trialPass(const Problem& P, const Work<T>& W, const Batch& B, int maxAlpha, hipStream_t s);
        static void finishRound(const Problem& P, const Work<T>& W, const Batch& B, int phase, hipStream_t s);
    };
}  // namespace pbk
