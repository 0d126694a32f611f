// pb_kernels.hip - HIP kernels of the MI355X frame-BA backend (gfx950, wave64).
//
// Stage map (reference function -> kernel):
//   effective cameras (Assembly.cpp:442-475, Projection.cpp:26-63) ........ k_prepare_cameras
//   linearizeTrack + eliminatePoint point side (ReferenceSchur.cpp:195-490)  k_linearize_points
//   accumulateObservationPrimaryTerms (ReferenceSchur.cpp:140-193) ......... k_camera_direct, k_group_reduce
//   Schur camera-camera blocks + merge + finalize (ReferenceSchur.cpp:437-468, 866-930)
//                                                                            k_schur_partial, k_schur_finalize
//   block-Jacobi PCG (the PlaMatrix GPU seam, BundleAdjustPlaMatrix.cpp:311-369)
//                                                                            k_block_jacobi, k_pcg_*
//   recoverReferencePointSteps (ReferenceSchur.cpp:982-1082) ............... k_back_substitute, k_step_stats
//   applyStep + evaluateObjective (Assembly.cpp:1010-1247) ................. k_trial_*, k_accept_*
//
// Every reduction runs in a fixed order (wave butterflies, per-block partials summed in index
// order), so results are bitwise reproducible run to run. No float atomics are used.
#include "pb_device.h"
#include "pb_kernels.h"
#include "pb_common.h"

#include <hip/hip_runtime.h>

namespace pbk
{
    using pbdev::Camera;
    using pbdev::kMaxActive;
    using pbdev::Linearization;

    // Workgroups are dispatched round-robin over the XCDs (workgroup b runs on XCD b % kXcds). For a grid whose
    // size is a multiple of kXcds this gives XCD x the contiguous range [x * n / kXcds, (x + 1) * n / kXcds) of
    // work indices, so neighbouring work items, which share data, share an L2.
    __device__ __forceinline__ int xcdLocal(int b, int n)
    {
        return (b % kXcds) * (n / kXcds) + b / kXcds;
    }

    // ---------------------------------------------------------------------------------------------
    // Effective camera: pose + the shared-group overlay of the intrinsic slots.
    __device__ void makeCamera(const Problem& P, int c, const double* R, const double* Cc, const double* gparams,
                               const double* anchorC, const double* scaleC, Camera& out)
    {
        const double* s = P.camSlots + c * 13;
        const double* ref = P.refSlots + c * 13;
        const int g = P.camGroup[c];
        const uint32_t en = g >= 0 ? P.groupEnabled[g] : 0u;
        const double* gp = g >= 0 ? gparams + g * 13 : nullptr;
        auto on = [&](int i) { return (en >> i) & 1u; };
        for (int i = 0; i < 9; ++i) out.R[i] = R[i];
        for (int i = 0; i < 3; ++i) out.C[i] = Cc[i];
        const double fx = on(0) ? gp[0] : s[0];
        const double aspect = on(1) ? exp(gp[1]) : exp(s[1]);
        out.fx = fx;
        out.fy = fx * aspect;
        out.cx = on(2) ? ref[2] + gp[2] : s[2];
        out.cy = on(3) ? ref[3] + gp[3] : s[3];
        out.k1 = on(4) ? gp[4] : s[4];
        out.k2 = on(5) ? gp[5] : s[5];
        out.k3 = on(6) ? gp[6] : s[6];
        out.p1 = on(7) ? gp[7] : s[7];
        out.p2 = on(8) ? gp[8] : s[8];
        out.b2 = on(9) ? gp[9] : s[9];
        out.k4 = on(10) ? gp[10] : s[10];
        const bool meta = P.camConv[c] == 1;
        out.p3 = meta ? (on(11) ? gp[11] : s[11]) : 0.0;
        out.p4 = meta ? (on(12) ? gp[12] : s[12]) : 0.0;
        out.metashape = meta ? 1 : 0;
        out.tangent = 0;
        for (int i = 0; i < 6; ++i) out.basis[i] = 0.0;
        if (anchorC)
        {
            // reference gauge tangent basis (Assembly.cpp:69-99) in target coordinates D = diag(1,-1,-1)
            const double o[3] = {anchorC[0], -anchorC[1], -anchorC[2]};
            const double q[3] = {scaleC[0], -scaleC[1], -scaleC[2]};
            double un[3] = {q[0] - o[0], q[1] - o[1], q[2] - o[2]};
            const double len = sqrt(un[0] * un[0] + un[1] * un[1] + un[2] * un[2]);
            const double il = len >= 1e-20 ? 1.0 / len : 0.0;
            for (double& v : un) v *= il;
            int ax = fabs(un[0]) > fabs(un[1]) ? 1 : 0;
            if (fabs(un[ax]) > fabs(un[2])) ax = 2;
            const double a[3] = {ax == 0 ? 1.0 : 0.0, ax == 1 ? 1.0 : 0.0, ax == 2 ? 1.0 : 0.0};
            double sec[3] = {a[1] * un[2] - a[2] * un[1], a[2] * un[0] - a[0] * un[2], a[0] * un[1] - a[1] * un[0]};
            const double sl = sqrt(sec[0] * sec[0] + sec[1] * sec[1] + sec[2] * sec[2]);
            const double isl = sl >= 1e-20 ? 1.0 / sl : 0.0;
            for (double& v : sec) v *= isl;
            const double fst[3] = {un[1] * sec[2] - un[2] * sec[1], un[2] * sec[0] - un[0] * sec[2],
                                   un[0] * sec[1] - un[1] * sec[0]};
            out.basis[0] = fst[0];
            out.basis[1] = -fst[1];
            out.basis[2] = -fst[2];
            out.basis[3] = sec[0];
            out.basis[4] = -sec[1];
            out.basis[5] = -sec[2];
            out.tangent = 1;
        }
    }

    // Anchor camera of the gauge when c is the gauge scale camera (of its window in a batch), else -1.
    __device__ __forceinline__ int gaugeAnchorOf(const Problem& P, int c)
    {
        if (P.camGaugeAnchor) return P.camGaugeAnchor[c];
        return (P.gaugeScale >= 0 && c == P.gaugeScale) ? P.gaugeAnchor : -1;
    }

    __device__ __forceinline__ double gaugeBaselineOf(const Problem& P, int c)
    {
        return P.nWin ? P.winBaseline[P.camWin[c]] : P.gaugeBaseline;
    }

    // LM damping and build-flag words of the window that owns item i (the scalar / global ones when unbatched)
    __device__ __forceinline__ double dampingOf(const Problem& P, const int32_t* winOf, int i, double damping)
    {
        return P.nWin ? P.winDamping[winOf[i]] : damping;
    }

    template <typename T>
    __device__ __forceinline__ int32_t* flagsOf(const Problem& P, const Work<T>& Wk, const int32_t* winOf, int i)
    {
        return P.nWin ? P.winFlags + kFlagCount * winOf[i] : Wk.flags;
    }

    // true when item i belongs to a window that does not build this round
    __device__ __forceinline__ bool idleWindow(const Problem& P, const int32_t* winOf, int i)
    {
        return P.nWin && !P.winActive[winOf[i]];
    }

    __global__ void k_prepare_cameras(Problem P)
    {
        const int c = blockIdx.x * blockDim.x + threadIdx.x;
        if (c >= P.nCam) return;
        const int anchor = gaugeAnchorOf(P, c);
        const double* anchorC = anchor >= 0 ? P.camC + 3 * anchor : nullptr;
        const double* scaleC = anchor >= 0 ? P.camC + 3 * c : nullptr;
        makeCamera(P, c, P.camR + 9 * c, P.camC + 3 * c, P.groupParams, anchorC, scaleC, P.cams[c]);
    }

    // ---------------------------------------------------------------------------------------------
    // Point side of the linearisation: Hpp, bp, cross blocks W (6x3 per entry), damped inverse.

    // The end of one point's linearisation, given the sums H (3x3) and bp over its observations: the point-only
    // constraints, the damped inverse, its factor Lf (L00, L10, L11, L20, L21, L22 of Hpp^-1 = L L^T), and the
    // stored Hinv (the factor), bp and y = L^T bp.
    template <typename T>
    __device__ __forceinline__ void finishPoint(const Problem& P, const Work<T>& Wk, int b, int t, double X0, double X1,
                                                double X2, double damping, int32_t* flags, double H[9], double bp[3],
                                                double Lf[6])
    {
        // point-only survey constraints (accumulatePointOnlyConstraint, ReferenceSchur.cpp:241-259): w J^T J and
        // -w J^T r; the adjoint build uses the root Q of the exact weight matrix instead (J <- Q J, weight 1)
        if (P.nCons)
            for (int k = P.consOffset[t]; k < P.consOffset[t + 1]; ++k)
            {
                double r[3], cost, w;
                if (!constraintAt(P, k, X0, X1, X2, r, cost, w)) continue;
                const int rows = P.consRows[k];
                double J[9];
#pragma unroll
                for (int i = 0; i < 9; ++i) J[i] = P.consJ[9 * k + i];
                if (P.adjoint)
                {
                    double q[2];
                    const double ss = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
                    pbdev::constraintWeightRoot(ss, w, P.consDelta[k], q);
                    double QJ[9];
#pragma unroll
                    for (int i = 0; i < 3; ++i)
#pragma unroll
                        for (int j = 0; j < 3; ++j)
                        {
                            double v = q[0] * J[3 * i + j];
#pragma unroll
                            for (int m = 0; m < 3; ++m) v += q[1] * r[i] * r[m] * J[3 * m + j];
                            QJ[3 * i + j] = i < rows ? v : 0.0;
                        }
#pragma unroll
                    for (int i = 0; i < 9; ++i) J[i] = QJ[i];
                    w = 1.0;
                }
#pragma unroll
                for (int i = 0; i < 3; ++i)
                {
                    if (i >= rows) break;
#pragma unroll
                    for (int c = 0; c < 3; ++c)
                    {
                        bp[c] -= w * J[3 * i + c] * r[i];
#pragma unroll
                        for (int j = 0; j < 3; ++j) H[c * 3 + j] += w * J[3 * i + c] * J[3 * i + j];
                    }
                }
            }
        double Hd[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) Hd[i] = H[i];
        Hd[0] *= 1.0 + damping;
        Hd[4] *= 1.0 + damping;
        Hd[8] *= 1.0 + damping;
        double inv[9];
        bool invertible = pbdev::inverse3(Hd, inv);
        if (!invertible && P.adjoint)
        {
            // the adjoint's weight matrices are rank-deficient for observations on the clamped branch of the
            // loss (a two-view outlier track can leave its point block singular): a ridge of 1e-9 of the largest
            // diagonal entry holds the directions the data leave free
            const double ridge = 1e-9 * fmax(fabs(Hd[0]), fmax(fabs(Hd[4]), fabs(Hd[8])));
            Hd[0] += ridge;
            Hd[4] += ridge;
            Hd[8] += ridge;
            invertible = ridge > 0.0 && pbdev::inverse3(Hd, inv, 1e-30);
        }
        if (!invertible)
        {
            setFlag(flags, kFlagSingular);
#pragma unroll
            for (int i = 0; i < 9; ++i) inv[i] = 0.0;
        }
        // Factor the inverse, Hpp^-1 = L L^T (L lower), and store U = W L in place of each cross block W: then
        // W Hpp^-1 W'^T = U U'^T, W Hpp^-1 bp = U z with z = L^T bp, and Hpp^-1 (bp - W^T dc) = L (z - U^T dc).
        {
            const double l00 = sqrt(inv[0]);
            const double l10 = inv[3] / l00, l20 = inv[6] / l00;
            const double l11 = sqrt(inv[4] - l10 * l10);
            const double l21 = (inv[7] - l20 * l10) / l11;
            const double l22 = sqrt(inv[8] - l20 * l20 - l21 * l21);
            Lf[0] = l00;
            Lf[1] = l10;
            Lf[2] = l11;
            Lf[3] = l20;
            Lf[4] = l21;
            Lf[5] = l22;
            bool ok = true;
#pragma unroll
            for (int i = 0; i < 6; ++i) ok = ok && pbdev::finite(Lf[i]);
            if (!ok)
            {
                setFlag(flags, kFlagSingular);
#pragma unroll
                for (int i = 0; i < 6; ++i) Lf[i] = 0.0;
            }
        }
#pragma unroll
        for (int i = 0; i < 6; ++i) Wk.Hinv[9 * b + i] = T(Lf[i]);
        Wk.bp[3 * b] = bp[0];
        Wk.bp[3 * b + 1] = bp[1];
        Wk.bp[3 * b + 2] = bp[2];
        Wk.y[3 * b] = Lf[0] * bp[0] + Lf[1] * bp[1] + Lf[3] * bp[2];
        Wk.y[3 * b + 1] = Lf[2] * bp[1] + Lf[4] * bp[2];
        Wk.y[3 * b + 2] = Lf[5] * bp[2];
    }

    // Cross block U = W L of one entry row: w0 w1 w2 the row of W.
    __device__ __forceinline__ void crossRow(double w0, double w1, double w2, const double Lf[6], double u[3])
    {
        u[0] = w0 * Lf[0] + w1 * Lf[1] + w2 * Lf[3];
        u[1] = w1 * Lf[2] + w2 * Lf[4];
        u[2] = w2 * Lf[5];
    }

    // One thread per point: the points the lane-per-observation kernel does not take (no observation, more than
    // a wave's lanes of observations, two observations in one camera).
    template <typename T, int KM>
    __global__ void __launch_bounds__(128) k_linearize_points(Problem P, Work<T> Wk, double damping)
    {
        const int idx = blockIdx.x * blockDim.x + threadIdx.x;
        if (idx >= P.nLinFallback) return;
        const int b = P.linFallback[idx];
        if (idleWindow(P, P.elimWin, b)) return;
        damping = dampingOf(P, P.elimWin, b, damping);
        int32_t* flags = flagsOf(P, Wk, P.elimWin, b);
        const int t = P.elimTrack[b];
        const double X0 = P.pointX[3 * t], X1 = P.pointX[3 * t + 1], X2 = P.pointX[3 * t + 2];
        const int e0 = P.entryOffset[b], e1 = P.entryOffset[b + 1];
        for (int e = e0; e < e1; ++e)
            for (int i = 0; i < 18; ++i) Wk.W[e * 18 + i] = T(0);
        double H[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        double bp[3] = {0, 0, 0};
        Linearization L;
        for (int o = P.obsOffset[t]; o < P.obsOffset[t + 1]; ++o)
        {
            const int c = P.obsCam[o];
            const int g = P.camGroup[c];
            const int K = g >= 0 ? P.groupK[g] : 0;
            const int8_t* act = g >= 0 ? P.groupActive + g * kMaxActive : nullptr;
            if (!pbdev::linearize<KM>(P.cams[c], X0, X1, X2, P.obsUV[2 * o], P.obsUV[2 * o + 1], P.obsW[o], P.loss,
                                      P.lossScale, K, act, L, P.refParam != 0))
            {
                setFlag(flags, kFlagInvalid);
                continue;
            }
            if (P.adjoint)
            {
                double q[3];
                pbdev::adjointWeightRoot(L, P.obsW[o], P.loss, P.lossScale, q);
                pbdev::applyWeightRoot<KM>(q, L);
            }
            const double w = L.w;
#pragma unroll
            for (int i = 0; i < 3; ++i)
            {
                bp[i] -= w * (L.Jp[i] * L.r[0] + L.Jp[3 + i] * L.r[1]);
#pragma unroll
                for (int j = 0; j < 3; ++j) H[i * 3 + j] += w * (L.Jp[i] * L.Jp[j] + L.Jp[3 + i] * L.Jp[3 + j]);
            }
            const int ep = P.obsPoseEntry[o];
            if (ep >= 0)
            {
                T* Wp = Wk.W + ep * 18;
#pragma unroll
                for (int a = 0; a < 6; ++a)
#pragma unroll
                    for (int i = 0; i < 3; ++i)
                        Wp[a * 3 + i] += T(w * (L.Jc[a] * L.Jp[i] + L.Jc[6 + a] * L.Jp[3 + i]));
            }
            const int eg = P.obsGroupEntry[o];
            if (eg >= 0)
            {
                // predicated rather than cut short, so the loop unrolls and Ji stays in registers
#pragma unroll
                for (int k = 0; k < KM; ++k)
                {
                    if (k < K)
                    {
                        T* Wg = Wk.W + (eg + k / 6) * 18 + (k % 6) * 3;
#pragma unroll
                        for (int i = 0; i < 3; ++i)
                            Wg[i] += T(w * (L.Ji[k] * L.Jp[i] + L.Ji[kMaxActive + k] * L.Jp[3 + i]));
                    }
                }
            }
        }
        double Lf[6];
        finishPoint(P, Wk, b, t, X0, X1, X2, damping, flags, H, bp, Lf);
        for (int e = e0; e < e1; ++e)
        {
            T* We = Wk.W + e * 18;
#pragma unroll
            for (int a = 0; a < 6; ++a)
            {
                double u[3];
                crossRow(double(We[a * 3]), double(We[a * 3 + 1]), double(We[a * 3 + 2]), Lf, u);
#pragma unroll
                for (int i = 0; i < 3; ++i) We[a * 3 + i] = T(u[i]);
            }
        }
    }

    // One lane per observation: each wave holds whole points (its lanes are their observations, in track order).
    // A lane linearises its observation; the point's first lane sums H and bp over the point's lanes in lane
    // order (the order of the thread-per-point kernel) and finishes the point; then each lane writes its cross
    // block U = W L straight to its pose entry (one observation per camera and point), and the first lane of
    // each intrinsic group of the point sums the group's contributions in lane order and writes the group's
    // entries. Each cross block is written once: no clearing and no read-modify-write pass over W.
    constexpr int kLinThreads = 128;

    template <typename T, int KM>
    __global__ void __launch_bounds__(kLinThreads) k_linearize_lanes(Problem P, Work<T> Wk, double damping)
    {
        constexpr int kWaves = kLinThreads / kWave;
        constexpr int kG = KM > 0 ? 3 * KM : 1;  // group contribution: K rows of 3
        // H upper triangle (00 01 02 11 12 22) and bp terms of each lane; then the point factor at its first lane
        __shared__ double sHb[kWaves][9][kWave];
        __shared__ T sG[kWaves][kG][kWave];
        __shared__ int32_t sKey[kWaves][kWave];
        const int wave = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
        const int wv = blockIdx.x * kWaves + wave;
        if (wv >= P.nLinWaves) return;
        const int o = P.linLaneObs[wv * kWave + lane];
        const int b = P.linLanePoint[wv * kWave + lane];
        // segment of this lane's point: [first, end) lanes
        const int prev = __shfl(b, lane > 0 ? lane - 1 : 0, kWave);
        const unsigned long long heads = __ballot(b >= 0 && (lane == 0 || prev != b));
        const unsigned long long filled = __ballot(b >= 0);
        const unsigned long long upTo = lane == kWave - 1 ? ~0ull : (2ull << lane) - 1;
        const int first = kWave - 1 - __clzll(heads & upTo);
        const unsigned long long after = heads & ~upTo;
        const int end = after ? __ffsll(static_cast<long long>(after)) - 1 : __popcll(filled);
        const bool live = b >= 0 && !idleWindow(P, P.elimWin, b);

        double hb[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        T wp[18];
#pragma unroll
        for (int i = 0; i < 18; ++i) wp[i] = T(0);
        int ep = -1, eg = -1, K = 0, t = 0;
        double X0 = 0.0, X1 = 0.0, X2 = 0.0;
        int32_t* flags = nullptr;
        if (live)
        {
            flags = flagsOf(P, Wk, P.elimWin, b);
            t = P.elimTrack[b];
            X0 = P.pointX[3 * t];
            X1 = P.pointX[3 * t + 1];
            X2 = P.pointX[3 * t + 2];
            const int c = P.obsCam[o];
            const int g = P.camGroup[c];
            K = g >= 0 ? P.groupK[g] : 0;
            ep = P.obsPoseEntry[o];
            eg = P.obsGroupEntry[o];
            const int8_t* act = g >= 0 ? P.groupActive + g * kMaxActive : nullptr;
            Linearization L;
            if (!pbdev::linearize<KM>(P.cams[c], X0, X1, X2, P.obsUV[2 * o], P.obsUV[2 * o + 1], P.obsW[o], P.loss,
                                      P.lossScale, K, act, L, P.refParam != 0))
            {
                setFlag(flags, kFlagInvalid);
#pragma unroll
                for (int k = 0; k < kG; ++k) sG[wave][k][lane] = T(0);
            }
            else
            {
                if (P.adjoint)
                {
                    double q[3];
                    pbdev::adjointWeightRoot(L, P.obsW[o], P.loss, P.lossScale, q);
                    pbdev::applyWeightRoot<KM>(q, L);
                }
                const double w = L.w;
                int h = 0;
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = i; j < 3; ++j) hb[h++] = w * (L.Jp[i] * L.Jp[j] + L.Jp[3 + i] * L.Jp[3 + j]);
#pragma unroll
                for (int i = 0; i < 3; ++i) hb[6 + i] = w * (L.Jp[i] * L.r[0] + L.Jp[3 + i] * L.r[1]);
#pragma unroll
                for (int a = 0; a < 6; ++a)
#pragma unroll
                    for (int i = 0; i < 3; ++i) wp[a * 3 + i] = T(w * (L.Jc[a] * L.Jp[i] + L.Jc[6 + a] * L.Jp[3 + i]));
#pragma unroll
                for (int k = 0; k < KM; ++k)
#pragma unroll
                    for (int i = 0; i < 3; ++i)
                        sG[wave][k * 3 + i][lane] =
                            k < K ? T(w * (L.Ji[k] * L.Jp[i] + L.Ji[kMaxActive + k] * L.Jp[3 + i])) : T(0);
            }
        }
#pragma unroll
        for (int i = 0; i < 9; ++i) sHb[wave][i][lane] = hb[i];
        sKey[wave][lane] = eg;
        waveLdsFence();
        if (live && lane == first)
        {
            double H[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            double bp[3] = {0, 0, 0};
            for (int j = first; j < end; ++j)
            {
                int h = 0;
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int m = i; m < 3; ++m) H[i * 3 + m] += sHb[wave][h++][j];
#pragma unroll
                for (int i = 0; i < 3; ++i) bp[i] -= sHb[wave][6 + i][j];
            }
            H[3] = H[1];
            H[6] = H[2];
            H[7] = H[5];
            double Lf[6];
            finishPoint(P, Wk, b, t, X0, X1, X2, dampingOf(P, P.elimWin, b, damping), flags, H, bp, Lf);
            // only this lane reads its point's slots: they take the factor
#pragma unroll
            for (int i = 0; i < 6; ++i) sHb[wave][i][lane] = Lf[i];
        }
        waveLdsFence();
        if (!live) return;
        double Lf[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) Lf[i] = sHb[wave][i][first];
        if (ep >= 0)
        {
            T* We = Wk.W + size_t(ep) * 18;
#pragma unroll
            for (int a = 0; a < 6; ++a)
            {
                double u[3];
                crossRow(double(wp[a * 3]), double(wp[a * 3 + 1]), double(wp[a * 3 + 2]), Lf, u);
#pragma unroll
                for (int i = 0; i < 3; ++i) We[a * 3 + i] = T(u[i]);
            }
        }
        if (KM > 0 && eg >= 0)
        {
            bool head = true;
            for (int j = first; j < lane; ++j) head = head && sKey[wave][j] != eg;
            if (head)
            {
                T g[kG];
#pragma unroll
                for (int k = 0; k < kG; ++k) g[k] = T(0);
                for (int j = lane; j < end; ++j)
                    if (sKey[wave][j] == eg)
#pragma unroll
                        for (int k = 0; k < kG; ++k) g[k] += sG[wave][k][j];
                // the group's units hold rows k = 0 .. K-1 of the contribution, six per unit; the rest are zero
                constexpr int kRows = (KM + 5) / 6 * 6;
                const int units = (K + 5) / 6;
#pragma unroll
                for (int k = 0; k < kRows; ++k)
                {
                    if (k / 6 >= units) break;
                    double u[3] = {0.0, 0.0, 0.0};
                    if (k < KM && k < K) crossRow(double(g[3 * k]), double(g[3 * k + 1]), double(g[3 * k + 2]), Lf, u);
                    T* We = Wk.W + size_t(eg + k / 6) * 18 + (k % 6) * 3;
#pragma unroll
                    for (int i = 0; i < 3; ++i) We[i] = T(u[i]);
                }
            }
        }
    }

    // ---------------------------------------------------------------------------------------------
    // Camera side: J^T W J over each camera's observations (pose-pose, pose-intrinsic,
    // intrinsic-intrinsic), the direct right-hand side and the robust cost.
    template <typename T, int KM>
    struct CamAcc
    {
        static constexpr int kSym = KM * (KM + 1) / 2;
        static constexpr int kCount = 21 + 6 * KM + kSym;  // T accumulators
        T h[kCount > 0 ? kCount : 1];
        double b[6 + KM];
        double cost;
    };

    template <typename T, int KM>
    __global__ void __launch_bounds__(256) k_camera_direct(Problem P, Work<T> Wk)
    {
        using Acc = CamAcc<T, KM>;
        const int c = blockIdx.x;
        if (idleWindow(P, P.camWin, c)) return;
        const int pu = P.camUnit[c];
        const int g = P.camGroup[c];
        const int K = g >= 0 ? P.groupK[g] : 0;
        const int8_t* act = g >= 0 ? P.groupActive + g * kMaxActive : nullptr;
        Acc A;
#pragma unroll
        for (int i = 0; i < Acc::kCount; ++i) A.h[i] = T(0);
#pragma unroll
        for (int i = 0; i < 6 + KM; ++i) A.b[i] = 0.0;
        A.cost = 0.0;
        const Camera& cam = P.cams[c];
        Linearization L;
        // the camera's observations in camera-major copies (point, pixel, weight read contiguously); only the
        // points are gathered. The next observation's point index is loaded while this one is linearised.
        const int kEnd = P.camObsOffset[c + 1];
        int k = P.camObsOffset[c] + threadIdx.x;
        int tNext = k < kEnd ? P.camObsPoint[k] : 0;
        for (; k < kEnd; k += blockDim.x)
        {
            const int t = tNext;
            if (k + int(blockDim.x) < kEnd) tNext = P.camObsPoint[k + blockDim.x];
            const double ow = P.camObsW[k];
            if (!pbdev::linearize<KM>(cam, P.pointX[3 * t], P.pointX[3 * t + 1], P.pointX[3 * t + 2], P.camObsUV[2 * k],
                                  P.camObsUV[2 * k + 1], ow, P.loss, P.lossScale, KM > 0 ? K : 0, act, L,
                                  P.refParam != 0))
            {
                setFlag(flagsOf(P, Wk, P.camWin, c), kFlagInvalid);
                continue;
            }
            if (P.adjoint)
            {
                double q[3];
                pbdev::adjointWeightRoot(L, ow, P.loss, P.lossScale, q);
                pbdev::applyWeightRoot<KM>(q, L);
            }
            const double w = L.w;
            A.cost += L.cost;
            int n = 0;
            if (pu >= 0)
            {
#pragma unroll
                for (int i = 0; i < 6; ++i)
                {
                    A.b[i] -= w * (L.Jc[i] * L.r[0] + L.Jc[6 + i] * L.r[1]);
#pragma unroll
                    for (int j = i; j < 6; ++j) A.h[n++] += T(w * (L.Jc[i] * L.Jc[j] + L.Jc[6 + i] * L.Jc[6 + j]));
                }
            }
            if constexpr (KM > 0)
            {
                n = 21;
#pragma unroll
                for (int i = 0; i < 6; ++i)
#pragma unroll
                    for (int k2 = 0; k2 < KM; ++k2)
                        A.h[n++] += (pu >= 0 && k2 < K)
                                        ? T(w * (L.Jc[i] * L.Ji[k2] + L.Jc[6 + i] * L.Ji[kMaxActive + k2]))
                                        : T(0);
#pragma unroll
                for (int i = 0; i < KM; ++i)
                {
                    if (i < K) A.b[6 + i] -= w * (L.Ji[i] * L.r[0] + L.Ji[kMaxActive + i] * L.r[1]);
#pragma unroll
                    for (int j = i; j < KM; ++j)
                        A.h[n++] += (j < K) ? T(w * (L.Ji[i] * L.Ji[j] + L.Ji[kMaxActive + i] * L.Ji[kMaxActive + j]))
                                            : T(0);
                }
            }
        }
        // deterministic block reduction: wave sums by recursive halving, then the four wave results in order
        __shared__ double lds[4][Acc::kCount + 6 + KM + 1];
        const int lane = threadIdx.x & 63;
        const int wv = threadIdx.x >> 6;
        {
            constexpr int kR = halvingCount<Acc::kCount>();
            double hs[kR];
            halvingSum<Acc::kCount, 32, double, T>(A.h, lane, hs);
#pragma unroll
            for (int j = 0; j < kR; ++j)
            {
                const int slot = halvingSlot<Acc::kCount>(lane, j);
                if (slot >= 0) lds[wv][slot] = hs[j];
            }
        }
        {
            constexpr int kB = 6 + KM + 1;
            double bc[kB];
#pragma unroll
            for (int i = 0; i < 6 + KM; ++i) bc[i] = A.b[i];
            bc[6 + KM] = A.cost;
            constexpr int kR = halvingCount<kB>();
            double bs[kR];
            halvingSum<kB, 32, double, double>(bc, lane, bs);
#pragma unroll
            for (int j = 0; j < kR; ++j)
            {
                const int slot = halvingSlot<kB>(lane, j);
                if (slot >= 0) lds[wv][Acc::kCount + slot] = bs[j];
            }
        }
        __syncthreads();
        constexpr int kTotal = Acc::kCount + 6 + KM + 1;
        __shared__ double red[kTotal];
        for (int i = threadIdx.x; i < kTotal; i += blockDim.x)
        {
            double v = lds[0][i];
            for (int q = 1; q < int(blockDim.x >> 6); ++q) v += lds[q][i];
            red[i] = v;
        }
        __syncthreads();
        const int tid = threadIdx.x;
        if (tid == 0) Wk.camCost[c] = red[kTotal - 1];
        if (pu >= 0)
        {
            // pose-pose block (full 6x6 from the upper triangle) and pose right-hand side
            if (tid < 36)
            {
                const int i = tid / 6, j = tid % 6;
                const int a = i < j ? i : j, bb = i < j ? j : i;
                const int idx = a * 6 - a * (a - 1) / 2 + (bb - a);
                Wk.D[P.camDiagBlock[c] * 36 + tid] = red[idx];
            }
            if (tid < 6) Wk.directRhs[pu * 6 + tid] = red[Acc::kCount + tid];
            if constexpr (KM > 0)
            {
                // pose-intrinsic blocks: unit j holds active slots [6j, 6j+6)
                for (int q = tid; q < 36 * ((KM + 5) / 6); q += blockDim.x)
                {
                    const int j = q / 36, r = (q % 36) / 6, col = q % 6;
                    const int k2 = 6 * j + col;
                    if (6 * j < K)
                        Wk.D[P.camGroupBlock[c * kGroupUnits + j] * 36 + r * 6 + col] = k2 < K ? red[21 + r * KM + k2] : 0.0;
                }
            }
        }
        if constexpr (KM > 0)
        {
            if (K > 0)
            {
                // intrinsic-intrinsic partial and intrinsic right-hand side of this camera
                constexpr int kSym = Acc::kSym;
                for (int q = tid; q < kSym; q += blockDim.x) Wk.camGroupPartial[c * kGroupPartialStride + q] = red[21 + 6 * KM + q];
                for (int q = tid; q < KM; q += blockDim.x)
                    Wk.camGroupPartial[c * kGroupPartialStride + kSym + q] = red[Acc::kCount + 6 + q];
            }
        }
    }

    // Sums the camera partials of each intrinsic group in camera order, adds the priors
    // (ReferenceSchur.cpp:492-536) and writes the group blocks and right-hand side.
    template <typename T, int KM>
    __global__ void k_group_reduce(Problem P, Work<T> Wk)
    {
        const int g = blockIdx.x;
        const int K = P.groupK[g];
        if (K == 0 || idleWindow(P, P.groupWin, g)) return;
        constexpr int kSym = KM * (KM + 1) / 2;
        __shared__ double red[kSym + KM];
        const double* gp = P.groupParams + g * 13;
        // one wave per packed entry at a time: lane l sums the group's cameras l, l + 64, ... (their loads all in
        // flight), then a butterfly over the lanes; the order is fixed, so the sum is reproducible
        {
            const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
            const int k0 = P.groupCamOffset[g], k1 = P.groupCamOffset[g + 1];
            for (int q = wv; q < kSym + KM; q += nw)
            {
                double v = 0.0;
                for (int k = k0 + lane; k < k1; k += kWave) v += Wk.camGroupPartial[P.groupCams[k] * kGroupPartialStride + q];
                v = waveSum(v);
                if (lane == 0) red[q] = v;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0)
        {
            double prior = 0.0;
            for (int k = 0; k < K; ++k)
            {
                const int slot = P.groupActive[g * kMaxActive + k];
                const double is = P.groupInvSigma[g * 13 + slot];
                const double res = is * (gp[slot] - P.groupPrior[g * 13 + slot]);
                prior += 0.5 * res * res;
                // diagonal of the packed upper triangle: row k starts at k*KM - k*(k-1)/2
                red[k * KM - k * (k - 1) / 2] += is * is;
                red[kSym + k] -= is * res;
            }
            Wk.groupPriorCost[g] = prior;
        }
        __syncthreads();
        const int m = (K + 5) / 6;
        const int u0 = P.groupUnit[g];
        for (int q = threadIdx.x; q < m * m * 36; q += blockDim.x)
        {
            const int bi = q / (m * 36), bj = (q / 36) % m, r = (q % 36) / 6, col = q % 6;
            if (bj < bi) continue;
            const int i = 6 * bi + r, j = 6 * bj + col;
            double v = 0.0;
            if (i < K && j < K)
            {
                const int a = i < j ? i : j, bb = i < j ? j : i;
                v = red[a * KM - a * (a - 1) / 2 + (bb - a)];
            }
            Wk.D[P.groupPairBlock[(g * kGroupUnits + bi) * kGroupUnits + bj] * 36 + r * 6 + col] = v;
        }
        for (int q = threadIdx.x; q < m * 6; q += blockDim.x)
            Wk.directRhs[(u0 + q / 6) * 6 + q % 6] = q < K ? red[kSym + q] : 0.0;
    }

    // Objective of the linearisation point: camera costs in camera order plus group priors.
    template <typename T>
    __global__ void k_finish_cost(Problem P, Work<T> Wk)
    {
        __shared__ double part[256];
        double v = 0.0;
        for (int c = threadIdx.x; c < P.nCam; c += blockDim.x) v += Wk.camCost[c];
        // survey constraints of every active track (Assembly.cpp:656-690; ReferenceSchur.cpp:331-346)
        for (int k = threadIdx.x; k < P.nCons; k += blockDim.x)
        {
            const int t = P.consPoint[k];
            double r[3], cost, w;
            if (constraintAt(P, k, P.pointX[3 * t], P.pointX[3 * t + 1], P.pointX[3 * t + 2], r, cost, w)) v += cost;
        }
        part[threadIdx.x] = v;
        __syncthreads();
        if (threadIdx.x == 0)
        {
            double s = 0.0;
            for (int i = 0; i < int(blockDim.x); ++i) s += part[i];
            for (int g = 0; g < P.nGroup; ++g) s += P.groupK[g] > 0 ? Wk.groupPriorCost[g] : 0.0;
            Wk.scalars[kScalarCost] = s;
        }
    }

    // ---------------------------------------------------------------------------------------------
    // Schur camera-camera blocks: each work item is a run of (entry, entry) pairs of one block.
    template <typename T>
    __global__ void __launch_bounds__(256) k_schur_partial(Problem P, Work<T> Wk)
    {
        const int item = xcdLocal(blockIdx.x, gridDim.x) * (blockDim.x / kWave) + (threadIdx.x / kWave);
        if (item >= P.nItems) return;
        const int lane = threadIdx.x & 63;
        const int blk = P.itemBlock[item];
        if (P.nWin && idleWindow(P, P.unitWin, P.blockRow[blk])) return;
        const bool diag = P.blockRow[blk] == P.blockCol[blk];
        T acc[36];
#pragma unroll
        for (int i = 0; i < 36; ++i) acc[i] = T(0);
        double accr[6] = {0, 0, 0, 0, 0, 0};
        // the cross blocks hold U = W L (Hpp^-1 = L L^T), so each pair contributes -U1 U2^T and the diagonal blocks
        // the right-hand side correction U1 z (z = L^T bp)
        // the next round's pair indices are loaded while this round's cross blocks are multiplied
        const int k1 = P.itemBegin[item + 1];
        int k = P.itemBegin[item] + lane;
        int e1n = k < k1 ? P.pairE1[k] : 0, e2n = k < k1 ? P.pairE2[k] : 0;
        for (; k < k1; k += kWave)
        {
            const int e1 = e1n, e2 = e2n;
            if (k + kWave < k1)
            {
                e1n = P.pairE1[k + kWave];
                e2n = P.pairE2[k + kWave];
            }
            T u1[18], u2[18];
#pragma unroll
            for (int i = 0; i < 18; ++i) u1[i] = Wk.W[e1 * 18 + i];
            // a diagonal block's pairs are mostly an observation with itself: its cross block is read once
            if (e2 != e1)
            {
#pragma unroll
                for (int i = 0; i < 18; ++i) u2[i] = Wk.W[e2 * 18 + i];
            }
            else
            {
#pragma unroll
                for (int i = 0; i < 18; ++i) u2[i] = u1[i];
            }
#pragma unroll
            for (int a = 0; a < 6; ++a)
#pragma unroll
                for (int j = 0; j < 6; ++j)
                    acc[a * 6 + j] -= u1[a * 3] * u2[j * 3] + u1[a * 3 + 1] * u2[j * 3 + 1] + u1[a * 3 + 2] * u2[j * 3 + 2];
            if (diag)
            {
                const int p = P.entryPoint[e1];
                const double z0 = Wk.y[3 * p], z1 = Wk.y[3 * p + 1], z2 = Wk.y[3 * p + 2];
#pragma unroll
                for (int a = 0; a < 6; ++a)
                    accr[a] += double(u1[a * 3]) * z0 + double(u1[a * 3 + 1]) * z1 + double(u1[a * 3 + 2]) * z2;
            }
        }
        // recursive halving: each of 36 lanes (6 for the right-hand side) ends with one finished sum
        T mine;
        double mr;
        halvingSum<36, 32, T, T>(acc, lane, &mine);
        halvingSum<6, 32, double, double>(accr, lane, &mr);
        const int slot = halvingSlot<36>(lane, 0), slotR = halvingSlot<6>(lane, 0);
        if (slot >= 0) Wk.itemPartial[item * 36 + slot] = mine;
        if (slotR >= 0) Wk.itemPartialRhs[item * 6 + slotR] = mr;
    }

    __device__ __forceinline__ bool rigMode(const Problem& P) { return P.nCap + P.nSen > 0; }

    // Direct part + Schur items per block, diagonal finalize (scale by 1 + lambda, exact zeros -> 1; a rig
    // problem keeps the camera-level sums and finalizes its rig system instead), both triangles of the
    // block-sparse matrix, and the reduced right-hand side.
    template <typename T>
    __global__ void k_schur_finalize(Problem P, Work<T> Wk, double damping)
    {
        const int blk = blockIdx.x * (blockDim.x / kWave) + (threadIdx.x / kWave);
        if (blk >= P.nBlocks) return;
        const int lane = threadIdx.x & 63;
        const int u = P.blockRow[blk], v = P.blockCol[blk];
        if (idleWindow(P, P.unitWin, u)) return;
        damping = dampingOf(P, P.unitWin, u, damping);
        const bool diag = u == v;
        const int it0 = P.blockItemOffset[blk], it1 = P.blockItemOffset[blk + 1];
        if (it1 - it0 > kFinalizeSerialItems)
        {
            // many items: lane l sums items l, l + 64, ..., then a recursive-halving sum over the lanes; the lane
            // holding value i keeps element i (i < 36) or right-hand side entry i - 36 (36 <= i < 42)
            double acc[42];
#pragma unroll
            for (int i = 0; i < 42; ++i) acc[i] = 0.0;
            for (int it = it0 + lane; it < it1; it += kWave)
            {
#pragma unroll
                for (int i = 0; i < 36; ++i) acc[i] += double(Wk.itemPartial[it * 36 + i]);
                if (diag)
#pragma unroll
                    for (int a = 0; a < 6; ++a) acc[36 + a] += Wk.itemPartialRhs[it * 6 + a];
            }
            double mine;
            halvingSum<42, 32, double, double>(acc, lane, &mine);
            const int i = halvingSlot<42>(lane, 0);
            if (i >= 0 && i < 36)
            {
                double s = Wk.D[blk * 36 + i] + mine;
                if (diag && (i / 6) == (i % 6) && !rigMode(P)) s = s == 0.0 ? 1.0 : s * (1.0 + damping);
                Wk.S[P.blockPosUp[blk] * 36 + i] = T(s);
                if (!diag) Wk.S[P.blockPosLo[blk] * 36 + (i % 6) * 6 + i / 6] = T(s);
            }
            else if (diag && i >= 36)
                Wk.rhs[u * 6 + i - 36] = Wk.directRhs[u * 6 + i - 36] - mine;
            return;
        }
        if (lane < 36)
        {
            double s = Wk.D[blk * 36 + lane];
            // at most kFinalizeSerialItems items: their loads all in flight, then summed in item order
            T v[kFinalizeSerialItems];
#pragma unroll
            for (int q = 0; q < kFinalizeSerialItems; ++q)
                if (it0 + q < it1) v[q] = Wk.itemPartial[(it0 + q) * 36 + lane];
#pragma unroll
            for (int q = 0; q < kFinalizeSerialItems; ++q)
                if (it0 + q < it1) s += double(v[q]);
            if (diag && (lane / 6) == (lane % 6) && !rigMode(P)) s = s == 0.0 ? 1.0 : s * (1.0 + damping);
            Wk.S[P.blockPosUp[blk] * 36 + lane] = T(s);
            if (!diag) Wk.S[P.blockPosLo[blk] * 36 + (lane % 6) * 6 + lane / 6] = T(s);
        }
        else if (diag && lane < 42)
        {
            const int a = lane - 36;
            double s = 0.0;
            for (int it = it0; it < it1; ++it) s += Wk.itemPartialRhs[it * 6 + a];
            Wk.rhs[u * 6 + a] = Wk.directRhs[u * 6 + a] - s;
        }
    }

    // 6x6 block-Jacobi preconditioner: Cholesky of each diagonal block in FP64, explicit inverse.
    template <typename T>
    __global__ void k_block_jacobi(Problem P, Work<T> Wk)
    {
        const int u = blockIdx.x * blockDim.x + threadIdx.x;
        if (u >= P.nUnits) return;
        const T* Sb = Wk.S + P.diagPos[u] * 36;
        double Lm[36];
#pragma unroll
        for (int i = 0; i < 36; ++i) Lm[i] = double(Sb[i]);
        bool ok = true;
#pragma unroll
        for (int j = 0; j < 6; ++j)
        {
            double d = Lm[j * 6 + j];
            for (int k = 0; k < j; ++k) d -= Lm[j * 6 + k] * Lm[j * 6 + k];
            if (!(d > 0.0) || !pbdev::finite(d))
            {
                ok = false;
                d = 1.0;
            }
            const double ljj = sqrt(d);
            Lm[j * 6 + j] = ljj;
            for (int i = j + 1; i < 6; ++i)
            {
                double s = Lm[i * 6 + j];
                for (int k = 0; k < j; ++k) s -= Lm[i * 6 + k] * Lm[j * 6 + k];
                Lm[i * 6 + j] = s / ljj;
            }
        }
        if (!ok) setFlag(Wk.flags, kFlagNotSpd);
        // inverse = L^-T L^-1, column by column
        double inv[36];
#pragma unroll
        for (int col = 0; col < 6; ++col)
        {
            double y[6];
            for (int i = 0; i < 6; ++i)
            {
                double s = (i == col) ? 1.0 : 0.0;
                for (int k = 0; k < i; ++k) s -= Lm[i * 6 + k] * y[k];
                y[i] = s / Lm[i * 6 + i];
            }
            for (int i = 5; i >= 0; --i)
            {
                double s = y[i];
                for (int k = i + 1; k < 6; ++k) s -= Lm[k * 6 + i] * y[k];
                y[i] = s / Lm[i * 6 + i];
            }
            for (int i = 0; i < 6; ++i) inv[i * 6 + col] = y[i];
        }
#pragma unroll
        for (int i = 0; i < 36; ++i) Wk.Minv[u * 36 + i] = T(inv[i]);
    }

    // ---------------------------------------------------------------------------------------------
    // PCG on the reduced system. Dot products accumulate in FP64 as per-block partials that every
    // consumer sums in index order.
    __device__ __forceinline__ double blockSum256(double v, double* lds)
    {
        v = waveSum(v);
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        if (lane == 0) lds[wv] = v;
        __syncthreads();
        double s = 0.0;
        if (threadIdx.x == 0)
            for (int i = 0; i < int(blockDim.x >> 6); ++i) s += lds[i];
        return s;
    }

    // Sum of n per-block partials by wave 0 of the calling block (fixed lane order and butterfly),
    // broadcast to the block. Every block gets bitwise the same value.
    __device__ __forceinline__ double blockSumPartials(const double* part, int n, double* slot)
    {
        if ((threadIdx.x >> 6) == 0)
        {
            double v = 0.0;
            for (int i = threadIdx.x; i < n; i += kWave) v += part[i];
            v = waveSum(v);
            if (threadIdx.x == 0) *slot = v;
        }
        __syncthreads();
        const double r = *slot;
        __syncthreads();
        return r;
    }

    // ---------------------------------------------------------------------------------------------
    // Fused PCG iteration (Chronopoulos-Gear recurrences): one launch per iteration.
    //
    // State entering iteration it (parity rd = it & 1), per unit row: r_it, u_it = M r_it, w_it = S u~_it,
    // Mw_it = M w_it, s_{it-1} = S p_{it-1} (by recurrence), Ms_{it-1} = M s_{it-1}, p_{it-1}, x_it, and the
    // row partials of gamma = (r, u), delta = (w, u), rho = (r, r). Every block sums the partials in the same
    // order, so all blocks derive bitwise the same alpha, beta and stop decision. The owner wave of a row
    // updates s = w + beta s, p = u + beta p, x += alpha p, r -= alpha s and stores u = M r, M s and, after the
    // row's SpMV, w and M w (M is block diagonal, so the owner holds its block). The SpMV w_{it+1} = S u~_{it+1}
    // needs u_{it+1} of every column unit; a wave evaluates it as u~ = u - alpha (Mw + beta Ms) from the
    // column's previous-parity values, which equals M r_{it+1} up to the rounding of one step (no
    // preconditioner block is read per stored block). r, u, w, Mw, s, Ms are double-buffered by parity; p and
    // x are touched by their owner only. it = -1 primes w_0 and the partials of iteration 0 (alpha = beta = 0).
    template <typename T>
    __device__ __forceinline__ T fmaT(T a, T b, T c);
    template <>
    __device__ __forceinline__ float fmaT<float>(float a, float b, float c)
    {
        return __builtin_fmaf(a, b, c);
    }
    template <>
    __device__ __forceinline__ double fmaT<double>(double a, double b, double c)
    {
        return __builtin_fma(a, b, c);
    }

    // x_{it+1} = x - alpha (y + beta z): r from (r, w, s) and u from (u, Mw, Ms)
    template <typename T>
    __device__ __forceinline__ T cgNext(T x, T y, T z, T a, T bt)
    {
        return fmaT(-a, fmaT(bt, z, y), x);
    }

    // sum_j m[j] v_j where v_j is held by lane base + j. Every lane of the wave must call it (it shuffles).
    template <typename T>
    __device__ __forceinline__ T rowTimes(const T (&m)[6], T v, int base)
    {
        T acc = T(0);
#pragma unroll
        for (int j = 0; j < 6; ++j) acc = fmaT(m[j], __shfl(v, min(base + j, kWave - 1), kWave), acc);
        return acc;
    }

    //
    // Iteration number: itRel, plus *itBase when itBase is given (launches replayed from a captured graph take
    // their absolute iteration from a device counter that k_pcg_advance moves on; the counter is even at every
    // replay and the graph holds an even number of launches, so itRel fixes the parity). The counter is loaded
    // with the other first-round-trip loads, and what depends on the iteration number comes after them.
    template <typename T>
    __global__ void __launch_bounds__(64 * kCgWaves) k_pcg_iter(Problem P, Work<T> Wk, int itRel, const int32_t* itBase,
                                                                int maxIt, int nParts, double relTol, double absTol,
                                                                double modelEta)
    {
        __shared__ double red3[kCgWaves][3];
        __shared__ T red[kCgWaves][kWave];
        __shared__ T segq[kCgWaves][6];
        const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
        const int rd = itRel & 1, wr = rd ^ 1;  // prime (it = -1) reads parity 1 and writes parity 0
        const bool started = itBase != nullptr || itRel >= 0;  // not the priming launch (launch-uniform)
        const int i6 = lane % 6, bslot = lane / 6, base = 6 * bslot;
        constexpr int kCgPrefetch = cgPrefetch<T>();

        // Round trip 1: everything that depends on nothing - the stop flag of the previous launch, the row
        // partials (every thread of the block loads its share), the scalars, this wave's slot and the column ids
        // of its first steps. A launch reads only what earlier launches wrote, so all blocks and waves see the
        // same values. Loads are unconditional (clamped to valid addresses and masked at their use) so that the
        // compiler issues them together instead of waiting at branch joins.
        const int slot = blockIdx.x * kCgWaves + wv;
        const int prevDone = Wk.iflags[kIflagDone0 + rd];
        const int it = itRel + (itBase ? *itBase : 0);
        const int4 si = P.cgSlot[slot];
        int col[kCgPrefetch];
#pragma unroll
        for (int st = 0; st < kCgPrefetch; ++st) col[st] = P.cgColDirect[slot * kCgDirect + st * 10 + min(bslot, 9)];
        double pg[kCgPartLoads], pd[kCgPartLoads], pq[kCgPartLoads];
        double thr2 = 0.0, qIt = 0.0, dqPrev = 0.0, invGammaPrev = 0.0, invAlphaPrev = 0.0;
        const double* part = Wk.cgPart[rd];
#pragma unroll
        for (int k = 0; k < kCgPartLoads; ++k)
        {
            const int i = min(int(threadIdx.x) + k * 64 * kCgWaves, nParts - 1);
            pg[k] = part[3 * i];
            pd[k] = part[3 * i + 1];
            pq[k] = part[3 * i + 2];
        }
        if (started)
        {
            thr2 = Wk.scalars[kScalarThreshold2];
            qIt = Wk.scalars[kScalarQ0 + rd];
            dqPrev = Wk.scalars[kScalarDq0 + rd];
            invGammaPrev = Wk.scalars[kScalarInvGamma0 + rd];
            invAlphaPrev = Wk.scalars[kScalarInvAlpha0 + rd];
        }
        const int row = si.x, kb = si.y, ke = si.z, lead = si.w;  // lead: waves of the row on its first wave
        const int nSteps = row >= 0 ? (ke - kb + 9) / 10 : 0;      // uniform over the wave
        const int rowSafe = row >= 0 ? row : 0;
        const T* uO = Wk.cgU[rd];
        const T* mwO = Wk.cgMW[rd];
        const T* msO = Wk.cgMS[rd];

        // Round trip 2: S rows and the column units' u, Mw, Ms, and the owner's own entries. The first kCgAlways
        // steps and the owner entries are loaded by every wave (so the compiler can count them and the partials
        // below wait for nothing else); further steps only by waves that have them. Lanes past the segment load
        // their own row's data and are masked at the use.
        bool act[kCgPrefetch];
        T Srow[kCgPrefetch][6], uC[kCgPrefetch], mwC[kCgPrefetch], msC[kCgPrefetch];
        auto loadStep = [&](int st) {
            const int k = act[st] ? kb + st * 10 + bslot : kb;
            const int c = act[st] ? col[st] : rowSafe;
#pragma unroll
            for (int j = 0; j < 6; ++j) Srow[st][j] = Wk.S[size_t(k) * 36 + i6 * 6 + j];
            uC[st] = uO[c * 6 + i6];
            mwC[st] = mwO[c * 6 + i6];
            msC[st] = msO[c * 6 + i6];
        };
#pragma unroll
        for (int st = 0; st < kCgPrefetch; ++st) act[st] = lane < 60 && col[st] >= 0;
#pragma unroll
        for (int st = 0; st < kCgAlways; ++st) loadStep(st);
        const bool own = lead && lane < 6;
        const int oi = lane < 6 ? lane : 0;
        const int gi = rowSafe * 6 + oi;
        const T rOld = Wk.cgR[rd][gi], uOld = uO[gi], wOld = Wk.cgW[rd][gi];
        const T sOld = Wk.cgS[rd][gi], pOld = Wk.p[gi], xOld = Wk.x[gi];
        T mOwn[6];
#pragma unroll
        for (int j = 0; j < 6; ++j) mOwn[j] = Wk.Minv[rowSafe * 36 + oi * 6 + j];
#pragma unroll
        for (int st = kCgAlways; st < kCgPrefetch; ++st)
            if (st < nSteps) loadStep(st);
        double g = 0.0, d = 0.0, q = 0.0;
        if (started)
        {
#pragma unroll
            for (int k = 0; k < kCgPartLoads; ++k)
                if (int(threadIdx.x) + k * 64 * kCgWaves < nParts)
                {
                    g += pg[k];
                    d += pd[k];
                    q += pq[k];
                }
            for (int i = threadIdx.x + kCgPartLoads * 64 * kCgWaves; i < nParts; i += 64 * kCgWaves)
            {
                g += part[3 * i];
                d += part[3 * i + 1];
                q += part[3 * i + 2];
            }
        }

        if (prevDone)
        {
            if (blockIdx.x == 0 && threadIdx.x == 0) Wk.iflags[kIflagDone0 + wr] = 1;
            return;
        }
        double alpha = 0.0, beta = 0.0;
        if (started)
        {
            g = waveSum(g);
            d = waveSum(d);
            q = waveSum(q);
            if (lane == 0)
            {
                red3[wv][0] = g;
                red3[wv][1] = d;
                red3[wv][2] = q;
            }
            __syncthreads();
            double gamma = 0.0, delta = 0.0, rho = 0.0;
#pragma unroll
            for (int w2 = 0; w2 < kCgWaves; ++w2)
            {
                gamma += red3[w2][0];
                delta += red3[w2][1];
                rho += red3[w2][2];
            }
            // the threshold, the reciprocals and the model values come from the previous launch, so that the
            // only division on the path of every block is alpha's
            double thrSq = thr2;
            if (it == 0)
            {
                const double thr = fmax(relTol * sqrt(rho), absTol);
                thrSq = thr * thr;
            }
            const bool conv = rho <= thrSq;
            // truncated-Newton test of the previous step on the quadratic model Q(x) = x'Sx/2 - b'x:
            // it (Q_it - Q_{it-1}) / Q_it < eta, multiplied out with Q_it < 0
            const bool modelStop = !conv && it >= kCgModelMinIterations && modelEta > 0.0 && qIt < 0.0 &&
                                   double(it) * dqPrev > modelEta * qIt;
            double den = delta;
            if (it > 0)
            {
                beta = gamma * invGammaPrev;
                den = delta - beta * gamma * invAlphaPrev;
            }
            alpha = gamma / den;
            const bool breakdown = !conv && !modelStop &&
                                   (!(den > 0.0) || !(gamma > 0.0) || !pbdev::finite(alpha) || !pbdev::finite(beta));
            const bool stop = conv || modelStop || breakdown || it >= maxIt;
            if (blockIdx.x == 0 && threadIdx.x == 0)
            {
                Wk.scalars[kScalarRnorm] = sqrt(rho);
                if (it == 0)
                {
                    Wk.scalars[kScalarThreshold] = sqrt(thrSq);
                    Wk.scalars[kScalarThreshold2] = thrSq;
                    Wk.scalars[kScalarBnorm] = sqrt(rho);
                }
                if (stop)
                {
                    Wk.iflags[kIflagIters] = it;
                    Wk.iflags[kIflagModelStop] = modelStop ? 1 : 0;
                    Wk.iflags[kIflagBreakdown] = breakdown ? 1 : 0;
                    Wk.iflags[kIflagDone0 + wr] = 1;
                }
                else
                {
                    const double dq = -0.5 * alpha * gamma;
                    Wk.scalars[kScalarQ0 + wr] = qIt + dq;
                    Wk.scalars[kScalarDq0 + wr] = dq;
                    Wk.scalars[kScalarInvGamma0 + wr] = 1.0 / gamma;
                    Wk.scalars[kScalarInvAlpha0 + wr] = 1.0 / alpha;
                }
            }
            if (stop) return;  // every block takes the same decision
        }
        else if (blockIdx.x == 0 && threadIdx.x == 0)
        {
            Wk.scalars[kScalarQ0] = 0.0;
            Wk.scalars[kScalarDq0] = 0.0;
        }
        const T a = T(alpha), bt = T(beta);

        // SpMV segment: w_{it+1} = S u_{it+1}, ten 6x6 blocks per step, lane (bslot, i6) takes row i6 of block bslot
        T acc = T(0);
#pragma unroll
        for (int st = 0; st < kCgPrefetch; ++st)
        {
            if (st < nSteps)
            {
                const T un = cgNext(uC[st], mwC[st], msC[st], a, bt);
                const T sum = rowTimes(Srow[st], un, base);
                if (act[st]) acc += sum;
            }
        }
        for (int st = kCgPrefetch; st < nSteps; ++st)  // rows longer than the prefetch window
        {
            const int k = kb + st * 10 + bslot;
            const bool on = lane < 60 && k < ke;
            const int kk = on ? k : kb;
            const int c = on ? P.colIdx[kk] : rowSafe;
            T Sr[6];
#pragma unroll
            for (int j = 0; j < 6; ++j) Sr[j] = Wk.S[size_t(kk) * 36 + i6 * 6 + j];
            const T un = cgNext(uO[c * 6 + i6], mwO[c * 6 + i6], msO[c * 6 + i6], a, bt);
            const T sum = rowTimes(Sr, un, base);
            if (on) acc += sum;
        }

        // owner update of row `row` (lanes 0..5 hold its six entries). The owner stores u = M r and M s computed
        // from its own block, so the column recurrence above starts every iteration from exact values and its
        // rounding does not accumulate.
        const T sn = fmaT(bt, sOld, wOld);
        const T rn = cgNext(rOld, wOld, sOld, a, bt);
        const T un = rowTimes(mOwn, rn, 0);
        const T msn = rowTimes(mOwn, sn, 0);
        if (own)
        {
            const T pn = fmaT(bt, pOld, uOld);
            Wk.cgS[wr][gi] = sn;
            Wk.cgMS[wr][gi] = msn;
            Wk.cgR[wr][gi] = rn;
            Wk.cgU[wr][gi] = un;
            Wk.p[gi] = pn;
            Wk.x[gi] = fmaT(a, pn, xOld);
        }
        // row sums: ten block slots of this wave (same wave, no barrier), then the row's waves
        red[wv][lane] = acc;
        __builtin_amdgcn_wave_barrier();
        if (lane < 6)
        {
            T s6 = T(0);
#pragma unroll
            for (int b = 0; b < 10; ++b) s6 += red[wv][b * 6 + lane];
            segq[wv][lane] = s6;
        }
        __syncthreads();
        if (lead)
        {
            T wsum = T(0);
            if (lane < 6)
                for (int w2 = 0; w2 < lead; ++w2) wsum += segq[wv + w2][lane];
            const T mw = rowTimes(mOwn, wsum, 0);  // M w_{it+1}, row of the owner's block
            double gg = 0.0, dd = 0.0, rr = 0.0;
            if (own)
            {
                Wk.cgW[wr][gi] = wsum;
                Wk.cgMW[wr][gi] = mw;
                gg = double(rn) * double(un);
                dd = double(wsum) * double(un);
                rr = double(rn) * double(rn);
            }
            // lanes 0..7 (6 and 7 hold zero): three butterfly steps, lane 0 writes the row partials
#pragma unroll
            for (int off = 1; off < 8; off <<= 1)
            {
                gg += __shfl_xor(gg, off, kWave);
                dd += __shfl_xor(dd, off, kWave);
                rr += __shfl_xor(rr, off, kWave);
            }
            if (lane == 0)
            {
                double* outp = Wk.cgPart[wr];
                outp[3 * row] = gg;
                outp[3 * row + 1] = dd;
                outp[3 * row + 2] = rr;
            }
        }
    }

    // r = b, u = M b and zero w, Mw, s, Ms, p, x (parity 1, read by the priming launch); resets the PCG flags.
    template <typename T>
    __global__ void k_pcg_init(Problem P, Work<T> Wk)
    {
        const int i = blockIdx.x * blockDim.x + threadIdx.x;
        if (i < P.nUnits * 6)
        {
            const int u = i / 6, row = i % 6;
            T z = T(0);
#pragma unroll
            for (int j = 0; j < 6; ++j) z = fmaT(Wk.Minv[u * 36 + row * 6 + j], T(Wk.rhs[u * 6 + j]), z);
            Wk.cgR[1][i] = T(Wk.rhs[i]);
            Wk.cgU[1][i] = z;
            Wk.cgW[1][i] = T(0);
            Wk.cgMW[1][i] = T(0);
            Wk.cgS[1][i] = T(0);
            Wk.cgMS[1][i] = T(0);
            Wk.p[i] = T(0);
            Wk.x[i] = T(0);
        }
        if (i == 0)
        {
            Wk.iflags[kIflagDone0] = 0;
            Wk.iflags[kIflagDone1] = 0;
            Wk.iflags[kIflagIters] = 0;
            Wk.iflags[kIflagBreakdown] = 0;
            Wk.iflags[kIflagModelStop] = 0;
            Wk.iflags[kIflagItBase] = 0;
        }
    }

    // Moves the iteration counter of graph-replayed PCG launches on by the launches of one replay.
    template <typename T>
    __global__ void k_pcg_advance(Work<T> Wk, int launches)
    {
        if (threadIdx.x == 0) Wk.iflags[kIflagItBase] += launches;
    }

    // ---------------------------------------------------------------------------------------------
    // Back-substitution dp = Hinv (bp - sum_e W_e^T dc_e); partials of bp.dp and |dp|^2.
    template <typename T>
    __global__ void k_back_substitute(Problem P, Work<T> Wk)
    {
        __shared__ double lds[8];
        const int b = blockIdx.x * blockDim.x + threadIdx.x;
        double dd = 0.0, ss = 0.0;
        if (b < P.nElim && !idleWindow(P, P.elimWin, b))
        {
            // dp = Hpp^-1 (bp - sum W^T dc) = L (z - sum U^T dc)
            double z[3] = {Wk.y[3 * b], Wk.y[3 * b + 1], Wk.y[3 * b + 2]};
            for (int e = P.entryOffset[b]; e < P.entryOffset[b + 1]; ++e)
            {
                const int u = P.entryUnit[e];
                const T* Up = Wk.W + e * 18;
#pragma unroll
                for (int a = 0; a < 6; ++a)
                {
                    const double xa = double(Wk.x[u * 6 + a]);
#pragma unroll
                    for (int i = 0; i < 3; ++i) z[i] -= double(Up[a * 3 + i]) * xa;
                }
            }
            const T* Lp = Wk.Hinv + 9 * b;
            const double d[3] = {double(Lp[0]) * z[0], double(Lp[1]) * z[0] + double(Lp[2]) * z[1],
                                 double(Lp[3]) * z[0] + double(Lp[4]) * z[1] + double(Lp[5]) * z[2]};
#pragma unroll
            for (int i = 0; i < 3; ++i)
            {
                Wk.dp[3 * b + i] = d[i];
                dd += Wk.bp[3 * b + i] * d[i];
                ss += d[i] * d[i];
            }
        }
        const double s1 = blockSum256(dd, lds);
        __syncthreads();
        const double s2 = blockSum256(ss, lds + 4);
        if (threadIdx.x == 0)
        {
            Wk.partDd[blockIdx.x] = s1;
            Wk.partSs[blockIdx.x] = s2;
        }
    }

    // Directional decrease dc.directRhs + sum bp.dp and the step RMS (BundleAdjustPlaMatrix.cpp:77-158).
    template <typename T>
    __global__ void k_step_stats(Problem P, Work<T> Wk, int nParts)
    {
        __shared__ double lds[8];
        double dd = 0.0, gs = 0.0;
        for (int u = threadIdx.x; u < P.nUnits; u += blockDim.x)
        {
            const int lim = P.unitWidth[u];
#pragma unroll
            for (int a = 0; a < 6; ++a)
            {
                const double x = double(Wk.x[u * 6 + a]);
                dd += x * Wk.directRhs[u * 6 + a];
                if (a < lim) gs += x * x;
            }
        }
        __shared__ double slot[2];
        const double ddp = blockSumPartials(Wk.partDd, nParts, slot);
        const double ssp = blockSumPartials(Wk.partSs, nParts, slot + 1);
        const double s1 = blockSum256(dd, lds);
        __syncthreads();
        const double s2 = blockSum256(gs, lds + 4);
        if (threadIdx.x == 0)
        {
            Wk.scalars[kScalarDd] = s1 + ddp;
            Wk.scalars[kScalarRmsGlobal] = P.globalDim > 0 ? sqrt(s2 / P.globalDim) : 0.0;
            Wk.scalars[kScalarRmsPoint] = P.nElim > 0 ? sqrt(ssp / (3.0 * P.nElim)) : 0.0;
        }
    }

    // ---------------------------------------------------------------------------------------------
    // Trial states for a batch of step lengths alpha_k: candidate group parameters (clamped) and cameras.
    template <typename T>
    __global__ void k_trial_groups(Problem P, Work<T> Wk, const double* alphas, int nAlpha)
    {
        const int id = blockIdx.x * blockDim.x + threadIdx.x;
        if (id >= nAlpha * P.nGroup) return;
        const int ai = id / P.nGroup, g = id % P.nGroup;
        double a = alphas ? alphas[ai] : 0.0;
        if (P.nWin)
        {
            const int w = P.groupWin[g];
            if (ai >= P.winNAlpha[w]) return;
            a = P.winAlpha[w * kMaxAlpha + ai];
        }
        const uint32_t en = P.groupEnabled[g];
        double* out = Wk.candGroupParams + (ai * P.nGroup + g) * 13;
        double prior = 0.0;
        for (int s = 0; s < 13; ++s)
        {
            double v = P.groupParams[g * 13 + s];
            if ((en >> s) & 1u)
            {
                const int k = P.slotToCompact[g * 13 + s];
                const int unit = P.groupUnit[g] + k / 6;
                v = fmin(fmax(v + a * double(Wk.x[unit * 6 + k % 6]), P.groupLower[g * 13 + s]),
                         P.groupUpper[g * 13 + s]);
                const double res = P.groupInvSigma[g * 13 + s] * (v - P.groupPrior[g * 13 + s]);
                prior += 0.5 * res * res;
            }
            out[s] = v;
        }
        Wk.candPriorCost[ai * P.nGroup + g] = P.groupK[g] > 0 ? prior : 0.0;
    }

    template <typename T>
    __global__ void k_trial_cameras(Problem P, Work<T> Wk, const double* alphas, int nAlpha)
    {
        const int id = blockIdx.x * blockDim.x + threadIdx.x;
        if (id >= nAlpha * P.nCam) return;
        const int ai = id / P.nCam, c = id % P.nCam;
        double a = alphas ? alphas[ai] : 0.0;
        if (P.nWin)
        {
            const int w = P.camWin[c];
            if (ai >= P.winNAlpha[w]) return;
            a = P.winAlpha[w * kMaxAlpha + ai];
        }
        const int u = P.camUnit[c];
        double R[9], Cc[3];
        const double* R0 = P.camR + 9 * c;
        const double* C0 = P.camC + 3 * c;
        const int anchor = gaugeAnchorOf(P, c);
        double anchorC[3], scaleC[3];
        if (anchor >= 0)
        {
            for (int i = 0; i < 3; ++i)
            {
                anchorC[i] = P.camC[3 * anchor + i];
                scaleC[i] = C0[i];
            }
        }
        if (rigMode(P))
        {
            // the composition of the camera's candidate capture and sensor (a camera whose capture and sensor
            // are both held keeps its pose)
            const int k = P.camCapture[c], q = P.camSensor[c];
            if (P.capUnit[k] >= 0 || P.senUnit[q] >= 0)
            {
                const long ck = long(ai) * P.nCap + k, cq = long(ai) * P.nSen + q;
                pbdev::composeRig(Wk.candCapR + ck * 9, Wk.candCapC + ck * 3, Wk.candSenR + cq * 9,
                                  Wk.candSenC + cq * 3, R, Cc);
            }
            else
            {
                for (int i = 0; i < 9; ++i) R[i] = R0[i];
                for (int i = 0; i < 3; ++i) Cc[i] = C0[i];
            }
        }
        else if (u >= 0)
        {
            double d[6];
#pragma unroll
            for (int i = 0; i < 6; ++i) d[i] = a * double(Wk.x[u * 6 + i]);
            if (P.refParam)
                pbdev::retractRotation(R0, d[0], d[1], d[2], R);
            else
                pbdev::retractExp(R0, d[0], d[1], d[2], R);
            // the default parameterisation applies every centre step in world axes, the gauge scale camera's
            // included (applyDeltaPose, Assembly.cpp:1133-1136)
            if (anchor >= 0 && P.refParam)
            {
                // centre moves along the tangent basis, then back onto the baseline sphere
                // (Assembly.cpp:101-128)
                const Camera& cur = P.cams[c];
                const double o[3] = {anchorC[0], -anchorC[1], -anchorC[2]};
                const double q[3] = {scaleC[0], -scaleC[1], -scaleC[2]};
                // cur.basis holds D * basis; recover basis in target coordinates
                const double b0[3] = {cur.basis[0], -cur.basis[1], -cur.basis[2]};
                const double b1[3] = {cur.basis[3], -cur.basis[4], -cur.basis[5]};
                double dir[3];
                for (int i = 0; i < 3; ++i) dir[i] = q[i] + b0[i] * d[3] + b1[i] * d[4] - o[i];
                const double len = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
                if (len < 1e-20)
                {
                    Cc[0] = anchorC[0];
                    Cc[1] = anchorC[1];
                    Cc[2] = anchorC[2];
                }
                else
                {
                    const double s = gaugeBaselineOf(P, c) / len;
                    const double t[3] = {o[0] + dir[0] * s, o[1] + dir[1] * s, o[2] + dir[2] * s};
                    Cc[0] = t[0];
                    Cc[1] = -t[1];
                    Cc[2] = -t[2];
                }
            }
            else
            {
                for (int i = 0; i < 3; ++i) Cc[i] = C0[i] + d[3 + i];
            }
        }
        else
        {
            for (int i = 0; i < 9; ++i) R[i] = R0[i];
            for (int i = 0; i < 3; ++i) Cc[i] = C0[i];
        }
        double* outR = Wk.candR + (ai * P.nCam + c) * 9;
        double* outC = Wk.candC + (ai * P.nCam + c) * 3;
        for (int i = 0; i < 9; ++i) outR[i] = R[i];
        for (int i = 0; i < 3; ++i) outC[i] = Cc[i];
        // effective camera of the candidate (the tangent basis is not needed for cost passes)
        makeCamera(P, c, R, Cc, Wk.candGroupParams + ai * P.nGroup * 13, nullptr, nullptr,
                   Wk.candCams[ai * P.nCam + c]);
    }

    // Objective of each candidate: one pass over the active observations (track order).
    template <typename T, int NA>
    __global__ void __launch_bounds__(256) k_trial_cost(Problem P, Work<T> Wk, const double* alphas, int nAlpha)
    {
        double acc[NA];
        int bad[NA];
#pragma unroll
        for (int a = 0; a < NA; ++a)
        {
            acc[a] = 0.0;
            bad[a] = 0;
        }
        double al[NA];
#pragma unroll
        for (int a = 0; a < NA; ++a) al[a] = a < nAlpha ? alphas[a] : 0.0;
        for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < P.nObs; o += gridDim.x * blockDim.x)
            trialObservation<NA>(P, Wk.candCams, Wk.dp, o, al, nAlpha, acc, bad);
        for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < P.nCons; k += gridDim.x * blockDim.x)
            trialConstraint<NA>(P, Wk.dp, k, al, nAlpha, acc);
        __shared__ double lds[NA][4];
        __shared__ int lbad[NA][4];
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        {
            double s;
            halvingSum<NA, 32, double, double>(acc, lane, &s);
            const int a = halvingSlot<NA>(lane, 0);
            if (a >= 0) lds[a][wv] = s;
        }
#pragma unroll
        for (int a = 0; a < NA; ++a)
        {
            const int anyBad = __any(bad[a]) ? 1 : 0;
            if (lane == 0) lbad[a][wv] = anyBad;
        }
        __syncthreads();
        if (threadIdx.x < NA && int(threadIdx.x) < nAlpha)
        {
            const int a = threadIdx.x;
            double s = 0.0;
            int bb = 0;
            for (int q = 0; q < 4; ++q)
            {
                s += lds[a][q];
                bb |= lbad[a][q];
            }
            Wk.trialPart[blockIdx.x * NA + a] = s;
            Wk.trialBad[blockIdx.x * NA + a] = bb;
        }
    }

    template <typename T, int NA>
    __global__ void __launch_bounds__(64 * NA) k_trial_reduce(Problem P, Work<T> Wk, int nAlpha, int nParts)
    {
        // one wave per step length: lane-strided partial sums in a fixed order, then the wave butterfly
        const int a = threadIdx.x >> 6, lane = threadIdx.x & 63;
        if (a >= nAlpha) return;
        double s = 0.0;
        int bb = 0;
        for (int i = lane; i < nParts; i += kWave)
        {
            s += Wk.trialPart[i * NA + a];
            bb |= Wk.trialBad[i * NA + a];
        }
        s = waveSum(s);
        const bool bad = __any(bb != 0);
        if (lane == 0)
        {
            for (int g = 0; g < P.nGroup; ++g) s += Wk.candPriorCost[a * P.nGroup + g];
            Wk.trialCost[a] = bad ? __builtin_inf() : s;
        }
    }

    // Commit candidate ai: cameras, groups and points.
    template <typename T>
    __global__ void k_accept(Problem P, Work<T> Wk, int ai, double alpha)
    {
        const int id = blockIdx.x * blockDim.x + threadIdx.x;
        if (id < P.nCam)
        {
            for (int i = 0; i < 9; ++i) P.camR[9 * id + i] = Wk.candR[(ai * P.nCam + id) * 9 + i];
            for (int i = 0; i < 3; ++i) P.camC[3 * id + i] = Wk.candC[(ai * P.nCam + id) * 3 + i];
        }
        if (id < P.nGroup * 13) P.groupParams[id] = Wk.candGroupParams[ai * P.nGroup * 13 + id];
        if (id < P.nCap)
        {
            for (int i = 0; i < 9; ++i) P.capR[9 * id + i] = Wk.candCapR[(long(ai) * P.nCap + id) * 9 + i];
            for (int i = 0; i < 3; ++i) P.capC[3 * id + i] = Wk.candCapC[(long(ai) * P.nCap + id) * 3 + i];
        }
        if (id < P.nSen)
        {
            for (int i = 0; i < 9; ++i) P.senR[9 * id + i] = Wk.candSenR[(long(ai) * P.nSen + id) * 9 + i];
            for (int i = 0; i < 3; ++i) P.senC[3 * id + i] = Wk.candSenC[(long(ai) * P.nSen + id) * 3 + i];
        }
        if (id < P.nElim)
        {
            const int t = P.elimTrack[id];
            const double s = stepScale(P, P.pointX[3 * t], P.pointX[3 * t + 1], P.pointX[3 * t + 2]);
            for (int i = 0; i < 3; ++i) P.pointX[3 * t + i] += s * alpha * Wk.dp[3 * id + i];
        }
    }

    // ---------------------------------------------------------------------------------------------
    // Camera rig. The camera-level reduced system S_cam (one pose unit per camera whose capture or sensor is
    // refined, then the intrinsic-group units) is mapped onto the rig's units: with T the chain maps (a pose
    // unit's rows go to its capture's unit through Mk and to its sensor's unit through Ms; group units map to
    // themselves), S_rig = T^T S_cam T and rhs_rig = T^T rhs_cam. Point elimination commutes with T, so this is
    // the Schur complement of the system assembled over captures and sensors (ReferenceSchur.cpp, Assembly.cpp
    // :207-284); the damped diagonal is applied to the rig system (path A damps the primary diagonal of the
    // equations it solves). The camera-level step is T x_rig.

    // Chain maps of every camera at the current rig state.
    __global__ void k_rig_maps(Problem P)
    {
        const int c = blockIdx.x * blockDim.x + threadIdx.x;
        if (c >= P.nCam) return;
        const int k = P.camCapture[c], q = P.camSensor[c];
        double Mk[36], Ms[36];
        pbdev::rigMaps(P.capR + 9 * k, P.senC + 3 * q, Mk, Ms);
        double* out = P.rigMap + 72L * c;
        for (int i = 0; i < 36; ++i)
        {
            out[i] = Mk[i];
            out[36 + i] = Ms[i];
        }
    }

    // One wave per upper block (u, v) of the rig system: lane (p, q) sums (M_l^T S_ab M_r)(p, q) over the
    // block's camera-level contributions in a fixed order, then the diagonal finalize of path A.
    template <typename T>
    __global__ void __launch_bounds__(256) k_rig_project(Problem P, Work<T> Wk, Work<T> Wr, double damping)
    {
        const int blk = blockIdx.x * (blockDim.x / kWave) + (threadIdx.x / kWave);
        if (blk >= P.nSysBlocks) return;
        const int lane = threadIdx.x & 63;
        if (lane >= 36) return;
        const int p = lane / 6, q = lane % 6;
        double acc = 0.0;
        for (int j = P.projOffset[blk]; j < P.projOffset[blk + 1]; ++j)
        {
            const T* Sb = Wk.S + long(P.projPos[j]) * 36;
            const int ml = P.projMapL[j], mr = P.projMapR[j];
            const double* ML = ml >= 0 ? P.rigMap + 36L * ml : nullptr;
            const double* MR = mr >= 0 ? P.rigMap + 36L * mr : nullptr;
            double t[6];
#pragma unroll
            for (int i = 0; i < 6; ++i)
            {
                double v = 0.0;
                if (MR)
                {
#pragma unroll
                    for (int k = 0; k < 6; ++k) v += double(Sb[i * 6 + k]) * MR[k * 6 + q];
                }
                else
                {
                    v = double(Sb[i * 6 + q]);
                }
                t[i] = v;
            }
            if (ML)
            {
#pragma unroll
                for (int i = 0; i < 6; ++i) acc += ML[i * 6 + p] * t[i];
            }
            else
            {
                acc += t[p];
            }
        }
        const int u = P.sysBlockRow[blk], v = P.sysBlockCol[blk];
        if (u == v && p == q) acc = acc == 0.0 ? 1.0 : acc * (1.0 + damping);
        Wr.S[long(P.sysPosUp[blk]) * 36 + lane] = T(acc);
        if (u != v) Wr.S[long(P.sysPosLo[blk]) * 36 + q * 6 + p] = T(acc);
    }

    // rhs_rig(u) = sum over the camera-level units a mapped to u of M^T rhs_cam(a).
    template <typename T>
    __global__ void k_rig_rhs(Problem P, Work<T> Wk, Work<T> Wr)
    {
        const int id = blockIdx.x * blockDim.x + threadIdx.x;
        if (id >= P.nSysUnits * 6) return;
        const int u = id / 6, r = id % 6;
        double acc = 0.0;
        for (int j = P.sysRhsOffset[u]; j < P.sysRhsOffset[u + 1]; ++j)
        {
            const int a = P.sysRhsUnit[j], m = P.sysRhsMap[j];
            const double* rc = Wk.rhs + 6L * a;
            if (m < 0)
            {
                acc += rc[r];
            }
            else
            {
                const double* M = P.rigMap + 36L * m;
#pragma unroll
                for (int i = 0; i < 6; ++i) acc += M[i * 6 + r] * rc[i];
            }
        }
        Wr.rhs[id] = acc;
    }

    // x_cam(a) = sum over the rig units of a of M x_rig(u).
    template <typename T>
    __global__ void k_rig_expand(Problem P, Work<T> Wr, Work<T> Wk)
    {
        const int id = blockIdx.x * blockDim.x + threadIdx.x;
        if (id >= P.nUnits * 6) return;
        const int a = id / 6, i = id % 6;
        double acc = 0.0;
        for (int e = 0; e < 2; ++e)
        {
            const int u = P.expUnit[2 * a + e];
            if (u < 0) continue;
            const int m = P.expMap[2 * a + e];
            if (m < 0)
            {
                acc += double(Wr.x[6L * u + i]);
            }
            else
            {
                const double* M = P.rigMap + 36L * m;
#pragma unroll
                for (int r = 0; r < 6; ++r) acc += M[i * 6 + r] * double(Wr.x[6L * u + r]);
            }
        }
        Wk.x[id] = T(acc);
    }

    // Step RMS over the rig units (referenceStepRms, BundleAdjustPlaMatrix.cpp:77-158: six entries per capture
    // and sensor block, the active slots of each group); fixed-order reduction.
    template <typename T>
    __global__ void k_rig_step_rms(Problem Pr, Work<T> Wr, double* scalars)
    {
        __shared__ double lds[4];
        double gs = 0.0;
        for (int u = threadIdx.x; u < Pr.nUnits; u += blockDim.x)
        {
            const int lim = Pr.unitWidth[u];
            for (int a = 0; a < lim; ++a)
            {
                const double x = double(Wr.x[u * 6 + a]);
                gs += x * x;
            }
        }
        gs = waveSum(gs);
        if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = gs;
        __syncthreads();
        if (threadIdx.x == 0)
        {
            double t = 0.0;
            for (int i = 0; i < int(blockDim.x >> 6); ++i) t += lds[i];
            scalars[kScalarRmsGlobal] = Pr.globalDim > 0 ? sqrt(t / Pr.globalDim) : 0.0;
        }
    }

    // Candidate capture and sensor poses of each step length: applyPoseDelta (world-frame / rig-frame left
    // retraction) of the refined ones (Assembly.cpp:1152-1198), the others unchanged.
    template <typename T>
    __global__ void k_rig_trial(Problem P, Work<T> Wk, const double* alphas, int nAlpha)
    {
        const int n = P.nCap + P.nSen;
        const int id = blockIdx.x * blockDim.x + threadIdx.x;
        if (id >= nAlpha * n) return;
        const int ai = id / n, j = id % n;
        const double a = alphas[ai];
        const bool capture = j < P.nCap;
        const int k = capture ? j : j - P.nCap;
        const double* R0 = capture ? P.capR + 9L * k : P.senR + 9L * k;
        const double* C0 = capture ? P.capC + 3L * k : P.senC + 3L * k;
        const int unit = capture ? P.capUnit[k] : P.senUnit[k];
        double* outR = capture ? Wk.candCapR + (long(ai) * P.nCap + k) * 9 : Wk.candSenR + (long(ai) * P.nSen + k) * 9;
        double* outC = capture ? Wk.candCapC + (long(ai) * P.nCap + k) * 3 : Wk.candSenC + (long(ai) * P.nSen + k) * 3;
        if (unit < 0)
        {
            for (int i = 0; i < 9; ++i) outR[i] = R0[i];
            for (int i = 0; i < 3; ++i) outC[i] = C0[i];
            return;
        }
        double d[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) d[i] = a * double(Wk.rigX[unit * 6 + i]);
        double R[9];
        pbdev::retractExp(R0, d[0], d[1], d[2], R);
        for (int i = 0; i < 9; ++i) outR[i] = R[i];
        for (int i = 0; i < 3; ++i) outC[i] = C0[i] + d[3 + i];
    }

    // ---------------------------------------------------------------------------------------------
    // Adjoint at convergence (pb_hip_backward). With the system built on the adjoint weight roots, H lambda =
    // d loss / d theta is solved through the same Schur path: the camera side of the right-hand side is the
    // loss gradient in the step parameters, the point side enters as z = L^T (s grad_X) (the point step is
    // X += s dp); then d loss / d y_i = M_i J_i lambda for every observation.
    template <typename T>
    __global__ void k_adjoint_rhs_cameras(Problem P, Work<T> Wk, const double* dlCenter, const double* dlRot)
    {
        const int c = blockIdx.x * blockDim.x + threadIdx.x;
        if (c >= P.nCam) return;
        const int u = P.camUnit[c];
        if (u < 0) return;
        const double g0 = dlCenter[3 * c], g1 = dlCenter[3 * c + 1], g2 = dlCenter[3 * c + 2];
        double* out = Wk.directRhs + 6 * u;
        // rotation step k: dR/dw_k at w = 0 is -D [e_k]x D R for the reference retraction (R' = D E^T D R,
        // E = Rx Ry Rz, D = diag(1, -1, -1)) and [e_k]x R for the default one (R' = Exp([w]x) R); dl/dw_k is its
        // inner product with dl/dR
        {
            const double* R = P.camR + 9 * c;
            const double* G = dlRot + 9 * c;
            const double sg[3] = {1.0, -1.0, -1.0};
#pragma unroll
            for (int kx = 0; kx < 3; ++kx)
            {
                // [e_k]x: entry (r, q) = -eps(k, r, q) ... written out: (k+1 mod 3, k+2 mod 3) = -1, (k+2, k+1) = +1
                double S[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                const int i1 = (kx + 1) % 3, i2 = (kx + 2) % 3;
                S[i1 * 3 + i2] = -1.0;
                S[i2 * 3 + i1] = 1.0;
                double acc = 0.0;
#pragma unroll
                for (int r = 0; r < 3; ++r)
#pragma unroll
                    for (int q = 0; q < 3; ++q)
                    {
                        double d = 0.0;
#pragma unroll
                        for (int m = 0; m < 3; ++m)
                        {
                            const double s = P.refParam ? -sg[r] * S[r * 3 + m] * sg[m] : S[r * 3 + m];
                            d += s * R[m * 3 + q];
                        }
                        acc += G[r * 3 + q] * d;
                    }
                out[kx] = acc;
            }
        }
        const Camera& cam = P.cams[c];
        if (cam.tangent)
        {
            // the gauge scale camera moves along its two tangent directions (basis holds dC / dt)
            out[3] = cam.basis[0] * g0 + cam.basis[1] * g1 + cam.basis[2] * g2;
            out[4] = cam.basis[3] * g0 + cam.basis[4] * g1 + cam.basis[5] * g2;
            out[5] = 0.0;
        }
        else
        {
            out[3] = g0;
            out[4] = g1;
            out[5] = g2;
        }
    }

    template <typename T>
    __global__ void k_adjoint_rhs_points(Problem P, Work<T> Wk, const double* dlPoint)
    {
        const int b = blockIdx.x * blockDim.x + threadIdx.x;
        if (b >= P.nElim) return;
        const int t = P.elimTrack[b];
        const double sc = stepScale(P, P.pointX[3 * t], P.pointX[3 * t + 1], P.pointX[3 * t + 2]);
        const double g0 = sc * dlPoint[3 * b], g1 = sc * dlPoint[3 * b + 1], g2 = sc * dlPoint[3 * b + 2];
        const T* Lp = Wk.Hinv + 9 * b;
        Wk.y[3 * b] = double(Lp[0]) * g0 + double(Lp[1]) * g1 + double(Lp[3]) * g2;
        Wk.y[3 * b + 1] = double(Lp[2]) * g1 + double(Lp[4]) * g2;
        Wk.y[3 * b + 2] = double(Lp[5]) * g2;
    }

    // Dense copy of the reduced system (column-major, leading dimension 6 * nUnits) and of its right-hand side:
    // one wave per block row, lanes over the 36 elements of each stored block.
    template <typename T>
    __global__ void k_dense_system(Problem P, Work<T> Wk, double* A, double* b)
    {
        const int u = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
        if (u >= P.nUnits) return;
        const int lane = threadIdx.x & 63;
        const long n = 6L * P.nUnits;
        for (int k = P.rowPtr[u]; k < P.rowPtr[u + 1]; ++k)
        {
            const int v = P.colIdx[k];
            if (lane < 36)
            {
                const int i = lane / 6, j = lane % 6;
                A[(6L * v + j) * n + 6 * u + i] = double(Wk.S[size_t(k) * 36 + lane]);
            }
        }
        if (lane < 6) b[6 * u + lane] = Wk.rhs[6 * u + lane];
    }

    template <typename T, int KM>
    __global__ void k_adjoint_observations(Problem P, Work<T> Wk, double* out)
    {
        const int o = blockIdx.x * blockDim.x + threadIdx.x;
        if (o >= P.nObs) return;
        const int c = P.obsCam[o];
        const int a = P.obsPoint[o];
        const int b = P.pointElim[a];
        const int g = P.camGroup[c];
        const int K = g >= 0 ? P.groupK[g] : 0;
        const int8_t* act = g >= 0 ? P.groupActive + g * kMaxActive : nullptr;
        Linearization L;
        if (!pbdev::linearize<KM>(P.cams[c], P.pointX[3 * a], P.pointX[3 * a + 1], P.pointX[3 * a + 2],
                                  P.obsUV[2 * o], P.obsUV[2 * o + 1], P.obsW[o], P.loss, P.lossScale, K, act, L,
                                  P.refParam != 0))
        {
            out[2 * o] = out[2 * o + 1] = 0.0;
            return;
        }
        double q[3];
        pbdev::adjointWeightRoot(L, P.obsW[o], P.loss, P.lossScale, q);
        double v0 = 0.0, v1 = 0.0;  // J lambda
        const int u = P.camUnit[c];
        if (u >= 0)
#pragma unroll
            for (int k = 0; k < 6; ++k)
            {
                const double xk = double(Wk.x[6 * u + k]);
                v0 += L.Jc[k] * xk;
                v1 += L.Jc[6 + k] * xk;
            }
        if (g >= 0)
        {
            const int gu = P.groupUnit[g];
#pragma unroll
            for (int k = 0; k < KM; ++k)
                if (k < K)
                {
                    const double xk = double(Wk.x[6 * (gu + k / 6) + k % 6]);
                    v0 += L.Ji[k] * xk;
                    v1 += L.Ji[kMaxActive + k] * xk;
                }
        }
        if (b >= 0)
#pragma unroll
            for (int i = 0; i < 3; ++i)
            {
                v0 += L.Jp[i] * Wk.dp[3 * b + i];
                v1 += L.Jp[3 + i] * Wk.dp[3 * b + i];
            }
        // M v = Q (Q v)
        const double t0 = q[0] * v0 + q[1] * v1, t1 = q[1] * v0 + q[2] * v1;
        out[2 * o] = q[0] * t0 + q[1] * t1;
        out[2 * o + 1] = q[1] * t0 + q[2] * t1;
    }

    // d loss / d P of control point k: its residual is r = J (X - P), so d g_X / d P = -J^T M J and
    // d loss / d P = J^T M J lambda_X with lambda_X = s dp the point part of the adjoint solution.
    template <typename T>
    __global__ void k_adjoint_constraints(Problem P, Work<T> Wk, double* out)
    {
        const int k = blockIdx.x * blockDim.x + threadIdx.x;
        if (k >= P.nCons) return;
        const int slot = P.consIndex[k];
        if (slot < 0) return;
        const int t = P.consPoint[k];
        const int b = P.pointElim[t];
        const double X0 = P.pointX[3 * t], X1 = P.pointX[3 * t + 1], X2 = P.pointX[3 * t + 2];
        double g[3] = {0.0, 0.0, 0.0};
        double r[3], cost, w;
        if (b >= 0 && constraintAt(P, k, X0, X1, X2, r, cost, w))
        {
            const int rows = P.consRows[k];
            const double* J = P.consJ + 9 * k;
            const double s = stepScale(P, X0, X1, X2);
            const double l[3] = {s * Wk.dp[3 * b], s * Wk.dp[3 * b + 1], s * Wk.dp[3 * b + 2]};
            double v[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) v[i] = i < rows ? J[3 * i] * l[0] + J[3 * i + 1] * l[1] + J[3 * i + 2] * l[2] : 0.0;
            double q[2];
            const double ss = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
            pbdev::constraintWeightRoot(ss, w, P.consDelta[k], q);
            // M v = Q (Q v), Q v = q0 v + q1 r (r . v)
#pragma unroll
            for (int pass = 0; pass < 2; ++pass)
            {
                const double rv = r[0] * v[0] + r[1] * v[1] + r[2] * v[2];
#pragma unroll
                for (int i = 0; i < 3; ++i) v[i] = q[0] * v[i] + q[1] * r[i] * rv;
            }
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
                for (int i = 0; i < 3; ++i)
                    if (i < rows) g[j] += J[3 * i + j] * v[i];
        }
#pragma unroll
        for (int j = 0; j < 3; ++j) out[3 * slot + j] = g[j];
    }

    // ---------------------------------------------------------------------------------------------
    // Launchers
    namespace
    {
        inline int blocksFor(long n, int tpb) { return int((n + tpb - 1) / tpb); }
    }

    template <typename T>
    void Launcher<T>::prepareCameras(const Problem& P, hipStream_t s)
    {
        if (P.nCam > 0) k_prepare_cameras<<<blocksFor(P.nCam, 128), 128, 0, s>>>(P);
    }

    template <typename T>
    void Launcher<T>::linearizePoints(const Problem& P, const Work<T>& W, double damping, hipStream_t s)
    {
        auto launch = [&](auto km)
        {
            constexpr int KM = decltype(km)::value;
            if (P.nLinWaves > 0)
                k_linearize_lanes<T, KM>
                    <<<blocksFor(P.nLinWaves, kLinThreads / kWave), kLinThreads, 0, s>>>(P, W, damping);
            if (P.nLinFallback > 0)
                k_linearize_points<T, KM><<<blocksFor(P.nLinFallback, 128), 128, 0, s>>>(P, W, damping);
        };
        if (P.nElim == 0) return;
        if (P.maxGroupK == 0)
            launch(std::integral_constant<int, 0>());
        else if (P.maxGroupK <= 6)
            launch(std::integral_constant<int, 6>());
        else if (P.maxGroupK <= 12)
            launch(std::integral_constant<int, 12>());
        else
            launch(std::integral_constant<int, 13>());
    }

    template <typename T>
    void Launcher<T>::cameraDirect(const Problem& P, const Work<T>& W, hipStream_t s)
    {
        if (P.nCam == 0) return;
        if (P.maxGroupK == 0)
            k_camera_direct<T, 0><<<P.nCam, 256, 0, s>>>(P, W);
        else if (P.maxGroupK <= 6)
            k_camera_direct<T, 6><<<P.nCam, 256, 0, s>>>(P, W);
        else if (P.maxGroupK <= 12)
            k_camera_direct<T, 12><<<P.nCam, 256, 0, s>>>(P, W);
        else
            k_camera_direct<T, 13><<<P.nCam, 256, 0, s>>>(P, W);
        if (P.nGroup > 0)
        {
            if (P.maxGroupK == 0)
                ;
            else if (P.maxGroupK <= 6)
                k_group_reduce<T, 6><<<P.nGroup, 512, 0, s>>>(P, W);
            else if (P.maxGroupK <= 12)
                k_group_reduce<T, 12><<<P.nGroup, 512, 0, s>>>(P, W);
            else
                k_group_reduce<T, 13><<<P.nGroup, 512, 0, s>>>(P, W);
        }
        k_finish_cost<T><<<1, 256, 0, s>>>(P, W);
    }

    template <typename T>
    void Launcher<T>::schurBlocks(const Problem& P, const Work<T>& W, double damping, hipStream_t s)
    {
        if (P.nItems > 0)
        {
            const int nb = (blocksFor(P.nItems, 4) + kXcds - 1) / kXcds * kXcds;  // a multiple of kXcds (xcdLocal)
            k_schur_partial<T><<<nb, 256, 0, s>>>(P, W);
        }
        if (P.nBlocks > 0) k_schur_finalize<T><<<blocksFor(P.nBlocks, 4), 256, 0, s>>>(P, W, damping);
        // a rig problem's preconditioner belongs to its rig system (rigProject)
        if (P.nUnits > 0 && P.nCap + P.nSen == 0) k_block_jacobi<T><<<blocksFor(P.nUnits, 64), 64, 0, s>>>(P, W);
    }

    template <typename T>
    void Launcher<T>::pcgInit(const Problem& P, const Work<T>& W, double relTol, double absTol, hipStream_t s)
    {
        if (P.cgBlocks == 0) return;
        k_pcg_init<T><<<std::max(1, blocksFor(long(P.nUnits) * 6, 256)), 256, 0, s>>>(P, W);
        k_pcg_iter<T><<<std::max(1, P.cgBlocks), 64 * kCgWaves, 0, s>>>(P, W, -1, nullptr, 0, P.nUnits, relTol, absTol,
                                                                        0.0);
    }

    template <typename T>
    void Launcher<T>::pcgIteration(const Problem& P, const Work<T>& W, int it, int maxIt, double relTol, double absTol,
                                   double modelEta, hipStream_t s, bool counted)
    {
        if (P.cgBlocks == 0) return;
        k_pcg_iter<T><<<std::max(1, P.cgBlocks), 64 * kCgWaves, 0, s>>>(
            P, W, it, counted ? W.iflags + kIflagItBase : nullptr, maxIt, P.nUnits, relTol, absTol, modelEta);
    }

    template <typename T>
    void Launcher<T>::pcgAdvance(const Problem& P, const Work<T>& W, int launches, hipStream_t s)
    {
        if (P.cgBlocks == 0) return;
        k_pcg_advance<T><<<1, 64, 0, s>>>(W, launches);
    }

    template <typename T>
    void Launcher<T>::backSubstitute(const Problem& P, const Work<T>& W, hipStream_t s)
    {
        const int nb = std::max(1, blocksFor(P.nElim, 256));
        k_back_substitute<T><<<nb, 256, 0, s>>>(P, W);
        k_step_stats<T><<<1, 256, 0, s>>>(P, W, nb);
    }

    template <typename T>
    void Launcher<T>::pointSteps(const Problem& P, const Work<T>& W, hipStream_t s)
    {
        if (P.nElim > 0) k_back_substitute<T><<<blocksFor(P.nElim, 256), 256, 0, s>>>(P, W);
    }

    template <typename T>
    void Launcher<T>::trialStates(const Problem& P, const Work<T>& W, const double* alphas, int nAlpha, hipStream_t s)
    {
        if (P.nGroup > 0) k_trial_groups<T><<<blocksFor(long(nAlpha) * P.nGroup, 64), 64, 0, s>>>(P, W, alphas, nAlpha);
        if (P.nCap + P.nSen > 0)
            k_rig_trial<T><<<blocksFor(long(nAlpha) * (P.nCap + P.nSen), 64), 64, 0, s>>>(P, W, alphas, nAlpha);
        k_trial_cameras<T><<<blocksFor(long(nAlpha) * P.nCam, 64), 64, 0, s>>>(P, W, alphas, nAlpha);
    }

    template <typename T>
    void Launcher<T>::adjointRhs(const Problem& P, const Work<T>& W, const double* dlCenter, const double* dlRot,
                                 const double* dlPoint, hipStream_t s)
    {
        if (P.nCam > 0) k_adjoint_rhs_cameras<T><<<blocksFor(P.nCam, 128), 128, 0, s>>>(P, W, dlCenter, dlRot);
        if (P.nElim > 0) k_adjoint_rhs_points<T><<<blocksFor(P.nElim, 256), 256, 0, s>>>(P, W, dlPoint);
    }

    template <typename T>
    void Launcher<T>::adjointObservations(const Problem& P, const Work<T>& W, double* out, hipStream_t s)
    {
        if (P.nObs == 0) return;
        const int nb = blocksFor(P.nObs, 128);
        if (P.maxGroupK == 0)
            k_adjoint_observations<T, 0><<<nb, 128, 0, s>>>(P, W, out);
        else if (P.maxGroupK <= 6)
            k_adjoint_observations<T, 6><<<nb, 128, 0, s>>>(P, W, out);
        else if (P.maxGroupK <= 12)
            k_adjoint_observations<T, 12><<<nb, 128, 0, s>>>(P, W, out);
        else
            k_adjoint_observations<T, 13><<<nb, 128, 0, s>>>(P, W, out);
    }

    template <typename T>
    void Launcher<T>::adjointConstraints(const Problem& P, const Work<T>& W, double* out, hipStream_t s)
    {
        if (P.nCons > 0) k_adjoint_constraints<T><<<blocksFor(P.nCons, 128), 128, 0, s>>>(P, W, out);
    }

    template <typename T>
    void Launcher<T>::denseSystem(const Problem& P, const Work<T>& W, double* A, double* b, hipStream_t s)
    {
        if (P.nUnits > 0) k_dense_system<T><<<blocksFor(P.nUnits, 4), 256, 0, s>>>(P, W, A, b);
    }

    template <typename T>
    void Launcher<T>::trialCosts(const Problem& P, const Work<T>& W, const double* alphas, int nAlpha,
                                 int costBlocks, hipStream_t s)
    {
        trialStates(P, W, alphas, nAlpha, s);
        trialObjective(P, W, alphas, nAlpha, costBlocks, s);
    }

    template <typename T>
    void Launcher<T>::trialObjective(const Problem& P, const Work<T>& W, const double* alphas, int nAlpha,
                                     int costBlocks, hipStream_t s)
    {
        k_trial_cost<T, kMaxAlpha><<<costBlocks, 256, 0, s>>>(P, W, alphas, nAlpha);
        k_trial_reduce<T, kMaxAlpha><<<1, 64 * kMaxAlpha, 0, s>>>(P, W, nAlpha, costBlocks);
    }

    template <typename T>
    __global__ void k_clear_build(Work<T> Wk, int nD, int nRhs, int unitAlpha)
    {
        const int i = blockIdx.x * blockDim.x + threadIdx.x;
        if (i < kFlagCount) Wk.flags[i] = 0;
        if (i < nD) Wk.D[i] = 0.0;
        if (i < nRhs) Wk.directRhs[i] = 0.0;
        if (i == 0 && unitAlpha) Wk.alphas[0] = 1.0;
    }

    template <typename T>
    void Launcher<T>::clearBuild(const Problem& P, const Work<T>& W, bool unitAlpha, hipStream_t s)
    {
        const int nD = std::max(1, P.nBlocks) * 36, nRhs = std::max(1, P.nUnits) * 6;
        const int n = std::max(std::max(nD, nRhs), int(kFlagCount));
        k_clear_build<T><<<blocksFor(n, 256), 256, 0, s>>>(W, nD, nRhs, unitAlpha ? 1 : 0);
    }

    template <typename T>
    __global__ void k_publish_step(Work<T> Wk, const int32_t* words, const double* norm, double* pin, int flags,
                                   int dense, int step, int trial)
    {
        const int i = threadIdx.x;
        if (i < kScalarCount)
        {
            pin[i] = Wk.scalars[i];
            pin[step + i] = Wk.scalars[i];
        }
        if (i < kFlagCount) reinterpret_cast<int32_t*>(pin + flags)[i] = Wk.flags[i];
        if (i < kDenseWordCount) reinterpret_cast<int32_t*>(pin + dense)[i] = words[i];
        if (i < 2) pin[dense + 4 + i] = norm[i];
        if (i == 0) pin[trial] = Wk.trialCost[0];
    }

    template <typename T>
    void Launcher<T>::publishStep(const Work<T>& W, const int32_t* denseWords, const double* denseNorm,
                                  double* hostPin, int flags, int dense, int step, int trial, hipStream_t s)
    {
        static_assert(kScalarCount <= 64 && kFlagCount <= 64 && kDenseWordCount <= 64, "one wave publishes");
        k_publish_step<T><<<1, 64, 0, s>>>(W, denseWords, denseNorm, hostPin, flags, dense, step, trial);
    }

    template <typename T>
    void Launcher<T>::accept(const Problem& P, const Work<T>& W, int ai, double alpha, hipStream_t s)
    {
        const long n = std::max<long>(std::max<long>(std::max<long>(P.nCam, long(P.nGroup) * 13), P.nElim),
                                      std::max(P.nCap, P.nSen));
        if (n > 0) k_accept<T><<<blocksFor(n, 256), 256, 0, s>>>(P, W, ai, alpha);
    }

    template <typename T>
    void Launcher<T>::rigMaps(const Problem& P, hipStream_t s)
    {
        if (P.nCam > 0) k_rig_maps<<<blocksFor(P.nCam, 64), 64, 0, s>>>(P);
    }

    template <typename T>
    void Launcher<T>::rigProject(const Problem& P, const Problem& Pr, const Work<T>& W, const Work<T>& Wr,
                                 double damping, hipStream_t s)
    {
        if (P.nSysBlocks > 0) k_rig_project<T><<<blocksFor(P.nSysBlocks, 4), 256, 0, s>>>(P, W, Wr, damping);
        if (P.nSysUnits > 0)
        {
            k_rig_rhs<T><<<blocksFor(long(P.nSysUnits) * 6, 128), 128, 0, s>>>(P, W, Wr);
            k_block_jacobi<T><<<blocksFor(Pr.nUnits, 64), 64, 0, s>>>(Pr, Wr);
        }
    }

    template <typename T>
    void Launcher<T>::rigExpand(const Problem& P, const Work<T>& Wr, const Work<T>& W, hipStream_t s)
    {
        if (P.nUnits > 0) k_rig_expand<T><<<blocksFor(long(P.nUnits) * 6, 128), 128, 0, s>>>(P, Wr, W);
    }

    template <typename T>
    void Launcher<T>::rigStepRms(const Problem& Pr, const Work<T>& Wr, const Work<T>& W, hipStream_t s)
    {
        k_rig_step_rms<T><<<1, 256, 0, s>>>(Pr, Wr, W.scalars);
    }

    template struct Launcher<float>;
    template struct Launcher<double>;
}  // namespace pbk
