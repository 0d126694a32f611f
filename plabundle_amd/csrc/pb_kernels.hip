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

#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

namespace pbk
{
    using pbdev::Camera;
    using pbdev::kMaxActive;
    using pbdev::Linearization;

    constexpr int kWave = 64;

    template <typename V>
    __device__ __forceinline__ V waveSum(V v)
    {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
        return v;
    }

    __device__ __forceinline__ void setFlag(int32_t* flags, int index)
    {
        // integer flag: any writer stores the same value, the order does not matter
        atomicOr(flags + index, 1);
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
        if (c == P.gaugeScale && anchorC)
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

    __global__ void k_prepare_cameras(Problem P)
    {
        const int c = blockIdx.x * blockDim.x + threadIdx.x;
        if (c >= P.nCam) return;
        const double* anchorC = P.gaugeScale >= 0 ? P.camC + 3 * P.gaugeAnchor : nullptr;
        const double* scaleC = P.gaugeScale >= 0 ? P.camC + 3 * P.gaugeScale : nullptr;
        makeCamera(P, c, P.camR + 9 * c, P.camC + 3 * c, P.groupParams, anchorC, scaleC, P.cams[c]);
    }

    // ---------------------------------------------------------------------------------------------
    // Point side of the linearisation: Hpp, bp, cross blocks W (6x3 per entry), damped inverse.
    template <typename T>
    __global__ void k_linearize_points(Problem P, Work<T> Wk, double damping)
    {
        const int b = blockIdx.x * blockDim.x + threadIdx.x;
        if (b >= P.nElim) return;
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
            if (!pbdev::linearize(P.cams[c], X0, X1, X2, P.obsUV[2 * o], P.obsUV[2 * o + 1], P.obsW[o], P.loss,
                                  P.lossScale, K, act, L))
            {
                setFlag(Wk.flags, kFlagInvalid);
                continue;
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
                for (int k = 0; k < K; ++k)
                {
                    T* Wg = Wk.W + (eg + k / 6) * 18 + (k % 6) * 3;
#pragma unroll
                    for (int i = 0; i < 3; ++i)
                        Wg[i] += T(w * (L.Ji[k] * L.Jp[i] + L.Ji[kMaxActive + k] * L.Jp[3 + i]));
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
        if (!pbdev::inverse3(Hd, inv))
        {
            setFlag(Wk.flags, kFlagSingular);
#pragma unroll
            for (int i = 0; i < 9; ++i) inv[i] = 0.0;
        }
#pragma unroll
        for (int i = 0; i < 9; ++i) Wk.Hinv[9 * b + i] = T(inv[i]);
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
            Wk.bp[3 * b + i] = bp[i];
            Wk.y[3 * b + i] = inv[3 * i] * bp[0] + inv[3 * i + 1] * bp[1] + inv[3 * i + 2] * bp[2];
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
        for (int k = P.camObsOffset[c] + threadIdx.x; k < P.camObsOffset[c + 1]; k += blockDim.x)
        {
            const int o = P.camObs[k];
            const int t = P.obsPoint[o];
            if (!pbdev::linearize(cam, P.pointX[3 * t], P.pointX[3 * t + 1], P.pointX[3 * t + 2], P.obsUV[2 * o],
                                  P.obsUV[2 * o + 1], P.obsW[o], P.loss, P.lossScale, KM > 0 ? K : 0, act, L))
            {
                setFlag(Wk.flags, kFlagInvalid);
                continue;
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
        // deterministic block reduction: wave butterflies, then the four wave results in order
        __shared__ double lds[4][Acc::kCount + 6 + KM + 1];
        const int lane = threadIdx.x & 63;
        const int wv = threadIdx.x >> 6;
#pragma unroll
        for (int i = 0; i < Acc::kCount; ++i)
        {
            const double v = waveSum(double(A.h[i]));
            if (lane == 0) lds[wv][i] = v;
        }
#pragma unroll
        for (int i = 0; i < 6 + KM; ++i)
        {
            const double v = waveSum(A.b[i]);
            if (lane == 0) lds[wv][Acc::kCount + i] = v;
        }
        {
            const double v = waveSum(A.cost);
            if (lane == 0) lds[wv][Acc::kCount + 6 + KM] = v;
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
                        Wk.D[P.camGroupBlock[c * 2 + j] * 36 + r * 6 + col] = k2 < K ? red[21 + r * KM + k2] : 0.0;
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
        if (K == 0) return;
        constexpr int kSym = KM * (KM + 1) / 2;
        __shared__ double red[kSym + KM];
        const double* gp = P.groupParams + g * 13;
        for (int q = threadIdx.x; q < kSym + KM; q += blockDim.x)
        {
            double v = 0.0;
            for (int k = P.groupCamOffset[g]; k < P.groupCamOffset[g + 1]; ++k)
                v += Wk.camGroupPartial[P.groupCams[k] * kGroupPartialStride + q];
            red[q] = v;
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
            Wk.D[P.groupPairBlock[g * 4 + bi * 2 + bj] * 36 + r * 6 + col] = v;
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
        const int item = blockIdx.x * (blockDim.x / kWave) + (threadIdx.x / kWave);
        if (item >= P.nItems) return;
        const int lane = threadIdx.x & 63;
        const int blk = P.itemBlock[item];
        const bool diag = P.blockRow[blk] == P.blockCol[blk];
        T acc[36];
#pragma unroll
        for (int i = 0; i < 36; ++i) acc[i] = T(0);
        double accr[6] = {0, 0, 0, 0, 0, 0};
        for (int k = P.itemBegin[item] + lane; k < P.itemBegin[item + 1]; k += kWave)
        {
            const int e1 = P.pairE1[k], e2 = P.pairE2[k];
            const int p = P.entryPoint[e1];
            T w1[18], w2[18], hi[9];
#pragma unroll
            for (int i = 0; i < 18; ++i)
            {
                w1[i] = Wk.W[e1 * 18 + i];
                w2[i] = Wk.W[e2 * 18 + i];
            }
#pragma unroll
            for (int i = 0; i < 9; ++i) hi[i] = Wk.Hinv[9 * p + i];
            // t = Hinv * W2^T (3x6)
            T tm[18];
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 6; ++j)
                    tm[i * 6 + j] = hi[i * 3] * w2[j * 3] + hi[i * 3 + 1] * w2[j * 3 + 1] + hi[i * 3 + 2] * w2[j * 3 + 2];
#pragma unroll
            for (int a = 0; a < 6; ++a)
#pragma unroll
                for (int j = 0; j < 6; ++j)
                    acc[a * 6 + j] -= w1[a * 3] * tm[j] + w1[a * 3 + 1] * tm[6 + j] + w1[a * 3 + 2] * tm[12 + j];
            if (diag)
            {
                const double y0 = Wk.y[3 * p], y1 = Wk.y[3 * p + 1], y2 = Wk.y[3 * p + 2];
#pragma unroll
                for (int a = 0; a < 6; ++a)
                    accr[a] += double(w1[a * 3]) * y0 + double(w1[a * 3 + 1]) * y1 + double(w1[a * 3 + 2]) * y2;
            }
        }
#pragma unroll
        for (int i = 0; i < 36; ++i) acc[i] = waveSum(acc[i]);
#pragma unroll
        for (int i = 0; i < 6; ++i) accr[i] = waveSum(accr[i]);
        T* out = Wk.itemPartial + item * 36;
        double* outr = Wk.itemPartialRhs + item * 6;
        // every lane holds every sum; lanes write disjoint elements
        T mine = T(0);
#pragma unroll
        for (int i = 0; i < 36; ++i)
            if (lane == i) mine = acc[i];
        if (lane < 36) out[lane] = mine;
        double mr = 0.0;
#pragma unroll
        for (int i = 0; i < 6; ++i)
            if (lane == 36 + i) mr = accr[i];
        if (lane >= 36 && lane < 42) outr[lane - 36] = mr;
    }

    // Direct part + Schur items per block, diagonal finalize (scale by 1 + lambda, exact zeros -> 1),
    // both triangles of the block-sparse matrix, and the reduced right-hand side.
    template <typename T>
    __global__ void k_schur_finalize(Problem P, Work<T> Wk, double damping)
    {
        const int blk = blockIdx.x * (blockDim.x / kWave) + (threadIdx.x / kWave);
        if (blk >= P.nBlocks) return;
        const int lane = threadIdx.x & 63;
        const int u = P.blockRow[blk], v = P.blockCol[blk];
        const bool diag = u == v;
        if (lane < 36)
        {
            double s = Wk.D[blk * 36 + lane];
            for (int it = P.blockItemOffset[blk]; it < P.blockItemOffset[blk + 1]; ++it)
                s += double(Wk.itemPartial[it * 36 + lane]);
            if (diag && (lane / 6) == (lane % 6)) s = s == 0.0 ? 1.0 : s * (1.0 + damping);
            Wk.S[P.blockPosUp[blk] * 36 + lane] = T(s);
            if (!diag) Wk.S[P.blockPosLo[blk] * 36 + (lane % 6) * 6 + lane / 6] = T(s);
        }
        else if (diag && lane < 42)
        {
            const int a = lane - 36;
            double s = 0.0;
            for (int it = P.blockItemOffset[blk]; it < P.blockItemOffset[blk + 1]; ++it)
                s += Wk.itemPartialRhs[it * 6 + a];
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

    __device__ __forceinline__ double sumPartials(const double* part, int n)
    {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += part[i];
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

    // x = 0, r = b, z = M r; partials of r.z and r.r (= ||b||^2).
    template <typename T>
    __global__ void k_pcg_start(Problem P, Work<T> Wk)
    {
        __shared__ double lds[8];
        const int u = blockIdx.x * blockDim.x + threadIdx.x;
        double rz = 0.0, rr = 0.0;
        if (u < P.nUnits)
        {
            T r[6];
#pragma unroll
            for (int i = 0; i < 6; ++i)
            {
                r[i] = T(Wk.rhs[u * 6 + i]);
                Wk.r[u * 6 + i] = r[i];
                Wk.x[u * 6 + i] = T(0);
                Wk.p[u * 6 + i] = T(0);  // p_old of iteration 0 (beta = 0)
            }
#pragma unroll
            for (int i = 0; i < 6; ++i)
            {
                T z = T(0);
#pragma unroll
                for (int j = 0; j < 6; ++j) z += Wk.Minv[u * 36 + i * 6 + j] * r[j];
                Wk.z[u * 6 + i] = z;
                rz += double(r[i]) * double(z);
                rr += double(r[i]) * double(r[i]);
            }
        }
        const double s1 = blockSum256(rz, lds);
        __syncthreads();
        const double s2 = blockSum256(rr, lds + 4);
        if (threadIdx.x == 0)
        {
            Wk.partRz[blockIdx.x] = s1;
            Wk.partRr[blockIdx.x] = s2;
        }
        if (blockIdx.x == 0 && threadIdx.x == 0)
        {
            Wk.iflags[kIflagDone] = 0;
            Wk.iflags[kIflagIters] = 0;
            Wk.iflags[kIflagBreakdown] = 0;
            Wk.iflags[kIflagModelStop] = 0;
            Wk.scalars[kScalarQ0] = 0.0;
        }
    }

    // Iteration it, first half: beta from the previous r.z, the new direction p = z + beta p_old
    // (formed on the fly for every column read), q segments of S p and partials of p.q.
    // One wave per row segment of at most kSpmvSegment blocks.
    template <typename T>
    __global__ void __launch_bounds__(256) k_pcg_spmv(Problem P, Work<T> Wk, int it, int nVecParts, double relTol,
                                                      double absTol)
    {
        if (Wk.iflags[kIflagDone] || Wk.iflags[kIflagModelStop]) return;
        __shared__ double slot[2];
        __shared__ T red[4][64];
        __shared__ double lds[8];
        const double rzCur = blockSumPartials(Wk.partRz, nVecParts, slot);
        const double rr = blockSumPartials(Wk.partRr, nVecParts, slot + 1);
        double thr;
        double beta = 0.0;
        if (it == 0)
        {
            thr = fmax(relTol * sqrt(rr), absTol);
        }
        else
        {
            thr = Wk.scalars[kScalarThreshold];
            const double rzPrev = Wk.scalars[((it - 1) & 1) ? kScalarRz1 : kScalarRz0];
            beta = rzPrev != 0.0 ? rzCur / rzPrev : 0.0;
        }
        const bool conv = sqrt(rr) <= thr;
        if (blockIdx.x == 0 && threadIdx.x == 0)
        {
            Wk.scalars[(it & 1) ? kScalarRz1 : kScalarRz0] = rzCur;
            Wk.scalars[kScalarRnorm] = sqrt(rr);
            if (it == 0)
            {
                Wk.scalars[kScalarThreshold] = thr;
                Wk.scalars[kScalarBnorm] = sqrt(rr);
            }
            Wk.iflags[kIflagIters] = it;
            if (conv) Wk.iflags[kIflagDone] = 1;
        }
        if (conv) return;  // every block takes the same decision
        const T* pOld = (it & 1) ? Wk.p2 : Wk.p;
        T* pNew = (it & 1) ? Wk.p : Wk.p2;
        const T bt = T(beta);
        const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
        const int seg = blockIdx.x * 4 + wv;
        double pq = 0.0;
        T acc = T(0);
        int u = -1;
        if (seg < P.nSegments)
        {
            u = P.segRow[seg];
            if (lane < 60)
            {
                const int row = lane % 6;
                for (int k = P.segBegin[seg] + lane / 6; k < P.segBegin[seg + 1]; k += 10)
                {
                    const int col = P.colIdx[k];
                    const T* Sb = Wk.S + k * 36 + row * 6;
                    const T* zc = Wk.z + col * 6;
                    const T* pc = pOld + col * 6;
                    T a = T(0);
#pragma unroll
                    for (int j = 0; j < 6; ++j) a += Sb[j] * (zc[j] + bt * pc[j]);
                    acc += a;
                }
            }
        }
        red[wv][lane] = acc;
        __syncthreads();
        if (seg < P.nSegments && lane < 6)
        {
            T q = T(0);
            for (int k = 0; k < 10; ++k) q += red[wv][lane + 6 * k];
            Wk.qseg[seg * 6 + lane] = q;
            const T pn = Wk.z[u * 6 + lane] + bt * pOld[u * 6 + lane];
            if (P.segFirst[seg]) pNew[u * 6 + lane] = pn;
            pq = double(pn) * double(q);
        }
        const double tot = blockSum256(pq, lds);
        if (threadIdx.x == 0) Wk.partPq[blockIdx.x] = tot;
    }

    // Iteration it, second half: alpha = r.z / p.q; q = sum of row segments; x += alpha p; r -= alpha q;
    // z = M r; partials of r.z and r.r for the next iteration.
    template <typename T>
    __global__ void k_pcg_update(Problem P, Work<T> Wk, int it, int nPq, double modelEta)
    {
        if (Wk.iflags[kIflagDone] || Wk.iflags[kIflagModelStop]) return;
        __shared__ double slot[1];
        __shared__ double lds[8];
        const double pq = blockSumPartials(Wk.partPq, nPq, slot);
        const double rz = Wk.scalars[(it & 1) ? kScalarRz1 : kScalarRz0];
        const bool breakdown = !(pq > 0.0) || !pbdev::finite(pq);
        const double alpha = breakdown ? 0.0 : rz / pq;
        const T* pNew = (it & 1) ? Wk.p : Wk.p2;
        const int u = blockIdx.x * blockDim.x + threadIdx.x;
        double rzn = 0.0, rr = 0.0;
        if (u < P.nUnits && !breakdown)
        {
            T q[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
            for (int sg = P.rowSegOffset[u]; sg < P.rowSegOffset[u + 1]; ++sg)
#pragma unroll
                for (int i = 0; i < 6; ++i) q[i] += Wk.qseg[sg * 6 + i];
            T r[6];
            const T a = T(alpha);
#pragma unroll
            for (int i = 0; i < 6; ++i)
            {
                Wk.x[u * 6 + i] += a * pNew[u * 6 + i];
                r[i] = Wk.r[u * 6 + i] - a * q[i];
                Wk.r[u * 6 + i] = r[i];
            }
#pragma unroll
            for (int i = 0; i < 6; ++i)
            {
                T z = T(0);
#pragma unroll
                for (int j = 0; j < 6; ++j) z += Wk.Minv[u * 36 + i * 6 + j] * r[j];
                Wk.z[u * 6 + i] = z;
                rzn += double(r[i]) * double(z);
                rr += double(r[i]) * double(r[i]);
            }
        }
        const double s1 = blockSum256(rzn, lds);
        __syncthreads();
        const double s2 = blockSum256(rr, lds + 4);
        if (threadIdx.x == 0)
        {
            Wk.partRz[blockIdx.x] = s1;
            Wk.partRr[blockIdx.x] = s2;
        }
        if (blockIdx.x == 0 && threadIdx.x == 0)
        {
            if (breakdown)
            {
                Wk.iflags[kIflagBreakdown] = 1;
                Wk.iflags[kIflagDone] = 1;
                Wk.iflags[kIflagIters] = it;
            }
            else
            {
                // quadratic model Q(x) = x'Sx/2 - b'x decreases by alpha r.z/2 per step; truncated-Newton test
                // i (Q_i - Q_{i-1}) / Q_i < eta
                const double dq = -0.5 * alpha * rz;
                const double qNew = Wk.scalars[(it & 1) ? kScalarQ1 : kScalarQ0] + dq;
                Wk.scalars[((it + 1) & 1) ? kScalarQ1 : kScalarQ0] = qNew;
                // Done is not raised here: sibling blocks of this launch may still be reading it. The next
                // SpMV launch and the host treat ModelStop as done.
                if (modelEta > 0.0 && qNew < 0.0 && double(it + 1) * dq / qNew < modelEta)
                {
                    Wk.iflags[kIflagModelStop] = 1;
                    Wk.iflags[kIflagIters] = it + 1;
                }
            }
        }
    }

    // ---------------------------------------------------------------------------------------------
    // Persistent preconditioned conjugate gradients (Chronopoulos-Gear form): one SpMV and one pair of
    // dot products per iteration, two grid barriers per iteration, all iterations in one cooperative
    // launch. Every block derives the scalars from the same partials in the same order, so all blocks
    // take identical decisions and execute the same number of barriers; the loop is bounded by maxIter.
    template <typename T>
    __device__ __forceinline__ void applyM(const Work<T>& Wk, int u, const T* r, T* out)
    {
        const T* M = Wk.Minv + u * 36;
#pragma unroll
        for (int i = 0; i < 6; ++i)
        {
            T z = T(0);
#pragma unroll
            for (int j = 0; j < 6; ++j) z += M[i * 6 + j] * r[j];
            out[i] = z;
        }
    }

    // w = S u for the rows of this block (wave per row, 60 lanes: row element = lane % 6, block slot = lane / 6);
    // returns this thread's share of (w.u).
    template <typename T>
    __device__ __forceinline__ double spmvRows(const Problem& P, const Work<T>& Wk, int ub, int ue, T (*red)[64])
    {
        const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
        double wu = 0.0;
        for (int u0 = ub; u0 < ue; u0 += 4)
        {
            const int u = u0 + wv;
            T acc = T(0);
            if (u < ue && lane < 60)
            {
                const int row = lane % 6;
                for (int k = P.rowPtr[u] + lane / 6; k < P.rowPtr[u + 1]; k += 10)
                {
                    const T* Sb = Wk.S + k * 36 + row * 6;
                    const T* pv = Wk.u + P.colIdx[k] * 6;
                    acc += Sb[0] * pv[0] + Sb[1] * pv[1] + Sb[2] * pv[2] + Sb[3] * pv[3] + Sb[4] * pv[4] + Sb[5] * pv[5];
                }
            }
            red[wv][lane] = acc;
            __syncthreads();
            if (u < ue && lane < 6)
            {
                T sum = T(0);
                for (int q = 0; q < 10; ++q) sum += red[wv][lane + 6 * q];
                Wk.w[u * 6 + lane] = sum;
                wu += double(sum) * double(Wk.u[u * 6 + lane]);
            }
            __syncthreads();
        }
        return wu;
    }

    template <typename T>
    __global__ void __launch_bounds__(256) k_pcg_persistent(Problem P, Work<T> Wk, int maxIter, double relTol,
                                                            double absTol)
    {
        namespace cg = cooperative_groups;
        cg::grid_group grid = cg::this_grid();
        __shared__ T red[4][64];
        __shared__ double lds[3][4];
        __shared__ double bc[4];
        const int nb = gridDim.x;
        const int ub = P.pcgRowStart[blockIdx.x], ue = P.pcgRowStart[blockIdx.x + 1];
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        // block partial of three values -> Wk.partPq[3 * block + k]
        auto publish = [&](double a, double b, double c) {
            a = waveSum(a);
            b = waveSum(b);
            c = waveSum(c);
            if (lane == 0)
            {
                lds[0][wv] = a;
                lds[1][wv] = b;
                lds[2][wv] = c;
            }
            __syncthreads();
            if (threadIdx.x < 3)
            {
                const int k = threadIdx.x;
                Wk.partPq[3 * blockIdx.x + k] = lds[k][0] + lds[k][1] + lds[k][2] + lds[k][3];
            }
            __syncthreads();
        };
        // every block sums the partials with the same fixed-order wave reduction, so all blocks hold
        // bitwise identical scalars
        auto gather = [&](double* out3) {
            if (wv == 0)
            {
                double a = 0.0, b = 0.0, c = 0.0;
                for (int i = lane; i < nb; i += 64)
                {
                    a += Wk.partPq[3 * i];
                    b += Wk.partPq[3 * i + 1];
                    c += Wk.partPq[3 * i + 2];
                }
                a = waveSum(a);
                b = waveSum(b);
                c = waveSum(c);
                if (lane == 0)
                {
                    bc[0] = a;
                    bc[1] = b;
                    bc[2] = c;
                }
            }
            __syncthreads();
            out3[0] = bc[0];
            out3[1] = bc[1];
            out3[2] = bc[2];
            __syncthreads();
        };

        // x = 0, r = b, u = M r
        double ru = 0.0, rr = 0.0;
        for (int i = threadIdx.x; i < (ue - ub); i += blockDim.x)
        {
            const int u = ub + i;
            T r[6], z[6];
#pragma unroll
            for (int k = 0; k < 6; ++k)
            {
                r[k] = T(Wk.rhs[u * 6 + k]);
                Wk.r[u * 6 + k] = r[k];
                Wk.x[u * 6 + k] = T(0);
            }
            applyM(Wk, u, r, z);
#pragma unroll
            for (int k = 0; k < 6; ++k) Wk.u[u * 6 + k] = z[k];
        }
        grid.sync();
        double wu = spmvRows(P, Wk, ub, ue, red);
        for (int i = threadIdx.x; i < (ue - ub); i += blockDim.x)
        {
            const int u = ub + i;
#pragma unroll
            for (int k = 0; k < 6; ++k)
            {
                const double rv = double(Wk.r[u * 6 + k]);
                ru += rv * double(Wk.u[u * 6 + k]);
                rr += rv * rv;
            }
        }
        publish(ru, wu, rr);
        grid.sync();
        double sc[3];
        gather(sc);
        double gamma = sc[0], delta = sc[1];
        const double bnorm = sqrt(sc[2]);
        const double thr = fmax(relTol * bnorm, absTol);
        double rnorm = bnorm;
        int iters = 0;
        bool done = rnorm <= thr;
        bool breakdown = false;
        double eta = delta;
        double alpha = 0.0;
        if (!done)
        {
            breakdown = !(eta > 0.0) || !pbdev::finite(eta);
            alpha = breakdown ? 0.0 : gamma / eta;
        }
        // p = u, s = w
        for (int i = threadIdx.x; i < (ue - ub) * 6; i += blockDim.x)
        {
            const int e = ub * 6 + i;
            Wk.p[e] = Wk.u[e];
            Wk.q[e] = Wk.w[e];
        }
        while (!done && !breakdown && iters < maxIter)
        {
            // x += alpha p; r -= alpha s; u = M r (rows of this block)
            const T a = T(alpha);
            for (int i = threadIdx.x; i < (ue - ub); i += blockDim.x)
            {
                const int u = ub + i;
                T r[6], z[6];
#pragma unroll
                for (int k = 0; k < 6; ++k)
                {
                    Wk.x[u * 6 + k] += a * Wk.p[u * 6 + k];
                    r[k] = Wk.r[u * 6 + k] - a * Wk.q[u * 6 + k];
                    Wk.r[u * 6 + k] = r[k];
                }
                applyM(Wk, u, r, z);
#pragma unroll
                for (int k = 0; k < 6; ++k) Wk.u[u * 6 + k] = z[k];
            }
            grid.sync();
            wu = spmvRows(P, Wk, ub, ue, red);
            ru = 0.0;
            rr = 0.0;
            for (int i = threadIdx.x; i < (ue - ub); i += blockDim.x)
            {
                const int u = ub + i;
#pragma unroll
                for (int k = 0; k < 6; ++k)
                {
                    const double rv = double(Wk.r[u * 6 + k]);
                    ru += rv * double(Wk.u[u * 6 + k]);
                    rr += rv * rv;
                }
            }
            publish(ru, wu, rr);
            grid.sync();
            gather(sc);
            ++iters;
            rnorm = sqrt(sc[2]);
            if (rnorm <= thr)
            {
                done = true;
                break;
            }
            const double gnew = sc[0];
            delta = sc[1];
            const double beta = gnew / gamma;
            eta = delta - beta * gnew / alpha;
            if (!(eta > 0.0) || !pbdev::finite(eta) || !pbdev::finite(beta))
            {
                breakdown = true;
                break;
            }
            alpha = gnew / eta;
            gamma = gnew;
            // p = u + beta p; s = w + beta s
            const T b = T(beta);
            for (int i = threadIdx.x; i < (ue - ub) * 6; i += blockDim.x)
            {
                const int e = ub * 6 + i;
                Wk.p[e] = Wk.u[e] + b * Wk.p[e];
                Wk.q[e] = Wk.w[e] + b * Wk.q[e];
            }
        }
        if (blockIdx.x == 0 && threadIdx.x == 0)
        {
            Wk.iflags[kIflagDone] = done ? 1 : 0;
            Wk.iflags[kIflagIters] = iters;
            Wk.iflags[kIflagBreakdown] = breakdown ? 1 : 0;
            Wk.scalars[kScalarRnorm] = rnorm;
            Wk.scalars[kScalarThreshold] = thr;
            Wk.scalars[kScalarBnorm] = bnorm;
        }
    }

    // ---------------------------------------------------------------------------------------------
    // Back-substitution dp = Hinv (bp - sum_e W_e^T dc_e); partials of bp.dp and |dp|^2.
    template <typename T>
    __global__ void k_back_substitute(Problem P, Work<T> Wk)
    {
        __shared__ double lds[8];
        const int b = blockIdx.x * blockDim.x + threadIdx.x;
        double dd = 0.0, ss = 0.0;
        if (b < P.nElim)
        {
            double z[3] = {Wk.bp[3 * b], Wk.bp[3 * b + 1], Wk.bp[3 * b + 2]};
            for (int e = P.entryOffset[b]; e < P.entryOffset[b + 1]; ++e)
            {
                const int u = P.entryUnit[e];
                const T* Wp = Wk.W + e * 18;
#pragma unroll
                for (int a = 0; a < 6; ++a)
                {
                    const double xa = double(Wk.x[u * 6 + a]);
#pragma unroll
                    for (int i = 0; i < 3; ++i) z[i] -= double(Wp[a * 3 + i]) * xa;
                }
            }
#pragma unroll
            for (int i = 0; i < 3; ++i)
            {
                const double d = double(Wk.Hinv[9 * b + 3 * i]) * z[0] + double(Wk.Hinv[9 * b + 3 * i + 1]) * z[1] +
                                 double(Wk.Hinv[9 * b + 3 * i + 2]) * z[2];
                Wk.dp[3 * b + i] = d;
                dd += Wk.bp[3 * b + i] * d;
                ss += d * d;
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
        const double s1 = blockSum256(dd, lds);
        __syncthreads();
        const double s2 = blockSum256(gs, lds + 4);
        if (threadIdx.x == 0)
        {
            const double ddp = sumPartials(Wk.partDd, nParts);
            const double ssp = sumPartials(Wk.partSs, nParts);
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
        const double a = alphas[ai];
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
        const double a = alphas[ai];
        const int u = P.camUnit[c];
        double R[9], Cc[3];
        const double* R0 = P.camR + 9 * c;
        const double* C0 = P.camC + 3 * c;
        double anchorC[3], scaleC[3];
        const bool gauge = P.gaugeScale >= 0;
        if (gauge)
        {
            for (int i = 0; i < 3; ++i)
            {
                anchorC[i] = P.camC[3 * P.gaugeAnchor + i];
                scaleC[i] = P.camC[3 * P.gaugeScale + i];
            }
        }
        if (u >= 0)
        {
            double d[6];
#pragma unroll
            for (int i = 0; i < 6; ++i) d[i] = a * double(Wk.x[u * 6 + i]);
            pbdev::retractRotation(R0, d[0], d[1], d[2], R);
            if (c == P.gaugeScale)
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
                    const double s = P.gaugeBaseline / len;
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
        for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < P.nObs; o += gridDim.x * blockDim.x)
        {
            const int t = P.obsPoint[o];
            const int c = P.obsCam[o];
            const int b = P.pointElim[t];
            const double X0 = P.pointX[3 * t], X1 = P.pointX[3 * t + 1], X2 = P.pointX[3 * t + 2];
            double d0 = 0.0, d1 = 0.0, d2 = 0.0;
            if (b >= 0)
            {
                const double s = pbdev::pointScale(X0, X1, X2);
                d0 = s * Wk.dp[3 * b];
                d1 = s * Wk.dp[3 * b + 1];
                d2 = s * Wk.dp[3 * b + 2];
            }
            const double ou = P.obsUV[2 * o], ov = P.obsUV[2 * o + 1], ow = P.obsW[o];
#pragma unroll
            for (int a = 0; a < NA; ++a)
            {
                if (a >= nAlpha) break;
                const double al = alphas[a];
                double u, v;
                if (!pbdev::project(Wk.candCams[a * P.nCam + c], X0 + al * d0, X1 + al * d1, X2 + al * d2, u, v))
                {
                    bad[a] = 1;
                    continue;
                }
                const double ru = u - ou, rv = v - ov;
                double cost, w;
                if (!pbdev::robust(ow * (ru * ru + rv * rv), P.loss, P.lossScale, cost, w))
                {
                    bad[a] = 1;
                    continue;
                }
                acc[a] += cost;
            }
        }
        __shared__ double lds[NA][4];
        __shared__ int lbad[NA][4];
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
        for (int a = 0; a < NA; ++a)
        {
            const double s = waveSum(acc[a]);
            const int anyBad = __any(bad[a]) ? 1 : 0;
            if (lane == 0)
            {
                lds[a][wv] = s;
                lbad[a][wv] = anyBad;
            }
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
    __global__ void k_trial_reduce(Problem P, Work<T> Wk, int nAlpha, int nParts)
    {
        const int a = threadIdx.x;
        if (a >= nAlpha) return;
        double s = 0.0;
        int bb = 0;
        for (int i = 0; i < nParts; ++i)
        {
            s += Wk.trialPart[i * NA + a];
            bb |= Wk.trialBad[i * NA + a];
        }
        for (int g = 0; g < P.nGroup; ++g) s += Wk.candPriorCost[a * P.nGroup + g];
        Wk.trialCost[a] = bb ? __builtin_inf() : s;
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
        if (id < P.nElim)
        {
            const int t = P.elimTrack[id];
            const double s = pbdev::pointScale(P.pointX[3 * t], P.pointX[3 * t + 1], P.pointX[3 * t + 2]);
            for (int i = 0; i < 3; ++i) P.pointX[3 * t + i] += s * alpha * Wk.dp[3 * id + i];
        }
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
        if (P.nElim > 0) k_linearize_points<T><<<blocksFor(P.nElim, 128), 128, 0, s>>>(P, W, damping);
    }

    template <typename T>
    void Launcher<T>::cameraDirect(const Problem& P, const Work<T>& W, hipStream_t s)
    {
        if (P.nCam == 0) return;
        if (P.maxGroupK == 0)
            k_camera_direct<T, 0><<<P.nCam, 256, 0, s>>>(P, W);
        else if (P.maxGroupK <= 6)
            k_camera_direct<T, 6><<<P.nCam, 256, 0, s>>>(P, W);
        else
            k_camera_direct<T, 12><<<P.nCam, 256, 0, s>>>(P, W);
        if (P.nGroup > 0)
        {
            if (P.maxGroupK == 0)
                ;
            else if (P.maxGroupK <= 6)
                k_group_reduce<T, 6><<<P.nGroup, 128, 0, s>>>(P, W);
            else
                k_group_reduce<T, 12><<<P.nGroup, 128, 0, s>>>(P, W);
        }
        k_finish_cost<T><<<1, 256, 0, s>>>(P, W);
    }

    template <typename T>
    void Launcher<T>::schurBlocks(const Problem& P, const Work<T>& W, double damping, hipStream_t s)
    {
        if (P.nItems > 0) k_schur_partial<T><<<blocksFor(P.nItems, 4), 256, 0, s>>>(P, W);
        if (P.nBlocks > 0) k_schur_finalize<T><<<blocksFor(P.nBlocks, 4), 256, 0, s>>>(P, W, damping);
        if (P.nUnits > 0) k_block_jacobi<T><<<blocksFor(P.nUnits, 64), 64, 0, s>>>(P, W);
    }

    template <typename T>
    int Launcher<T>::pcgStart(const Problem& P, const Work<T>& W, double relTol, double absTol, hipStream_t s)
    {
        const int nb = std::max(1, blocksFor(P.nUnits, 256));
        k_pcg_start<T><<<nb, 256, 0, s>>>(P, W);
        return nb;
    }

    template <typename T>
    void Launcher<T>::pcgSpmv(const Problem& P, const Work<T>& W, int it, double relTol, double absTol, hipStream_t s)
    {
        const int nbVec = std::max(1, blocksFor(P.nUnits, 256));
        k_pcg_spmv<T><<<std::max(1, blocksFor(P.nSegments, 4)), 256, 0, s>>>(P, W, it, nbVec, relTol, absTol);
    }

    template <typename T>
    void Launcher<T>::pcgUpdate(const Problem& P, const Work<T>& W, int it, double modelEta, hipStream_t s)
    {
        const int nbSpmv = std::max(1, blocksFor(P.nSegments, 4));
        const int nbVec = std::max(1, blocksFor(P.nUnits, 256));
        k_pcg_update<T><<<nbVec, 256, 0, s>>>(P, W, it, nbSpmv, modelEta);
    }

    template <typename T>
    int Launcher<T>::pcgPersistentBlocks()
    {
        int perCu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, k_pcg_persistent<T>, 256, 0) != hipSuccess) return 0;
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess) return 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
        return perCu * cus;
    }

    template <typename T>
    hipError_t Launcher<T>::pcgPersistent(const Problem& P, const Work<T>& W, int blocks, int maxIter, double relTol,
                                          double absTol, hipStream_t s)
    {
        Problem Pc = P;
        Work<T> Wc = W;
        void* args[] = {&Pc, &Wc, &maxIter, &relTol, &absTol};
        return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_pcg_persistent<T>), dim3(blocks), dim3(256),
                                          args, 0, s);
    }

    template <typename T>
    void Launcher<T>::backSubstitute(const Problem& P, const Work<T>& W, hipStream_t s)
    {
        const int nb = std::max(1, blocksFor(P.nElim, 256));
        k_back_substitute<T><<<nb, 256, 0, s>>>(P, W);
        k_step_stats<T><<<1, 256, 0, s>>>(P, W, nb);
    }

    template <typename T>
    void Launcher<T>::trialCosts(const Problem& P, const Work<T>& W, const double* alphas, int nAlpha,
                                 int costBlocks, hipStream_t s)
    {
        if (P.nGroup > 0) k_trial_groups<T><<<blocksFor(long(nAlpha) * P.nGroup, 64), 64, 0, s>>>(P, W, alphas, nAlpha);
        k_trial_cameras<T><<<blocksFor(long(nAlpha) * P.nCam, 64), 64, 0, s>>>(P, W, alphas, nAlpha);
        k_trial_cost<T, kMaxAlpha><<<costBlocks, 256, 0, s>>>(P, W, alphas, nAlpha);
        k_trial_reduce<T, kMaxAlpha><<<1, 64, 0, s>>>(P, W, nAlpha, costBlocks);
    }

    template <typename T>
    void Launcher<T>::accept(const Problem& P, const Work<T>& W, int ai, double alpha, hipStream_t s)
    {
        const long n = std::max<long>(std::max<long>(P.nCam, long(P.nGroup) * 13), P.nElim);
        if (n > 0) k_accept<T><<<blocksFor(n, 256), 256, 0, s>>>(P, W, ai, alpha);
    }

    template struct Launcher<float>;
    template struct Launcher<double>;
}  // namespace pbk
