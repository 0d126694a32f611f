// pb_common.h - device helpers shared by the kernel files (wave64).
#pragma once

#include "pb_device.h"
#include "pb_kernels.h"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace pbk
{
    constexpr int kWave = 64;

    // butterfly sum over the 64 lanes; every lane receives the total, in the same order on every lane
    template <typename V>
    __device__ __forceinline__ V waveSum(V v)
    {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
        return v;
    }

    // Orders LDS accesses of one wave (lanes exchange values through LDS without a workgroup barrier).
    __device__ __forceinline__ void waveLdsFence()
    {
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }

    __device__ __forceinline__ void setFlag(int32_t* flags, int index)
    {
        // integer flag: any writer stores the same value, the order does not matter
        atomicOr(flags + index, 1);
    }

    // Point scale of the step parameterisation (applyStep, Assembly.cpp:1213-1232).
    __device__ __forceinline__ double stepScale(const Problem& P, double X0, double X1, double X2)
    {
        return P.refParam ? pbdev::pointScale(X0, X1, X2) : 1.0;
    }

    // Survey constraint k at point X (BundleAdjustPlaMatrixConstraints.cpp:66-116): residual r = J (X - target),
    // Huber cost and IRLS weight. False when the loss is not finite; the reference then skips the constraint.
    __device__ __forceinline__ bool constraintAt(const Problem& P, int k, double X0, double X1, double X2, double r[3],
                                                 double& cost, double& w)
    {
        const int rows = P.consRows[k];
        const double* J = P.consJ + 9 * k;
        const double d0 = X0 - P.consTarget[3 * k], d1 = X1 - P.consTarget[3 * k + 1], d2 = X2 - P.consTarget[3 * k + 2];
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
            r[i] = i < rows ? J[3 * i] * d0 + J[3 * i + 1] * d1 + J[3 * i + 2] * d2 : 0.0;
            s += r[i] * r[i];
        }
        return pbdev::huber(s, P.consDelta[k], cost, w);
    }

    // Robust cost of observation o at the trial states alpha_a (a < nAlpha): candidate cameras candCams
    // [NA*nCam] and the point moved by alpha_a * s * dp (applyStep, Assembly.cpp:1010-1247). Adds the cost to
    // acc[a]; a projection behind the camera or an invalid loss argument marks bad[a].
    template <int NA>
    __device__ __forceinline__ void trialObservation(const Problem& P, const pbdev::Camera* candCams, const double* dp,
                                                     int o, const double* al, int nAlpha, double* acc, int* bad)
    {
        const int t = P.obsPoint[o];
        const int c = P.obsCam[o];
        const int b = P.pointElim[t];
        const double X0 = P.pointX[3 * t], X1 = P.pointX[3 * t + 1], X2 = P.pointX[3 * t + 2];
        double d0 = 0.0, d1 = 0.0, d2 = 0.0;
        if (b >= 0)
        {
            const double s = stepScale(P, X0, X1, X2);
            d0 = s * dp[3 * b];
            d1 = s * dp[3 * b + 1];
            d2 = s * dp[3 * b + 2];
        }
        const double ou = P.obsUV[2 * o], ov = P.obsUV[2 * o + 1], ow = P.obsW[o];
#pragma unroll
        for (int a = 0; a < NA; ++a)
        {
            if (a >= nAlpha) break;
            double u, v;
            if (!pbdev::project(candCams[a * P.nCam + c], X0 + al[a] * d0, X1 + al[a] * d1, X2 + al[a] * d2, u, v))
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

    // Cost of survey constraint k at the trial states alpha_a: the point moves by alpha_a * s * dp. A constraint
    // whose loss is not finite adds nothing (the reference skips it).
    template <int NA>
    __device__ __forceinline__ void trialConstraint(const Problem& P, const double* dp, int k, const double* al,
                                                    int nAlpha, double* acc)
    {
        const int t = P.consPoint[k];
        const int b = P.pointElim[t];
        const double X0 = P.pointX[3 * t], X1 = P.pointX[3 * t + 1], X2 = P.pointX[3 * t + 2];
        double d0 = 0.0, d1 = 0.0, d2 = 0.0;
        if (b >= 0)
        {
            const double s = stepScale(P, X0, X1, X2);
            d0 = s * dp[3 * b];
            d1 = s * dp[3 * b + 1];
            d2 = s * dp[3 * b + 2];
        }
#pragma unroll
        for (int a = 0; a < NA; ++a)
        {
            if (a >= nAlpha) break;
            double r[3], cost, w;
            if (constraintAt(P, k, X0 + al[a] * d0, X1 + al[a] * d1, X2 + al[a] * d2, r, cost, w)) acc[a] += cost;
        }
    }
}  // namespace pbk
