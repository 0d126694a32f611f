// pb_common.h - device helpers shared by the kernel files (wave64).
#pragma once

#include "pb_device.h"
#include "pb_kernels.h"

#include <hip/hip_runtime.h>

#include <cstdint>

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

    __device__ __forceinline__ void setFlag(int32_t* flags, int index)
    {
        // integer flag: any writer stores the same value, the order does not matter
        atomicOr(flags + index, 1);
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
            const double s = pbdev::pointScale(X0, X1, X2);
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
}  // namespace pbk
