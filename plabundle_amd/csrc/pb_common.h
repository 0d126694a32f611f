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

    // Wave sums of N values held by every lane, by recursive halving: at the stage with partner lane l ^ O
    // (O = 32, 16, ..., 1) each lane keeps one half of its remaining values (the upper half on lanes with bit O
    // set; zero-padded to ceil(n / 2)) and adds the partner's copy of that half. After six stages lane l holds
    // halvingCount<N>() values, value j being the sum of value halvingSlot<N>(l, j) (padding where that is -1).
    // About N exchanges instead of the 6 N of N butterflies; the order of the additions is fixed, so the sums
    // are reproducible.
    template <int N>
    constexpr int halvingCount()
    {
        int n = N;
        for (int s = 0; s < 6; ++s) n = (n + 1) / 2;
        return n;
    }

    // (the values are exchanged in their own type VI and added in VO)
    template <int N, int O, typename VO, typename VI>
    __device__ __forceinline__ void halvingSum(const VI* v, int lane, VO* out)
    {
        constexpr int H = (N + 1) / 2;
        VO a[H];
        const bool up = (lane & O) != 0;
#pragma unroll
        for (int k = 0; k < H; ++k)
        {
            const VI lo = v[k], hi = k + H < N ? v[k + H] : VI(0);
            a[k] = VO(up ? hi : lo) + VO(__shfl_xor(up ? lo : hi, O, kWave));
        }
        if constexpr (O == 1)
        {
#pragma unroll
            for (int k = 0; k < H; ++k) out[k] = a[k];
        }
        else
        {
            halvingSum<H, O / 2, VO, VO>(a, lane, out);
        }
    }

    template <int N>
    __device__ __forceinline__ int halvingSlot(int lane, int j)
    {
        int ns[6], offs[6];
        int n = N;
#pragma unroll
        for (int s = 0; s < 6; ++s)
        {
            const int h = (n + 1) / 2;
            ns[s] = n;
            offs[s] = (lane & (32 >> s)) ? h : 0;
            n = h;
        }
        // value j after the last stage is value j + sum(offs[s..5]) of stage s, which exists when below ns[s]
        int slot = j;
        bool held = true;
#pragma unroll
        for (int s = 5; s >= 0; --s)
        {
            slot += offs[s];
            held = held && slot < ns[s];
        }
        return held ? slot : -1;
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
