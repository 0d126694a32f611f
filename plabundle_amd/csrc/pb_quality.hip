// pb_quality.hip - post-solve quality on the device: strict reprojection RMS of every active track at the input
// and at the refined state, the adaptive point filter and the track statistics of finalizeBundleAdjustResult
// (src/internal/BundleAdjustQuality.cpp:302-502). The median of the filter candidates comes from a device radix
// sort (rocPRIM) of their RMS values.
#include "pb_common.h"
#include "pb_device.h"
#include "pb_kernels.h"

#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include <stdexcept>
#include <string>

namespace pbk
{
    namespace
    {
        constexpr int kSummaryThreads = 1024;

        // strictTrackRms (BundleAdjustQuality.cpp:302-363) over the usable observations of active track a: a
        // failed projection, a flagged weighted observation, fewer than four residuals or a single distinct
        // camera give inf.
        __device__ double strictRms(const Problem& P, const pbdev::Camera* cams, int a, double X0, double X1, double X2)
        {
            const double inf = __builtin_inf();
            if (!pbdev::finite(X0) || !pbdev::finite(X1) || !pbdev::finite(X2)) return inf;
            if (P.trackStrictBad[a]) return inf;
            double sum = 0.0;
            int residuals = 0, first = -1;
            bool second = false;
            for (int o = P.obsOffset[a]; o < P.obsOffset[a + 1]; ++o)
            {
                const int c = P.obsCam[o];
                double u, v;
                if (!pbdev::project(cams[c], X0, X1, X2, u, v)) return inf;
                const double du = u - P.obsUV[2 * o], dv = v - P.obsUV[2 * o + 1];
                sum += P.obsW[o] * (du * du + dv * dv);
                residuals += 2;
                if (first < 0)
                    first = c;
                else if (c != first)
                    second = true;
            }
            if (residuals < 4 || !second) return inf;
            return sqrt(sum / residuals);
        }
    }  // namespace

    __global__ void k_quality_tracks(Problem P, Quality Q)
    {
        const int a = blockIdx.x * blockDim.x + threadIdx.x;
        if (a >= P.nActive) return;
        const double before =
            strictRms(P, P.inputCams, a, P.pointX0[3 * a], P.pointX0[3 * a + 1], P.pointX0[3 * a + 2]);
        const double after = strictRms(P, P.cams, a, P.pointX[3 * a], P.pointX[3 * a + 1], P.pointX[3 * a + 2]);
        Q.before[a] = before;
        Q.after[a] = after;
        Q.key[a] = pbdev::finite(after) ? after : __builtin_inf();
    }

    // One workgroup: candidate count, median of the sorted candidates (even count: mean of the middle two),
    // threshold max(floor, factor * median) (adaptivePointFilterThreshold, BundleAdjustQuality.cpp:373-392), then
    // the kept tracks and their RMS sums, all reductions in a fixed order.
    __global__ void __launch_bounds__(kSummaryThreads) k_quality_summary(Problem P, Quality Q, int filter,
                                                                          double floorPx, double factor)
    {
        __shared__ double red[3][kSummaryThreads / kWave];
        __shared__ double thr;
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        constexpr int kWaves = kSummaryThreads / kWave;
        double m = 0.0;
        for (int a = threadIdx.x; a < P.nActive; a += blockDim.x) m += pbdev::finite(Q.key[a]) ? 1.0 : 0.0;
        m = waveSum(m);
        if (lane == 0) red[0][wv] = m;
        __syncthreads();
        if (threadIdx.x == 0)
        {
            double cnt = 0.0;
            for (int q = 0; q < kWaves; ++q) cnt += red[0][q];
            const long k = long(cnt);
            const double median = k == 0 ? 0.0 : (k % 2 ? Q.sorted[k / 2] : 0.5 * (Q.sorted[k / 2 - 1] + Q.sorted[k / 2]));
            double t = pbdev::finite(floorPx) && floorPx >= 0.0 ? floorPx : 0.0;
            if (k > 0 && pbdev::finite(factor) && factor > 0.0) t = fmax(t, factor * median);
            thr = filter ? t : __builtin_inf();
            Q.out[0] = cnt;
            Q.out[1] = median;
            Q.out[2] = thr;
        }
        __syncthreads();
        const double threshold = thr;
        double kept = 0.0, sb = 0.0, sa = 0.0;
        for (int a = threadIdx.x; a < P.nActive; a += blockDim.x)
        {
            const double after = Q.after[a];
            const bool ok = pbdev::finite(after) && !(after > threshold);
            Q.valid[a] = ok ? 1 : 0;
            if (ok)
            {
                kept += 1.0;
                sb += Q.before[a];
                sa += after;
            }
        }
        kept = waveSum(kept);
        sb = waveSum(sb);
        sa = waveSum(sa);
        __syncthreads();
        if (lane == 0)
        {
            red[0][wv] = kept;
            red[1][wv] = sb;
            red[2][wv] = sa;
        }
        __syncthreads();
        if (threadIdx.x == 0)
        {
            double k = 0.0, b = 0.0, c = 0.0;
            for (int q = 0; q < kWaves; ++q)
            {
                k += red[0][q];
                b += red[1][q];
                c += red[2][q];
            }
            Q.out[3] = k;
            Q.out[4] = b;
            Q.out[5] = c;
        }
    }

    size_t QualityLauncher::sortScratchBytes(int n)
    {
        size_t bytes = 0;
        (void)hipGetLastError();
        const hipError_t e = rocprim::radix_sort_keys(nullptr, bytes, static_cast<const double*>(nullptr),
                                                      static_cast<double*>(nullptr), size_t(n));
        if (e != hipSuccess)
            throw std::runtime_error(std::string("radix sort scratch query failed: ") + hipGetErrorString(e));
        return bytes;
    }

    void QualityLauncher::run(const Problem& P, const Quality& Q, int filter, double floorPx, double factor,
                              hipStream_t s)
    {
        if (P.nActive > 0)
        {
            // rocPRIM checks hipGetLastError after its launches: clear what an earlier, already reported call
            // left there, then make this launch's own failure visible
            (void)hipGetLastError();
            k_quality_tracks<<<(P.nActive + 255) / 256, 256, 0, s>>>(P, Q);
            size_t bytes = Q.tempBytes;
            const hipError_t e = rocprim::radix_sort_keys(Q.temp, bytes, static_cast<const double*>(Q.key), Q.sorted,
                                                          size_t(P.nActive), 0, 64, s);
            if (e != hipSuccess)
                throw std::runtime_error(std::string("radix sort of the track RMS values failed: ") +
                                         hipGetErrorString(e));
        }
        k_quality_summary<<<1, kSummaryThreads, 0, s>>>(P, Q, filter, floorPx, factor);
    }
}  // namespace pbk
