// pb_batch.hip - batched windows (pb_hip_solve_batch): per-window LM control and the dense window solve.
//
// The windows of a batch form one block-diagonal problem. Linearisation, Schur reduction and point
// back-substitution run through the shared grid kernels (every launch covers every window); the kernels here
// keep one LM controller per window on the device (BundleAdjustPlaMatrix.cpp:162-631 with
// useReferenceOnlineSchur, run per window) and solve each window's reduced camera system exactly with a dense
// FP64 Cholesky factorisation, one workgroup per window.
//
// Round of every running window:
//   k_batch_round_begin   activity, flags and line-search state of the round
//   (shared build)        effective cameras, linearisation, camera terms, Schur blocks with the window damping
//   k_batch_after_build   objective of the linearisation point, build failures -> reject or stop
//   k_batch_tiles         dense S of the window, MFMA tile Cholesky, tile forward/backward solves -> camera step
//                         (k_batch_factor: the left-looking panel factorisation with autotuned launch shapes)
//   (shared back-subst.)  point steps
//   k_batch_step_stats    directional derivative and step RMS values per window
//   trial passes          k_batch_alphas, trial states (shared), k_batch_trial_cost, k_batch_decide
//   k_batch_control       accept / reject, damping update, convergence and budget tests, precision phase
//   k_batch_apply         commits the accepted trial state of each window
// A mixed batch runs each round twice: the kernels instantiated for FP32 assembly on the windows in the FP32
// phase, then the FP64 kernels on the windows in the FP64 phase (every window ends in FP64).
#include "pb_common.h"
#include "pb_device.h"
#include "pb_kernels.h"
#include "pb_tile.h"

#include <hip/hip_runtime.h>

#include <stdexcept>

namespace pbk
{
    namespace
    {
        int gridFor(long n, int per) { return static_cast<int>(std::max<long>(1, (n + per - 1) / per)); }

        __device__ __forceinline__ unsigned long long deviceClock() { return wall_clock64(); }
    }  // namespace

    __global__ void k_batch_start(Batch B)
    {
        if (threadIdx.x == 0 && blockIdx.x == 0) B.doneClock[B.nWin] = deviceClock();
    }

    __global__ void k_batch_round_begin(Problem P, Batch B, int phase)
    {
        const int w = blockIdx.x * blockDim.x + threadIdx.x;
        if (w >= B.nWin) return;
        B.active[w] = (B.status[w] == kWinRunning && B.phase[w] == phase) ? 1 : 0;
        B.reject[w] = 0;
        B.searching[w] = 0;
        B.nAlpha[w] = 0;
        B.acceptIdx[w] = -1;
#pragma unroll
        for (int f = 0; f < kFlagCount; ++f) P.winFlags[w * kFlagCount + f] = 0;
    }

    // Objective at the linearisation point (camera costs in camera order plus the group priors) and the
    // reaction to build failures, as the controller does after each build.
    template <typename T>
    __global__ void k_batch_after_build(Problem P, Work<T> Wk, Batch B)
    {
        const int lane = threadIdx.x & 63;
        const int w = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
        if (w >= B.nWin || !B.active[w]) return;
        double v = 0.0;
        for (int c = B.camOffset[w] + lane; c < B.camOffset[w + 1]; c += kWave) v += Wk.camCost[c];
        v = waveSum(v);
        if (lane != 0) return;
        for (int g = B.groupOffset[w]; g < B.groupOffset[w + 1]; ++g) v += P.groupK[g] > 0 ? Wk.groupPriorCost[g] : 0.0;
        B.cost[w] = v;
        int32_t* cnt = B.counters + w * kWinCounterCount;
        ++cnt[kWinLinearizations];
        ++cnt[kWinEvaluations];
        if (cnt[kWinLinearizations] == 1) B.initialCost[w] = v;
        const int32_t* f = P.winFlags + w * kFlagCount;
        const bool empty = B.poseUnitOffset[w + 1] == B.poseUnitOffset[w] &&
                           B.groupUnitOffset[w + 1] == B.groupUnitOffset[w] && B.elimOffset[w + 1] == B.elimOffset[w];
        if (f[kFlagInvalid])
        {
            // an invalid projection in a build fails the solve (the single-problem path reports NUMERIC)
            B.status[w] = kWinNumeric;
            B.active[w] = 0;
            B.doneClock[w] = deviceClock();
        }
        else if (empty)
        {
            // nothing to refine: converged at the initial objective
            B.status[w] = kWinConverged;
            B.active[w] = 0;
            B.doneClock[w] = deviceClock();
        }
        else if (f[kFlagSingular] || f[kFlagNotSpd])
        {
            B.reject[w] = 1;
        }
    }

    // ---------------------------------------------------------------------------------------------
    // Dense window solve. The window's reduced system (6 * units rows, both triangles of the block-sparse S)
    // is gathered into a column-major FP64 matrix A, factored in place by a left-looking blocked Cholesky with
    // column panels of NB, and solved by a blocked forward substitution and a right-looking blocked backward
    // substitution. Thread t of the NT-thread workgroup owns the panel rows j0 + t + r * NT (r < R, R * NT =
    // kMaxDense). Every value is computed by one thread in an order that NT and NB do not change:
    //   L(i,j) = A(i,j) - L(i,0) L(j,0) - L(i,1) L(j,1) - ... - L(i,j-1) L(j,j-1), then sqrt (i = j) or / L(j,j)
    //   y_i    = b_i - L(i,0) y_0 - ... - L(i,i-1) y_{i-1}, then / L(i,i)
    //   x_i    = y_i - L(n-1,i) x_{n-1} - ... - L(i+1,i) x_{i+1}, then / L(i,i)
    // (each "-" one fused multiply-subtract), so every launch shape gives bitwise the same step and the batch
    // driver can pick the shape by time alone. A failed pivot (not positive, or not finite) rejects the step, as
    // a breakdown of the iterative solve does on the single-problem path.

    // acc[r][jj] -= sum_{k < j0} L(row_r, k) L(j0 + jj, k) for the RA active row slots of this thread, k ascending;
    // Bt holds L(j0 + jj, k) at k * NB + jj. The loads of the next group of columns are issued before the
    // arithmetic of the current one (two register buffers), so their latency overlaps it.
    template <int NT, int NB, int RA, int R>
    __device__ __forceinline__ void panelUpdate(const double* A, const double* Bt, int n, int j0, int tid,
                                                double (&acc)[R][NB])
    {
        constexpr int L0 = (R * NB >= 32 || NT >= 512) ? 8 : 16;
        constexpr int L = L0 < NB ? L0 : NB;  // columns per load group (j0 is a multiple of NB, so of L)
        if (j0 == 0) return;
        int row[RA];
#pragma unroll
        for (int r = 0; r < RA; ++r) row[r] = j0 + tid + r * NT;
        double a0[L][RA], a1[L][RA];
        auto load = [&](double (&a)[L][RA], int k0) {
#pragma unroll
            for (int u = 0; u < L; ++u)
#pragma unroll
                for (int r = 0; r < RA; ++r) a[u][r] = row[r] < n ? A[size_t(k0 + u) * n + row[r]] : 0.0;
        };
        auto update = [&](const double (&a)[L][RA], int k0) {
#pragma unroll
            for (int u = 0; u < L; ++u)
            {
                const double2* b2 = reinterpret_cast<const double2*>(Bt + (k0 + u) * NB);
#pragma unroll
                for (int q = 0; q < NB / 2; ++q)
                {
                    const double2 b = b2[q];
#pragma unroll
                    for (int r = 0; r < RA; ++r)
                    {
                        acc[r][2 * q] -= a[u][r] * b.x;
                        acc[r][2 * q + 1] -= a[u][r] * b.y;
                    }
                }
            }
        };
        load(a0, 0);
        for (int k0 = 0; k0 < j0; k0 += 2 * L)
        {
            if (k0 + L < j0) load(a1, k0 + L);
            update(a0, k0);
            if (k0 + L >= j0) break;
            if (k0 + 2 * L < j0) load(a0, k0 + 2 * L);
            update(a1, k0 + L);
        }
    }

    // Two workgroups per CU (every window of a 512-window batch resident at once) need NT / 128 waves per SIMD.
    template <typename T, int NT, int NB>
    __global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(NT >= 256 ? NT / 128 : 1)))
    k_batch_factor(Problem P, Work<T> Wk, Batch B)
    {
        constexpr int R = kMaxDense / NT;
        static_assert(R * NT == kMaxDense && NB % 2 == 0 && NB <= kWave, "factorisation shape");
        extern __shared__ double lds[];
        const int w = blockIdx.x;
        if (!B.active[w] || B.reject[w]) return;
        const int pu0 = B.poseUnitOffset[w], pu1 = B.poseUnitOffset[w + 1];
        const int gu0 = B.groupUnitOffset[w];
        const int nPose = pu1 - pu0;
        const int nU = nPose + (B.groupUnitOffset[w + 1] - gu0);
        const int n = 6 * nU;
        if (n == 0) return;
        const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
        double* Bt = lds;                         // [maxDim][NB]  L(j0 + jj, k) at k * NB + jj
        double* Dg = Bt + size_t(B.maxDim) * NB;  // [NB][NB]      diagonal block of the panel, row-major
        double* vec = Dg + NB * NB;               // [maxDim]      right-hand side / solution
        __shared__ int failed;
        if (tid == 0) failed = 0;
        double* A = B.dense + B.denseOffset[w];
        auto unitOf = [&](int r) { return r < nPose ? pu0 + r : gu0 + (r - nPose); };
        auto localOf = [&](int u) { return (u >= pu0 && u < pu1) ? u - pu0 : nPose + (u - gu0); };

        // gather: zero, then every element of the stored blocks of the window's rows (the pose rows and the
        // group rows are two contiguous ranges of the block-sparse rows) goes to its place, all loads independent
        for (long i = tid; i < long(n) * n; i += NT) A[i] = 0.0;
        for (int i = tid; i < n; i += NT) vec[i] = Wk.rhs[unitOf(i / 6) * 6 + i % 6];
        __syncthreads();
        {
            const int kp0 = P.rowPtr[pu0], nbp = P.rowPtr[pu1] - kp0;
            const int kg0 = P.rowPtr[gu0], nbg = P.rowPtr[B.groupUnitOffset[w + 1]] - kg0;
            const long total = long(nbp + nbg) * 36;
            for (long e0 = tid; e0 < total; e0 += kDenseLoads * NT)
            {
                double v[kDenseLoads];
                size_t at[kDenseLoads];
#pragma unroll
                for (int u = 0; u < kDenseLoads; ++u)
                {
                    const long e = e0 + long(u) * NT;
                    const long ec = e < total ? e : total - 1;
                    const int bi = int(ec / 36), el = int(ec % 36);
                    const int k = bi < nbp ? kp0 + bi : kg0 + (bi - nbp);
                    const int r = localOf(P.nnzRow[k]), c = localOf(P.colIdx[k]);
                    at[u] = size_t(6 * c + el % 6) * n + 6 * r + el / 6;
                    v[u] = double(Wk.S[size_t(k) * 36 + el]);
                }
#pragma unroll
                for (int u = 0; u < kDenseLoads; ++u)
                    if (e0 + long(u) * NT < total) A[at[u]] = v[u];
            }
        }
        __syncthreads();

        // factorisation
        for (int j0 = 0; j0 < n; j0 += NB)
        {
            const int jb = min(NB, n - j0);
            // staging: groups of kDenseLoads independent loads per thread before the LDS stores
            for (int q0 = tid; q0 < j0 * NB; q0 += kDenseLoads * NT)
            {
                double v[kDenseLoads];
#pragma unroll
                for (int u = 0; u < kDenseLoads; ++u)
                {
                    const int q = q0 + u * NT, k = q / NB, jj = q % NB;
                    v[u] = (q < j0 * NB && jj < jb) ? A[size_t(k) * n + j0 + jj] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < kDenseLoads; ++u)
                    if (q0 + u * NT < j0 * NB) Bt[q0 + u * NT] = v[u];
            }
            __syncthreads();
            double acc[R][NB];
#pragma unroll
            for (int r = 0; r < R; ++r)
            {
                const int i = j0 + tid + r * NT;
#pragma unroll
                for (int jj = 0; jj < NB; ++jj) acc[r][jj] = (i < n && jj < jb) ? A[size_t(j0 + jj) * n + i] : 0.0;
            }
            // row slots of this wave that reach into the matrix (wave-uniform)
            const int first = j0 + (tid & ~63);
            int ra = 0;
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (first + r * NT < n) ra = r + 1;
            if constexpr (R >= 4)
                if (ra == 4) panelUpdate<NT, NB, 4>(A, Bt, n, j0, tid, acc);
            if constexpr (R >= 3)
                if (ra == 3) panelUpdate<NT, NB, 3>(A, Bt, n, j0, tid, acc);
            if constexpr (R >= 2)
                if (ra == 2) panelUpdate<NT, NB, 2>(A, Bt, n, j0, tid, acc);
            if (ra == 1) panelUpdate<NT, NB, 1>(A, Bt, n, j0, tid, acc);
            // the diagonal block goes to LDS; wave 0 factors it in registers (lane l holds row l)
#pragma unroll
            for (int r = 0; r < R; ++r)
            {
                const int i = j0 + tid + r * NT;
                if (i < j0 + jb)
                {
#pragma unroll
                    for (int jj = 0; jj < NB; ++jj) Dg[(i - j0) * NB + jj] = acc[r][jj];
                }
            }
            __syncthreads();
            if (wv == 0)
            {
                double rowv[NB];
#pragma unroll
                for (int jj = 0; jj < NB; ++jj)
                    rowv[jj] = lane < jb ? (jj < jb ? Dg[lane * NB + jj] : 0.0) : (jj == lane ? 1.0 : 0.0);
                bool bad = false;
                // columns j >= jb of a short last panel are identity (rows of lanes >= jb too): no effect
#pragma unroll
                for (int j = 0; j < NB; ++j)
                {
                    const double d = __shfl(rowv[j], j, kWave);
                    const bool ok = d > 0.0 && pbdev::finite(d);
                    bad = bad || (j < jb && !ok);
                    const double ljj = ok ? sqrt(d) : 1.0;
                    if (lane == j) rowv[j] = ljj;
                    if (lane > j) rowv[j] /= ljj;
#pragma unroll
                    for (int q = j + 1; q < NB; ++q)
                    {
                        const double lq = __shfl(rowv[j], q, kWave);
                        if (lane > j) rowv[q] -= rowv[j] * lq;
                    }
                }
                if (lane < jb)
                {
#pragma unroll
                    for (int jj = 0; jj < NB; ++jj) Dg[lane * NB + jj] = jj <= lane ? rowv[jj] : 0.0;
                }
                if (lane == 0 && bad) failed = 1;
            }
            __syncthreads();
            // rows below the diagonal block: L(i, panel) = acc L_D^-T (forward substitution over the panel)
#pragma unroll
            for (int r = 0; r < R; ++r)
            {
                const int i = j0 + tid + r * NT;
                if (i >= n) continue;
                if (i < j0 + jb)
                {
#pragma unroll
                    for (int jj = 0; jj < NB; ++jj) acc[r][jj] = Dg[(i - j0) * NB + jj];
                }
                else
                {
#pragma unroll
                    for (int j = 0; j < NB; ++j)
                    {
                        double s = acc[r][j];
#pragma unroll
                        for (int q = 0; q < j; ++q) s -= acc[r][q] * Dg[j * NB + q];
                        acc[r][j] = j < jb ? s / Dg[j * NB + j] : 0.0;
                    }
                }
#pragma unroll
                for (int jj = 0; jj < NB; ++jj)
                    if (jj < jb) A[size_t(j0 + jj) * n + i] = acc[r][jj];
            }
            __syncthreads();
        }
        if (failed)
        {
            if (tid == 0) B.reject[w] = 1;
            return;
        }

        // forward substitution L y = b: wave 0 solves the panel's triangle, then every row below subtracts the
        // panel's terms
        for (int j0 = 0; j0 < n; j0 += NB)
        {
            const int jb = min(NB, n - j0);
            if (wv == 0)
            {
                double rowv[NB];
#pragma unroll
                for (int jj = 0; jj < NB; ++jj)
                    rowv[jj] = (lane < jb && jj <= lane) ? A[size_t(j0 + jj) * n + j0 + lane] : 0.0;
                double val = lane < jb ? vec[j0 + lane] : 0.0;
#pragma unroll
                for (int j = 0; j < NB; ++j)
                {
                    if (j < jb)
                    {
                        const double xj = __shfl(val, j, kWave) / __shfl(rowv[j], j, kWave);
                        if (lane > j) val -= rowv[j] * xj;
                        if (lane == j) val = xj;
                    }
                }
                if (lane < jb) vec[j0 + lane] = val;
            }
            __syncthreads();
            for (int i = j0 + jb + tid; i < n; i += NT)
            {
                double s = vec[i];
#pragma unroll
                for (int jj = 0; jj < NB; ++jj)
                    if (jj < jb) s -= A[size_t(j0 + jj) * n + i] * vec[j0 + jj];
                vec[i] = s;
            }
            __syncthreads();
        }
        // backward substitution L^T x = y, panels from the last: wave 0 solves the panel's triangle (lane l
        // holds column l of the panel's L, i.e. row l of L^T), then every row above subtracts the panel's terms,
        // highest index first
        for (int j0 = ((n - 1) / NB) * NB; j0 >= 0; j0 -= NB)
        {
            const int jb = min(NB, n - j0);
            if (wv == 0)
            {
                double colv[NB];  // colv[jj] = L(j0 + jj, j0 + lane) for jj >= lane
#pragma unroll
                for (int jj = 0; jj < NB; ++jj)
                    colv[jj] = (lane < jb && jj < jb && jj >= lane) ? A[size_t(j0 + lane) * n + j0 + jj] : 0.0;
                double val = lane < jb ? vec[j0 + lane] : 0.0;
#pragma unroll
                for (int j = NB - 1; j >= 0; --j)
                {
                    if (j < jb)
                    {
                        const double xj = __shfl(val, j, kWave) / __shfl(colv[j], j, kWave);
                        if (lane < j) val -= colv[j] * xj;
                        if (lane == j) val = xj;
                    }
                }
                if (lane < jb) vec[j0 + lane] = val;
            }
            __syncthreads();
            for (int i = tid; i < j0; i += NT)
            {
                double s = vec[i];
#pragma unroll
                for (int jj = NB - 1; jj >= 0; --jj)
                    if (jj < jb) s -= A[size_t(i) * n + j0 + jj] * vec[j0 + jj];
                vec[i] = s;
            }
            __syncthreads();
        }
        for (int i = tid; i < n; i += NT) Wk.x[unitOf(i / 6) * 6 + i % 6] = T(vec[i]);
    }

    // Tile factorisation of each window's system (the default): the system gathered into a column-major matrix of
    // order np = 64 nt with identity padding, factored right-looking in 64 x 64 tiles by the LDS tile routines of
    // the band route (one-wave panel Cholesky of the diagonal tile, its triangular inverse, MFMA 16x16x4 products
    // written straight to memory); each diagonal tile is replaced by the inverse of its factor, so the forward
    // and backward solves are tile products: y_k = Linv_kk (b_k - sum_{j<k} L_kj y_j), x_k = Linv_kk^T (y_k -
    // sum_{i>k} L_ik^T x_i). Two LDS tiles per workgroup, two workgroups per CU. Each value is computed in one
    // fixed order (tile k ascending), so repeated batches are bitwise identical; the panel factorisation
    // (k_batch_factor) orders its sums differently and agrees to rounding.
    constexpr int kTileLdsBytes = int(2 * kDenseTile * (kDenseTile + 1) * sizeof(double) +
                                      (kDenseTile + kMaxDense + 256) * sizeof(double));

    template <typename T>
    __global__ void __launch_bounds__(256) k_batch_tiles(Problem P, Work<T> Wk, Batch B)
    {
        using namespace tile;
        extern __shared__ __align__(16) unsigned char tileLds[];
        double(*b0)[kLd] = reinterpret_cast<double(*)[kLd]>(tileLds);
        double(*b1)[kLd] = b0 + kT;
        double* col = reinterpret_cast<double*>(b1 + kT);  // [64] pivot column / tile vector
        double* vec = col + kT;                            // [kMaxDense] right-hand side, then solution
        double* part = vec + kMaxDense;                    // [256] partial sums
        const int w = blockIdx.x;
        if (!B.active[w] || B.reject[w]) return;
        const int pu0 = B.poseUnitOffset[w], pu1 = B.poseUnitOffset[w + 1];
        const int gu0 = B.groupUnitOffset[w];
        const int nPose = pu1 - pu0;
        const int nU = nPose + (B.groupUnitOffset[w + 1] - gu0);
        const int n = 6 * nU;
        if (n == 0) return;
        const int nt = (n + kT - 1) / kT, np = nt * kT;
        const int tid = threadIdx.x;
        double* A = B.dense + B.denseOffset[w];
        auto unitOf = [&](int r) { return r < nPose ? pu0 + r : gu0 + (r - nPose); };
        auto localOf = [&](int u) { return (u >= pu0 && u < pu1) ? u - pu0 : nPose + (u - gu0); };

        // gather: identity, then every element of the stored blocks of the window's rows to its place
        for (long i = tid; i < long(np) * np; i += 256) A[i] = (i % np == i / np && i % np >= n) ? 1.0 : 0.0;
        for (int i = tid; i < np; i += 256) vec[i] = i < n ? Wk.rhs[unitOf(i / 6) * 6 + i % 6] : 0.0;
        __syncthreads();
        {
            const int kp0 = P.rowPtr[pu0], nbp = P.rowPtr[pu1] - kp0;
            const int kg0 = P.rowPtr[gu0], nbg = P.rowPtr[B.groupUnitOffset[w + 1]] - kg0;
            const long total = long(nbp + nbg) * 36;
            for (long e0 = tid; e0 < total; e0 += kDenseLoads * 256)
            {
                double v[kDenseLoads];
                size_t at[kDenseLoads];
#pragma unroll
                for (int u = 0; u < kDenseLoads; ++u)
                {
                    const long e = e0 + long(u) * 256;
                    const long ec = e < total ? e : total - 1;
                    const int bi = int(ec / 36), el = int(ec % 36);
                    const int k = bi < nbp ? kp0 + bi : kg0 + (bi - nbp);
                    const int r = localOf(P.nnzRow[k]), c = localOf(P.colIdx[k]);
                    at[u] = size_t(6 * c + el % 6) * np + 6 * r + el / 6;
                    v[u] = double(Wk.S[size_t(k) * 36 + el]);
                }
#pragma unroll
                for (int u = 0; u < kDenseLoads; ++u)
                    if (e0 + long(u) * 256 < total) A[at[u]] = v[u];
            }
        }
        __syncthreads();

        // right-looking tile Cholesky
        bool bad = false;
        for (int k = 0; k < nt; ++k)
        {
            loadTile(b0, A, np, k * kT, k * kT);
            __syncthreads();
            cholTile(b0, col, bad);
            invTile(b0, b1, col);  // b1 = L_kk^-1 (synchronised inside)
            storeTile(A, np, k * kT, k * kT, b1);
            for (int i = k + 1; i < nt; ++i)
            {
                __syncthreads();
                loadTile(b0, A, np, i * kT, k * kT);
                __syncthreads();
                tileProductGlobal(A + size_t(k * kT) * np + i * kT, np, b0, false, b1, true, false);  // L_ik
            }
            for (int j = k + 1; j < nt; ++j)
            {
                __syncthreads();
                loadTile(b1, A, np, j * kT, k * kT);  // L_jk
                for (int i = j; i < nt; ++i)
                {
                    __syncthreads();
                    loadTile(b0, A, np, i * kT, k * kT);  // L_ik
                    __syncthreads();
                    tileProductGlobal(A + size_t(j * kT) * np + i * kT, np, b0, false, b1, true, true);
                }
            }
            __syncthreads();
        }
        __shared__ int failed;
        if (tid == 0) failed = 0;
        __syncthreads();
        if (bad) failed = 1;
        __syncthreads();
        if (failed)
        {
            if (tid == 0) B.reject[w] = 1;
            return;
        }

        // forward: y_k = Linv_kk (b_k - sum_{j<k} L_kj y_j); thread t sums row t & 63 over every fourth column
        const int rr = tid & 63, pp = tid >> 6;
        for (int k = 0; k < nt; ++k)
        {
            const int r0 = k * kT;
            double sacc = 0.0;
            for (int c = pp; c < r0; c += 4) sacc += A[size_t(c) * np + r0 + rr] * vec[c];
            part[tid] = sacc;
            __syncthreads();
            if (tid < kT) col[tid] = vec[r0 + tid] - ((part[tid] + part[64 + tid]) + (part[128 + tid] + part[192 + tid]));
            __syncthreads();
            if (tid < kT)
            {
                double y = 0.0;
                for (int c = 0; c <= tid; ++c) y += A[size_t(r0 + c) * np + r0 + tid] * col[c];
                vec[r0 + tid] = y;
            }
            __syncthreads();
        }
        // backward: x_k = Linv_kk^T (y_k - sum_{i>k} L_ik^T x_i)
        for (int k = nt - 1; k >= 0; --k)
        {
            const int r0 = k * kT;
            double sacc = 0.0;
            for (int i = r0 + kT + pp; i < np; i += 4) sacc += A[size_t(r0 + rr) * np + i] * vec[i];
            part[tid] = sacc;
            __syncthreads();
            if (tid < kT) col[tid] = vec[r0 + tid] - ((part[tid] + part[64 + tid]) + (part[128 + tid] + part[192 + tid]));
            __syncthreads();
            if (tid < kT)
            {
                double x = 0.0;
                for (int c = tid; c < kT; ++c) x += A[size_t(r0 + tid) * np + r0 + c] * col[c];
                vec[r0 + tid] = x;
            }
            __syncthreads();
        }
        for (int i = tid; i < n; i += 256) Wk.x[unitOf(i / 6) * 6 + i % 6] = T(vec[i]);
    }

    // Directional decrease dc.directRhs + sum bp.dp and the two step RMS values of each window
    // (BundleAdjustPlaMatrix.cpp:77-158), in a fixed reduction order.
    template <typename T>
    __global__ void __launch_bounds__(256) k_batch_step_stats(Problem P, Work<T> Wk, Batch B)
    {
        const int w = blockIdx.x;
        if (!B.active[w] || B.reject[w]) return;
        double dd = 0.0, gs = 0.0, ss = 0.0;
        const int pu0 = B.poseUnitOffset[w], nPose = B.poseUnitOffset[w + 1] - pu0;
        const int gu0 = B.groupUnitOffset[w], nU = nPose + B.groupUnitOffset[w + 1] - gu0;
        for (int r = threadIdx.x; r < nU; r += blockDim.x)
        {
            const int u = r < nPose ? pu0 + r : gu0 + (r - nPose);
            const int lim = P.unitWidth[u];
#pragma unroll
            for (int a = 0; a < 6; ++a)
            {
                const double x = double(Wk.x[u * 6 + a]);
                dd += x * Wk.directRhs[u * 6 + a];
                if (a < lim) gs += x * x;
            }
        }
        for (int b = B.elimOffset[w] + threadIdx.x; b < B.elimOffset[w + 1]; b += blockDim.x)
        {
#pragma unroll
            for (int i = 0; i < 3; ++i)
            {
                const double d = Wk.dp[3 * b + i];
                dd += Wk.bp[3 * b + i] * d;
                ss += d * d;
            }
        }
        __shared__ double red[3][4];
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        dd = waveSum(dd);
        gs = waveSum(gs);
        ss = waveSum(ss);
        if (lane == 0)
        {
            red[0][wv] = dd;
            red[1][wv] = gs;
            red[2][wv] = ss;
        }
        __syncthreads();
        if (threadIdx.x != 0) return;
        dd = gs = ss = 0.0;
        for (int q = 0; q < int(blockDim.x / kWave); ++q)
        {
            dd += red[0][q];
            gs += red[1][q];
            ss += red[2][q];
        }
        const int gd = B.globalDim[w], ne = B.elimOffset[w + 1] - B.elimOffset[w];
        B.dd[w] = dd;
        B.rmsG[w] = gd > 0 ? sqrt(gs / gd) : 0.0;
        B.rmsP[w] = ne > 0 ? sqrt(ss / (3.0 * ne)) : 0.0;
        if (!pbdev::finite(dd))
        {
            B.reject[w] = 1;
            return;
        }
        B.searching[w] = 1;
        B.lsCount[w] = 0;
        B.alphaNext[w] = 1.0;
    }

    // Step lengths of the next trial pass: alphaNext * 2^-k for the windows still searching.
    __global__ void k_batch_alphas(Batch B, int maxAlpha)
    {
        const int w = blockIdx.x * blockDim.x + threadIdx.x;
        if (w == 0) B.counts[2] = 0;
        if (w >= B.nWin) return;
        if (!B.searching[w])
        {
            B.nAlpha[w] = 0;
            return;
        }
        const int n = min(maxAlpha, B.lineSearchSteps - B.lsCount[w]);
        for (int k = 0; k < n; ++k) B.alpha[w * kMaxAlpha + k] = ldexp(B.alphaNext[w], -k);
        B.nAlpha[w] = n;
    }

    // Trial objectives: each workgroup covers one chunk of one window's observations.
    template <typename T, int NA>
    __global__ void __launch_bounds__(256) k_batch_trial_cost(Problem P, Work<T> Wk, Batch B)
    {
        const int ch = blockIdx.x;
        const int w = B.chunkWin[ch];
        const int nA = B.nAlpha[w];
        if (nA == 0) return;
        double acc[NA], al[NA];
        int bad[NA];
#pragma unroll
        for (int a = 0; a < NA; ++a)
        {
            acc[a] = 0.0;
            bad[a] = 0;
            al[a] = a < nA ? B.alpha[w * kMaxAlpha + a] : 0.0;
        }
        for (int o = B.chunkBegin[ch] + threadIdx.x; o < B.chunkEnd[ch]; o += blockDim.x)
            trialObservation<NA>(P, Wk.candCams, Wk.dp, o, al, nA, acc, bad);
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
        if (threadIdx.x < NA)
        {
            const int a = threadIdx.x;
            double s = 0.0;
            int bb = 0;
            for (int q = 0; q < int(blockDim.x / kWave); ++q)
            {
                s += lds[a][q];
                bb |= lbad[a][q];
            }
            B.trialPart[ch * kMaxAlpha + a] = s;
            B.trialBad[ch * kMaxAlpha + a] = bb;
        }
    }

    // Sums each window's chunk partials per step length (lane-strided, then the wave butterfly) and runs the
    // Armijo test over the pass's step lengths in sequential order (the first that passes is accepted, as in
    // the reference's backtracking loop).
    template <typename T, int NA>
    __global__ void k_batch_decide(Problem P, Work<T> Wk, Batch B)
    {
        const int lane = threadIdx.x & 63;
        const int w = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
        if (w >= B.nWin) return;
        const int nA = B.nAlpha[w];
        if (nA == 0) return;
        double s[NA];
        int bad[NA];
#pragma unroll
        for (int a = 0; a < NA; ++a)
        {
            s[a] = 0.0;
            bad[a] = 0;
        }
        for (int k = B.winChunkOffset[w] + lane; k < B.winChunkOffset[w + 1]; k += kWave)
        {
#pragma unroll
            for (int a = 0; a < NA; ++a)
            {
                if (a >= nA) break;
                s[a] += B.trialPart[k * kMaxAlpha + a];
                bad[a] |= B.trialBad[k * kMaxAlpha + a];
            }
        }
        bool anyBad[NA];
#pragma unroll
        for (int a = 0; a < NA; ++a)
        {
            s[a] = waveSum(s[a]);
            anyBad[a] = __any(bad[a] != 0);
        }
        if (lane != 0) return;
        int32_t* cnt = B.counters + w * kWinCounterCount;
        const double current = B.cost[w], dd = B.dd[w];
        bool accepted = false;
#pragma unroll
        for (int a = 0; a < NA; ++a)
        {
            if (a >= nA || accepted) break;
            double cand = s[a];
            for (int g = B.groupOffset[w]; g < B.groupOffset[w + 1]; ++g) cand += Wk.candPriorCost[a * P.nGroup + g];
            if (anyBad[a]) cand = __builtin_inf();
            ++cnt[kWinEvaluations];
            ++B.lsCount[w];
            const double alpha = B.alpha[w * kMaxAlpha + a];
            const double limit = current - B.armijo * alpha * dd;
            if (pbdev::finite(cand) && cand <= limit && cand < current)
            {
                accepted = true;
                B.acceptIdx[w] = a;
                B.acceptAlpha[w] = alpha;
                B.acceptCost[w] = cand;
            }
        }
        if (accepted || B.lsCount[w] >= B.lineSearchSteps)
        {
            B.searching[w] = 0;
            return;
        }
        B.alphaNext[w] = B.alpha[w * kMaxAlpha + nA - 1] * 0.5;
        atomicAdd(B.counts + 2, 1);
    }

    // Controller step of each window of this phase: counts, damping, acceptance, convergence and budget. In the
    // FP32 phase of a mixed batch the stop and the failure exits move the window to FP64 instead.
    __global__ void k_batch_control(Batch B, int phase)
    {
        const int w = blockIdx.x * blockDim.x + threadIdx.x;
        if (w >= B.nWin || !B.active[w]) return;
        int32_t* cnt = B.counters + w * kWinCounterCount;
        ++cnt[kWinIterations];
        if (phase == 1) ++cnt[kWinFp64Iterations];
        double& lambda = B.damping[w];
        if (B.reject[w])
        {
            ++cnt[kWinRejected];
            lambda = lambda == 0.0 ? 1e-3 : lambda * 10.0;
            if (phase == 0) B.phase[w] = 1;
        }
        else if (B.acceptIdx[w] >= 0)
        {
            ++cnt[kWinAccepted];
            const double before = B.cost[w];
            B.cost[w] = B.acceptCost[w];
            lambda /= 10.0;
            const double a = B.acceptAlpha[w];
            const bool stop = a * B.rmsG[w] <= B.rmsTol && a * B.rmsP[w] <= B.rmsTol;
            if (phase == 0)
            {
                if (stop || before - B.acceptCost[w] < B.switchDecrease * before) B.phase[w] = 1;
            }
            else if (stop)
            {
                B.status[w] = kWinConverged;
            }
        }
        else
        {
            // line search exhausted: the controller stops (BundleAdjustPlaMatrix.cpp:532-536)
            ++cnt[kWinRejected];
            if (phase == 0)
                B.phase[w] = 1;
            else
                B.status[w] = kWinNoConvergence;
        }
        if (B.status[w] == kWinRunning && cnt[kWinIterations] >= B.maxIterations) B.status[w] = kWinNoConvergence;
        if (B.status[w] != kWinRunning) B.doneClock[w] = deviceClock();
    }

    // Running windows per phase (one workgroup, fixed reduction order).
    __global__ void k_batch_count(Batch B)
    {
        __shared__ int red[2][4];
        int c0 = 0, c1 = 0;
        for (int w = threadIdx.x; w < B.nWin; w += blockDim.x)
            if (B.status[w] == kWinRunning) (B.phase[w] == 0 ? c0 : c1) += 1;
        c0 = waveSum(c0);
        c1 = waveSum(c1);
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        if (lane == 0)
        {
            red[0][wv] = c0;
            red[1][wv] = c1;
        }
        __syncthreads();
        if (threadIdx.x == 0)
        {
            int s0 = 0, s1 = 0;
            for (int q = 0; q < int(blockDim.x / kWave); ++q)
            {
                s0 += red[0][q];
                s1 += red[1][q];
            }
            B.counts[0] = s0;
            B.counts[1] = s1;
        }
    }

    // Commits the accepted candidate of each window: cameras, group parameters and points.
    template <typename T>
    __global__ void k_batch_apply(Problem P, Work<T> Wk, Batch B)
    {
        const int id = blockIdx.x * blockDim.x + threadIdx.x;
        if (id < P.nCam)
        {
            const int ai = B.acceptIdx[P.camWin[id]];
            if (ai >= 0)
            {
                for (int i = 0; i < 9; ++i) P.camR[9 * id + i] = Wk.candR[(ai * P.nCam + id) * 9 + i];
                for (int i = 0; i < 3; ++i) P.camC[3 * id + i] = Wk.candC[(ai * P.nCam + id) * 3 + i];
            }
        }
        if (id < P.nGroup * 13)
        {
            const int ai = B.acceptIdx[P.groupWin[id / 13]];
            if (ai >= 0) P.groupParams[id] = Wk.candGroupParams[ai * P.nGroup * 13 + id];
        }
        if (id < P.nElim)
        {
            const int w = P.elimWin[id];
            if (B.acceptIdx[w] >= 0)
            {
                const double alpha = B.acceptAlpha[w];
                const int t = P.elimTrack[id];
                const double s = stepScale(P, P.pointX[3 * t], P.pointX[3 * t + 1], P.pointX[3 * t + 2]);
                for (int i = 0; i < 3; ++i) P.pointX[3 * t + i] += s * alpha * Wk.dp[3 * id + i];
            }
        }
    }

    // ---------------------------------------------------------------------------------------------
    template <typename T>
    void BatchLauncher<T>::start(const Batch& B, hipStream_t s)
    {
        k_batch_start<<<1, 64, 0, s>>>(B);
    }

    template <typename T>
    void BatchLauncher<T>::count(const Batch& B, hipStream_t s)
    {
        k_batch_count<<<1, 256, 0, s>>>(B);
    }

    template <typename T>
    void BatchLauncher<T>::roundBegin(const Problem& P, const Work<T>&, const Batch& B, int phase, hipStream_t s)
    {
        k_batch_round_begin<<<gridFor(B.nWin, 256), 256, 0, s>>>(P, B, phase);
    }

    template <typename T>
    void BatchLauncher<T>::build(const Problem& P, const Work<T>& W, hipStream_t s)
    {
        // the shared kernels read each window's damping and activity from P
        Launcher<T>::prepareCameras(P, s);
        Launcher<T>::linearizePoints(P, W, 0.0, s);
        Launcher<T>::cameraDirect(P, W, s);
        Launcher<T>::schurBlocks(P, W, 0.0, s);
    }

    template <typename T>
    void BatchLauncher<T>::afterBuild(const Problem& P, const Work<T>& W, const Batch& B, hipStream_t s)
    {
        k_batch_after_build<T><<<gridFor(B.nWin, 4), 256, 0, s>>>(P, W, B);
    }

    namespace
    {
        size_t factorLds(const Batch& B, int shape)
        {
            const size_t nb = size_t(kFactorShape[shape].panel);
            return sizeof(double) * (size_t(B.maxDim) * nb + nb * nb + size_t(B.maxDim));
        }

        template <int V, typename T>
        constexpr auto factorKernel()
        {
            return &k_batch_factor<T, kFactorShape[V].threads, kFactorShape[V].panel>;
        }

        template <typename T>
        const void* factorFunction(int shape)
        {
            switch (shape)
            {
                case 0: return reinterpret_cast<const void*>(factorKernel<0, T>());
                case 1: return reinterpret_cast<const void*>(factorKernel<1, T>());
                case 2: return reinterpret_cast<const void*>(factorKernel<2, T>());
                case 3: return reinterpret_cast<const void*>(factorKernel<3, T>());
                case 4: return reinterpret_cast<const void*>(factorKernel<4, T>());
                default: return reinterpret_cast<const void*>(factorKernel<5, T>());
            }
        }
        static_assert(kFactorShapes == 6, "factorFunction and BatchLauncher::factor list every shape");
    }  // namespace

    template <typename T>
    void BatchLauncher<T>::configure(const Batch& B, bool (&usable)[kFactorShapes])
    {
        if (hipFuncSetAttribute(reinterpret_cast<const void*>(&k_batch_tiles<T>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, kTileLdsBytes) != hipSuccess)
            throw std::runtime_error("hipFuncSetAttribute (tile window factorisation LDS)");
        for (int v = 0; v < kFactorShapes; ++v)
        {
            const size_t lds = factorLds(B, v);
            usable[v] = lds <= size_t(kFactorLdsLimit);
            // above the default dynamic LDS limit (kMaxDense = 512 with 16-wide panels needs 66 KiB of the CU's
            // 160 KiB)
            if (usable[v] && lds > 64 * 1024)
                usable[v] = hipFuncSetAttribute(factorFunction<T>(v), hipFuncAttributeMaxDynamicSharedMemorySize,
                                                int(lds)) == hipSuccess;
        }
        if (!usable[0]) throw std::runtime_error("hipFuncSetAttribute (dense window solve LDS)");
    }

    template <typename T>
    void BatchLauncher<T>::factor(const Problem& P, const Work<T>& W, const Batch& B, int shape, hipStream_t s)
    {
        if (shape == kFactorTiles)
        {
            k_batch_tiles<T><<<B.nWin, 256, kTileLdsBytes, s>>>(P, W, B);
            return;
        }
        const size_t lds = factorLds(B, shape);
        switch (shape)
        {
            case 0: factorKernel<0, T>()<<<B.nWin, kFactorShape[0].threads, lds, s>>>(P, W, B); break;
            case 1: factorKernel<1, T>()<<<B.nWin, kFactorShape[1].threads, lds, s>>>(P, W, B); break;
            case 2: factorKernel<2, T>()<<<B.nWin, kFactorShape[2].threads, lds, s>>>(P, W, B); break;
            case 3: factorKernel<3, T>()<<<B.nWin, kFactorShape[3].threads, lds, s>>>(P, W, B); break;
            case 4: factorKernel<4, T>()<<<B.nWin, kFactorShape[4].threads, lds, s>>>(P, W, B); break;
            case 5: factorKernel<5, T>()<<<B.nWin, kFactorShape[5].threads, lds, s>>>(P, W, B); break;
            default: throw std::invalid_argument("factorisation shape out of range");
        }
    }

    template <typename T>
    void BatchLauncher<T>::backSubstitute(const Problem& P, const Work<T>& W, const Batch& B, hipStream_t s)
    {
        Launcher<T>::pointSteps(P, W, s);
        k_batch_step_stats<T><<<B.nWin, 256, 0, s>>>(P, W, B);
    }

    template <typename T>
    void BatchLauncher<T>::trialPass(const Problem& P, const Work<T>& W, const Batch& B, int maxAlpha, hipStream_t s)
    {
        k_batch_alphas<<<gridFor(B.nWin, 256), 256, 0, s>>>(B, maxAlpha);
        Launcher<T>::trialStates(P, W, nullptr, maxAlpha, s);  // step lengths per window from P.winAlpha
        if (B.nChunks > 0) k_batch_trial_cost<T, kMaxAlpha><<<B.nChunks, 256, 0, s>>>(P, W, B);
        k_batch_decide<T, kMaxAlpha><<<gridFor(B.nWin, 4), 256, 0, s>>>(P, W, B);
    }

    template <typename T>
    void BatchLauncher<T>::finishRound(const Problem& P, const Work<T>& W, const Batch& B, int phase, hipStream_t s)
    {
        k_batch_control<<<gridFor(B.nWin, 256), 256, 0, s>>>(B, phase);
        const long n = std::max<long>(std::max<long>(P.nCam, long(P.nGroup) * 13), P.nElim);
        if (n > 0) k_batch_apply<T><<<gridFor(n, 256), 256, 0, s>>>(P, W, B);
    }

    template struct BatchLauncher<float>;
    template struct BatchLauncher<double>;
}  // namespace pbk
