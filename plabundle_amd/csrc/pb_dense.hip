// pb_dense.hip - dense solver of the reduced camera system (gfx950, wave64), FP64 or FP32.
//
// The reduced system S (block-sparse 6x6 rows) is gathered into a dense column-major matrix whose order is
// padded to a multiple of kDenseTile with an identity block, so every kernel works on whole 64 x 64 tiles:
//   k_dense_gather   one workgroup per lower tile: zero, scatter the stored blocks that meet the tile, store
//   k_dense_panel    step k of the right-looking Cholesky: every workgroup of the launch factors diagonal tile
//                    k in registers and LDS (redundantly, so no workgroup waits for another) and solves its own
//                    tile of panel column k (the 128 x 64 panel [D; X] is eliminated column by column)
//   k_dense_update   step k trailing update A_ij -= L_ik L_jk^T of the lower tiles, MFMA 16x16x4 (f64 or f32)
//   k_dense_lower / k_dense_upper
//                    L y = b and L^T x = y in one launch each: workgroup t takes tile t in dispatch order (an
//                    atomic ticket), streams its off-diagonal tiles while it waits for the tiles it depends on
//                    (agent-scope release/acquire on an in-order progress word tagged with the solve's epoch,
//                    bounded spin), then solves its diagonal tile in one wave and publishes
// Only the lower triangle is read (the trailing update rewrites whole diagonal tiles, which stay symmetric).
// This replaces the reference's DenseCpu route of the reduced solve (BundleAdjustPlaMatrixRuntime.cpp:22-25,
// called at BundleAdjustPlaMatrix.cpp:349-369) on the device.
#include "pb_common.h"
#include "pb_kernels.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <stdexcept>

namespace pbk
{
    namespace
    {
        constexpr int kT = kDenseTile;

        template <typename T>
        struct Mfma;
        // Result element e of lane l is row kLaneRow * (l >> 4) + kItemRow * e, column l & 15 of the 16 x 16 block
        // (the two instructions lay their results out differently; measured on gfx950).
        template <>
        struct Mfma<double>
        {
            static constexpr int kLaneRow = 1, kItemRow = 4;
            typedef double V __attribute__((ext_vector_type(4)));
            static __device__ __forceinline__ V op(double a, double b, V c)
            {
                return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
            }
        };
        template <>
        struct Mfma<float>
        {
            static constexpr int kLaneRow = 4, kItemRow = 1;
            typedef float V __attribute__((ext_vector_type(4)));
            static __device__ __forceinline__ V op(float a, float b, V c)
            {
                return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
            }
        };

        __device__ __forceinline__ double readLane(double v, int j)
        {
            const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(v));
            const unsigned lo = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(b & 0xffffffffu), j));
            const unsigned hi = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(b >> 32), j));
            return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
        }
        __device__ __forceinline__ float readLane(float v, int j)
        {
            return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
        }

        // 1/p from the hardware reciprocal refined by Newton's iteration r <- r + r (1 - p r) (two steps in FP64,
        // where the estimate carries about half the bits; one in FP32). Keeps the IEEE division off the serial
        // chain of the panel factorisation.
        __device__ __forceinline__ double recip(double p)
        {
            double r = __builtin_amdgcn_rcp(p);
            r = __builtin_fma(r, __builtin_fma(-p, r, 1.0), r);
            return __builtin_fma(r, __builtin_fma(-p, r, 1.0), r);
        }
        __device__ __forceinline__ float recip(float p)
        {
            const float r = __builtin_amdgcn_rcpf(p);
            return __builtin_fmaf(r, __builtin_fmaf(-p, r, 1.0f), r);
        }

        template <typename T>
        __device__ __forceinline__ bool positiveFinite(T v)
        {
            return v > T(0) && v < T(INFINITY);
        }

        // Row-major enumeration of the lower triangle of an m x m tile grid: b = row (row + 1) / 2 + col.
        __device__ __forceinline__ void lowerTile(int b, int& row, int& col)
        {
            int r = int((sqrtf(8.0f * float(b) + 1.0f) - 1.0f) * 0.5f);
            while (r * (r + 1) / 2 > b) --r;
            while ((r + 1) * (r + 2) / 2 <= b) ++r;
            row = r;
            col = b - r * (r + 1) / 2;
        }

        __device__ __forceinline__ int xcdMapped(int b, int n)
        {
            return (b % kXcds) * (n / kXcds) + b / kXcds;
        }
    }  // namespace

    // ---------------------------------------------------------------------------------------------
    // Gather: lower tile (ti, tj) of the padded dense matrix from the block-sparse rows of S.
    template <typename TW, typename TA>
    __global__ void __launch_bounds__(256) k_dense_gather(Problem P, const TW* S, TA* A, int lda, int nt, int grid)
    {
        __shared__ TA tile[kT][kT + 1];  // [column][row]
        const int b = xcdMapped(blockIdx.x, grid);
        if (b >= nt * (nt + 1) / 2) return;
        int ti, tj;
        lowerTile(b, ti, tj);
        const int ri = ti * kT, cj = tj * kT;
        const int n = 6 * P.nUnits;
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        for (int c = wv; c < kT; c += 4) tile[c][lane] = (ri + lane == cj + c && ri + lane >= n) ? TA(1) : TA(0);
        __syncthreads();
        if (ri < n && cj < n)
        {
            const int u0 = ri / 6, u1 = min((ri + kT - 1) / 6, P.nUnits - 1);
            const int v0 = cj / 6, v1 = min((cj + kT - 1) / 6, P.nUnits - 1);
            for (int u = u0 + wv; u <= u1; u += 4)
                for (int k = P.rowPtr[u] + lane; k < P.rowPtr[u + 1]; k += kWave)
                {
                    const int v = P.colIdx[k];
                    if (v < v0 || v > v1) continue;
                    const TW* blk = S + size_t(k) * 36;
                    for (int e = 0; e < 36; ++e)
                    {
                        const int r = 6 * u + e / 6 - ri, c = 6 * v + e % 6 - cj;
                        if (r >= 0 && r < kT && c >= 0 && c < kT) tile[c][r] = TA(blk[e]);
                    }
                }
        }
        __syncthreads();
        for (int c = wv; c < kT; c += 4) A[size_t(cj + c) * lda + ri + lane] = tile[c][lane];
    }

    // Sets up one factorisation or solve: a new epoch for the tile flags, fresh tickets, and (factorisation)
    // cleared failure words.
    __global__ void k_dense_begin(int32_t* words, int resetFailure)
    {
        if (threadIdx.x == 0)
        {
            words[kDenseEpoch] += 1;
            words[kDenseTicketLower] = 0;
            words[kDenseTicketUpper] = 0;
            if (resetFailure)
            {
                words[kDenseNotSpd] = 0;
                words[kDenseTimeout] = 0;
            }
        }
    }

    // ---------------------------------------------------------------------------------------------
    // Panel step k. Workgroup b takes tile row i = k + b; wave q of kPanelWaves holds the columns
    // q, q + kPanelWaves, ... of the diagonal tile D (lane = row) and of its own tile X. Column j of the stacked
    // panel [D; X] is eliminated with one barrier: the wave holding column j publishes it in LDS
    // (double-buffered by parity), every thread subtracts d_rj d_cj / d_jj from its entries of the columns c > j.
    // The published diagonal-tile column is stored permuted (entry of row c at [c % W][c / W]) so that wave q
    // reads the entries of its columns as one contiguous run. The columns are scaled by 1/sqrt(d_jj) at the end.
    constexpr int kPanelWaves = 8;
    template <typename T>
    __global__ void __launch_bounds__(64 * kPanelWaves) k_dense_panel(T* A, int lda, int k, int32_t* words)
    {
        constexpr int W = kPanelWaves, M = kT / W;           // waves, columns per thread
        constexpr int kRun = M + 16 / int(sizeof(T));        // M entries + padding to a 16-byte multiple
        __shared__ __align__(16) T colD[2][W][kRun];
        __shared__ T colX[2][kT];
        __shared__ T pivot[kT];
        const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
        const bool diagWg = blockIdx.x == 0;
        const size_t c0 = size_t(k) * kT, r0 = size_t(k + blockIdx.x) * kT;
        T d[M], x[M];
#pragma unroll
        for (int m = 0; m < M; ++m)
        {
            const size_t col = (c0 + q + W * m) * size_t(lda);
            d[m] = A[col + c0 + lane];
            x[m] = diagWg ? T(0) : A[col + r0 + lane];
        }
        bool bad = false;
#pragma unroll
        for (int j = 0; j < kT; ++j)
        {
            const int buf = j & 1;
            if (q == j % W)
            {
                colD[buf][lane % W][lane / W] = d[j / W];
                colX[buf][lane] = x[j / W];
            }
            __syncthreads();
            T piv = colD[buf][j % W][j / W];
            if (!positiveFinite(piv))
            {
                bad = true;
                piv = T(1);
            }
            const T ip = recip(piv);
            if (threadIdx.x == 0) pivot[j] = piv;
            const T dr = colD[buf][lane % W][lane / W] * ip, xr = colX[buf][lane] * ip;
            // the entries d_cj of this wave's columns, read together; columns c <= j are finished (a uniform
            // branch per column skips them)
            T dc[M];
#pragma unroll
            for (int m = 0; m < M; ++m) dc[m] = colD[buf][q][m];
#pragma unroll
            for (int m = 0; m < M; ++m)
            {
                if (q + W * m > j)
                {
                    d[m] -= dr * dc[m];
                    x[m] -= xr * dc[m];
                }
            }
        }
        __syncthreads();
        if (diagWg && bad && threadIdx.x == 0) atomicOr(words + kDenseNotSpd, 1);
#pragma unroll
        for (int m = 0; m < M; ++m)
        {
            const int c = q + W * m;
            const T s = sqrt(pivot[c]);
            const size_t col = (c0 + c) * size_t(lda);
            if (diagWg)
                A[col + c0 + lane] = lane > c ? d[m] / s : (lane == c ? s : T(0));
            else
                A[col + r0 + lane] = x[m] / s;
        }
    }

    // ---------------------------------------------------------------------------------------------
    // Update A_ij -= sum_{p = kFirst}^{kFirst + kCount - 1} L_ip L_jp^T of a set of lower tiles:
    //   columns mode (j1 > j0): tile columns j in [j0, j1), rows i in [j, nt) - the remaining columns of a block
    //                           after one of its panel steps;
    //   triangle mode (j1 == 0): every lower tile with j0 <= j <= i < nt - the trailing matrix after a block.
    // Wave w computes the 32 x 32 quadrant (w >> 1, w & 1) of the 64 x 64 tile as 2 x 2 MFMA 16x16x4 blocks;
    // the two panel slices are staged in LDS 32 columns at a time. Operand lanes: A(i, kk) from lane i + 16 kk,
    // B(kk, j) from lane j + 16 kk; result layout per Mfma<T>.
    template <typename T>
    __global__ void __launch_bounds__(256) k_dense_update(T* A, int lda, int kFirst, int kCount, int j0, int j1, int nt,
                                                          int grid)
    {
        using V = typename Mfma<T>::V;
        constexpr int kStage = 32;  // panel columns staged in LDS per pass
        __shared__ T sa[kStage][kT + 2], sb[kStage][kT + 2];  // [panel column][row] of L_ip and L_jp
        const int b = xcdMapped(blockIdx.x, grid);
        int ti, tj;
        if (j1 > j0)
        {
            const int rows = nt - j0;
            if (b >= (j1 - j0) * rows) return;
            tj = j0 + b / rows;
            ti = j0 + b % rows;
            if (ti < tj) return;
        }
        else
        {
            const int m = nt - j0;
            if (b >= m * (m + 1) / 2) return;
            int row, col;
            lowerTile(b, row, col);
            ti = j0 + row;
            tj = j0 + col;
        }
        const size_t rI = size_t(ti) * kT, rJ = size_t(tj) * kT;
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        const int qi = (w >> 1) * 32, qj = (w & 1) * 32;
        const int li = lane & 15, lk = lane >> 4;
        constexpr int kLR = Mfma<T>::kLaneRow, kIR = Mfma<T>::kItemRow;
        V acc[2][2];
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
            for (int bj = 0; bj < 2; ++bj)
            {
                const T* src = A + (rJ + qj + 16 * bj + li) * size_t(lda) + rI + qi + 16 * bi + kLR * lk;
#pragma unroll
                for (int e = 0; e < 4; ++e) acc[bi][bj][e] = src[kIR * e];
            }
        const int passes = kCount * (kT / kStage);
        for (int pass = 0; pass < passes; ++pass)
        {
            if (pass) __syncthreads();  // every wave is done with the previous slice
            const size_t kc = size_t(kFirst) * kT + size_t(pass) * kStage;
#pragma unroll
            for (int i = 0; i < kStage / 4; ++i)
            {
                const int c = w + 4 * i;
                const T* colp = A + (kc + c) * size_t(lda);
                sa[c][lane] = colp[rI + lane];
                sb[c][lane] = colp[rJ + lane];
            }
            __syncthreads();
#pragma unroll
            for (int kk = 0; kk < kStage; kk += 4)
            {
                const T a0 = -sa[kk + lk][qi + li], a1 = -sa[kk + lk][qi + 16 + li];
                const T b0 = sb[kk + lk][qj + li], b1 = sb[kk + lk][qj + 16 + li];
                acc[0][0] = Mfma<T>::op(a0, b0, acc[0][0]);
                acc[0][1] = Mfma<T>::op(a0, b1, acc[0][1]);
                acc[1][0] = Mfma<T>::op(a1, b0, acc[1][0]);
                acc[1][1] = Mfma<T>::op(a1, b1, acc[1][1]);
            }
        }
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
            for (int bj = 0; bj < 2; ++bj)
            {
                T* dst = A + (rJ + qj + 16 * bj + li) * size_t(lda) + rI + qi + 16 * bi + kLR * lk;
#pragma unroll
                for (int e = 0; e < 4; ++e) dst[kIR * e] = acc[bi][bj][e];
            }
    }

    // ---------------------------------------------------------------------------------------------
    // Triangular solves in one launch. Tiles complete strictly in sweep order (a tile waits for every tile
    // before it), so one 64-bit progress word per sweep, (epoch << 32) | tiles done, tells a waiting tile how
    // many predecessors it may use; it polls only when it has caught up with that frontier.
    __device__ __forceinline__ int awaitProgress(const unsigned long long* word, unsigned epoch, int need,
                                                 int spinLimit, bool& ok)
    {
        int spins = 0;
        unsigned long long v = __hip_atomic_load(word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        while (unsigned(v >> 32) != epoch || int(v & 0xffffffffu) < need)
        {
            if (++spins > spinLimit)
            {
                ok = false;
                return need;
            }
            __builtin_amdgcn_s_sleep(1);
            v = __hip_atomic_load(word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        return int(v & 0xffffffffu);
    }

    __device__ __forceinline__ void publishProgress(unsigned long long* word, unsigned epoch, int done)
    {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_store(word, (static_cast<unsigned long long>(epoch) << 32) | unsigned(done), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }

    constexpr int kSweepChunk = 4;  // off-diagonal tiles whose loads a thread issues together

    // L y = b (x holds b on entry, y on exit). Tile i needs y_k of every k < i.
    template <typename T>
    __global__ void __launch_bounds__(256) k_dense_lower(const T* A, int lda, int nt, T* x, int32_t* words,
                                                         unsigned long long* progress, int spinLimit)
    {
        __shared__ int sTile;
        __shared__ T part[4][kT];
        const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
        if (threadIdx.x == 0) sTile = atomicAdd(words + kDenseTicketLower, 1);
        __syncthreads();
        const int i = sTile;
        const unsigned epoch = unsigned(words[kDenseEpoch]);
        const size_t ri = size_t(i) * kT;
        // reciprocal of this lane's diagonal entry, off the serial chain of the tile solve
        const T dinv = q == 0 ? recip(A[(ri + lane) * size_t(lda) + ri + lane]) : T(0);
        T acc = T(0);
        bool ok = true;
        int avail = 0;
        for (int k0 = 0; k0 < i; k0 += kSweepChunk)
        {
            const int kn = min(kSweepChunk, i - k0);
            T l[kSweepChunk][16];
#pragma unroll
            for (int u = 0; u < kSweepChunk; ++u)
                if (u < kn)
                {
                    const size_t kc = size_t(k0 + u) * kT + q * 16;
#pragma unroll
                    for (int c = 0; c < 16; ++c) l[u][c] = A[(kc + c) * size_t(lda) + ri + lane];
                }
            if (avail < k0 + kn && ok) avail = awaitProgress(progress, epoch, k0 + kn, spinLimit, ok);
#pragma unroll
            for (int u = 0; u < kSweepChunk; ++u)
                if (u < kn)
                {
                    const size_t kc = size_t(k0 + u) * kT + q * 16;
#pragma unroll
                    for (int c = 0; c < 16; ++c) acc += l[u][c] * x[kc + c];
                }
        }
        part[q][lane] = acc;
        __syncthreads();
        if (q == 0)
        {
            T rhs = x[ri + lane] - ((part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane]));
            T lrow[kT];
#pragma unroll
            for (int j = 0; j < kT; ++j) lrow[j] = A[(ri + j) * size_t(lda) + ri + lane];
            T y = T(0);
#pragma unroll
            for (int j = 0; j < kT; ++j)
            {
                const T yj = readLane(rhs, j) * readLane(dinv, j);
                if (lane == j) y = yj;
                rhs -= lrow[j] * yj;
            }
            x[ri + lane] = y;
            if (lane == 0) publishProgress(progress, epoch, i + 1);
        }
        if (!ok && threadIdx.x == 0) atomicOr(words + kDenseTimeout, 1);
    }

    // L^T x = y in place. Workgroup t takes tile i = nt - 1 - t; tile i needs x_k of every k > i, which are
    // published in the order nt - 1, nt - 2, ... (progress counts the tiles done from the bottom).
    template <typename T>
    __global__ void __launch_bounds__(256) k_dense_upper(const T* A, int lda, int nt, T* x, int32_t* words,
                                                         unsigned long long* progress, int spinLimit)
    {
        __shared__ int sTile;
        __shared__ T part[4][kT];
        const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
        if (threadIdx.x == 0) sTile = atomicAdd(words + kDenseTicketUpper, 1);
        __syncthreads();
        const int t = sTile, i = nt - 1 - t;
        const unsigned epoch = unsigned(words[kDenseEpoch]);
        const size_t ri = size_t(i) * kT;
        const T dinv = q == 0 ? recip(A[(ri + lane) * size_t(lda) + ri + lane]) : T(0);
        T acc = T(0);
        bool ok = true;
        int avail = 0;
        // tile k = nt - 1 - u is the u-th published; this tile needs u < t
        for (int u0 = 0; u0 < t; u0 += kSweepChunk)
        {
            const int un = min(kSweepChunk, t - u0);
            T l[kSweepChunk][16];
#pragma unroll
            for (int v = 0; v < kSweepChunk; ++v)
                if (v < un)
                {
                    // (L_ki^T)(lane, c) = L(kc + c, ri + lane): 16 consecutive entries of column ri + lane
                    const size_t kc = size_t(nt - 1 - (u0 + v)) * kT + q * 16;
#pragma unroll
                    for (int c = 0; c < 16; ++c) l[v][c] = A[(ri + lane) * size_t(lda) + kc + c];
                }
            if (avail < u0 + un && ok) avail = awaitProgress(progress, epoch, u0 + un, spinLimit, ok);
#pragma unroll
            for (int v = 0; v < kSweepChunk; ++v)
                if (v < un)
                {
                    const size_t kc = size_t(nt - 1 - (u0 + v)) * kT + q * 16;
#pragma unroll
                    for (int c = 0; c < 16; ++c) acc += l[v][c] * x[kc + c];
                }
        }
        part[q][lane] = acc;
        __syncthreads();
        if (q == 0)
        {
            T rhs = x[ri + lane] - ((part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane]));
            T lcol[kT];  // lcol[j] = L(ri + j, ri + lane) = (L^T)(lane, j)
#pragma unroll
            for (int j = 0; j < kT; ++j) lcol[j] = A[(ri + lane) * size_t(lda) + ri + j];
            T y = T(0);
#pragma unroll
            for (int j = kT - 1; j >= 0; --j)
            {
                const T yj = readLane(rhs, j) * readLane(dinv, j);
                if (lane == j) y = yj;
                rhs -= lcol[j] * yj;
            }
            x[ri + lane] = y;
            if (lane == 0) publishProgress(progress, epoch, t + 1);
        }
        if (!ok && threadIdx.x == 0) atomicOr(words + kDenseTimeout, 1);
    }

    // ---------------------------------------------------------------------------------------------
    // Refinement helpers (FP64): r = b - S x with the block-sparse rows of S, and vector conversions.
    template <typename TW>
    __global__ void k_dense_residual(Problem P, const TW* S, const double* b, const double* x, double* r,
                                     double* rowNorm)
    {
        const int u = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
        if (u >= P.nUnits) return;
        const int lane = threadIdx.x & 63;
        double acc[6] = {0, 0, 0, 0, 0, 0};
        for (int k = P.rowPtr[u] + lane; k < P.rowPtr[u + 1]; k += kWave)
        {
            const int v = P.colIdx[k];
            const TW* blk = S + size_t(k) * 36;
            double xv[6];
#pragma unroll
            for (int j = 0; j < 6; ++j) xv[j] = x[6 * v + j];
#pragma unroll
            for (int a = 0; a < 6; ++a)
#pragma unroll
                for (int j = 0; j < 6; ++j) acc[a] = fma(double(blk[a * 6 + j]), xv[j], acc[a]);
        }
        double nrm = 0.0;
#pragma unroll
        for (int a = 0; a < 6; ++a)
        {
            const double s = waveSum(acc[a]);
            const double ra = b[6 * u + a] - s;
            nrm += ra * ra;
            if (lane == a) r[6 * u + a] = ra;
        }
        if (lane == 0) rowNorm[u] = nrm;
    }

    // out[0] = sum of v[0..n) (of v^2 with square) in a fixed order (one workgroup)
    __global__ void __launch_bounds__(256) k_dense_sum(const double* v, int n, int square, double* out)
    {
        __shared__ double part[4];
        double s = 0.0;
        for (int i = threadIdx.x; i < n; i += 256) s += square ? v[i] * v[i] : v[i];
        s = waveSum(s);
        if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
        __syncthreads();
        if (threadIdx.x == 0) out[0] = (part[0] + part[1]) + (part[2] + part[3]);
    }

    template <typename TA>
    __global__ void k_dense_to(const double* src, int n, int npad, TA* dst)
    {
        const int i = blockIdx.x * blockDim.x + threadIdx.x;
        if (i < npad) dst[i] = i < n ? TA(src[i]) : TA(0);
    }

    // x += c (first n entries); with overwrite, x = c
    template <typename TA>
    __global__ void k_dense_accumulate(double* x, const TA* c, int n, int overwrite)
    {
        const int i = blockIdx.x * blockDim.x + threadIdx.x;
        if (i < n) x[i] = overwrite ? double(c[i]) : x[i] + double(c[i]);
    }

    template <typename TO>
    __global__ void k_dense_store(const double* src, int n, TO* dst)
    {
        const int i = blockIdx.x * blockDim.x + threadIdx.x;
        if (i < n) dst[i] = TO(src[i]);
    }

    // ---------------------------------------------------------------------------------------------
    namespace
    {
        int blocks(long n, int per) { return int((n + per - 1) / per); }
        int roundXcd(int n) { return (n + kXcds - 1) / kXcds * kXcds; }
        constexpr int kSpinLimit = 1 << 20;  // ~1 s of polling before a solve gives up on a tile (never reached)
    }

    template <typename TA>
    template <typename TW>
    void DenseLauncher<TA>::gather(const Problem& P, const TW* S, TA* A, int nt, hipStream_t s)
    {
        const int tiles = nt * (nt + 1) / 2;
        const int grid = roundXcd(tiles);
        k_dense_gather<TW, TA><<<grid, 256, 0, s>>>(P, S, A, nt * kT, nt, grid);
    }

    // Blocked right-looking factorisation: blocks of kDenseBlock tile columns; inside a block every panel step
    // is followed by the update of the block's remaining columns only, and the trailing matrix is updated once
    // per block with the block's kDenseBlock * 64 columns (a quarter of the passes over the trailing matrix of
    // one update per tile column).
    template <typename TA>
    void DenseLauncher<TA>::begin(const DenseState& st, hipStream_t s)
    {
        k_dense_begin<<<1, 64, 0, s>>>(st.words, 1);
    }

    template <typename TA>
    void DenseLauncher<TA>::factor(TA* A, int nt, const DenseState& st, hipStream_t s, bool reset)
    {
        const int lda = nt * kT;
        k_dense_begin<<<1, 64, 0, s>>>(st.words, reset ? 1 : 0);
        for (int kb = 0; kb < nt; kb += kDenseBlock)
        {
            const int ke = std::min(nt, kb + kDenseBlock);
            for (int k = kb; k < ke; ++k)
            {
                k_dense_panel<TA><<<nt - k, 64 * kPanelWaves, 0, s>>>(A, lda, k, st.words);
                if (k + 1 < ke)
                {
                    const int grid = roundXcd((ke - k - 1) * (nt - k - 1));
                    k_dense_update<TA><<<grid, 256, 0, s>>>(A, lda, k, 1, k + 1, ke, nt, grid);
                }
            }
            const int m = nt - ke;
            if (m > 0)
            {
                const int grid = roundXcd(m * (m + 1) / 2);
                k_dense_update<TA><<<grid, 256, 0, s>>>(A, lda, kb, ke - kb, ke, 0, nt, grid);
            }
        }
    }

    template <typename TA>
    void DenseLauncher<TA>::solve(const TA* A, int nt, TA* x, const DenseState& st, hipStream_t s)
    {
        const int lda = nt * kT;
        k_dense_begin<<<1, 64, 0, s>>>(st.words, 0);
        k_dense_lower<TA><<<nt, 256, 0, s>>>(A, lda, nt, x, st.words, st.progress, kSpinLimit);
        k_dense_upper<TA><<<nt, 256, 0, s>>>(A, lda, nt, x, st.words, st.progress + 1, kSpinLimit);
    }

    template <typename TA>
    template <typename TW>
    void DenseLauncher<TA>::residual(const Problem& P, const TW* S, const double* b, const double* x, double* r,
                                     double* rowNorm, double* norm, hipStream_t s)
    {
        k_dense_residual<TW><<<blocks(P.nUnits, 4), 256, 0, s>>>(P, S, b, x, r, rowNorm);
        k_dense_sum<<<1, 256, 0, s>>>(rowNorm, P.nUnits, 0, norm);
    }

    template <typename TA>
    void DenseLauncher<TA>::norm2(const double* v, int n, double* out, hipStream_t s)
    {
        k_dense_sum<<<1, 256, 0, s>>>(v, n, 1, out);
    }

    template <typename TA>
    void DenseLauncher<TA>::toDense(const double* src, int n, int npad, TA* dst, hipStream_t s)
    {
        k_dense_to<TA><<<blocks(npad, 256), 256, 0, s>>>(src, n, npad, dst);
    }

    template <typename TA>
    void DenseLauncher<TA>::accumulate(double* x, const TA* c, int n, bool overwrite, hipStream_t s)
    {
        k_dense_accumulate<TA><<<blocks(n, 256), 256, 0, s>>>(x, c, n, overwrite ? 1 : 0);
    }

    template <typename TA>
    template <typename TO>
    void DenseLauncher<TA>::store(const double* src, int n, TO* dst, hipStream_t s)
    {
        k_dense_store<TO><<<blocks(n, 256), 256, 0, s>>>(src, n, dst);
    }

    template struct DenseLauncher<float>;
    template struct DenseLauncher<double>;
    template void DenseLauncher<float>::gather<float>(const Problem&, const float*, float*, int, hipStream_t);
    template void DenseLauncher<float>::gather<double>(const Problem&, const double*, float*, int, hipStream_t);
    template void DenseLauncher<double>::gather<float>(const Problem&, const float*, double*, int, hipStream_t);
    template void DenseLauncher<double>::gather<double>(const Problem&, const double*, double*, int, hipStream_t);
    template void DenseLauncher<float>::residual<float>(const Problem&, const float*, const double*, const double*,
                                                        double*, double*, double*, hipStream_t);
    template void DenseLauncher<float>::residual<double>(const Problem&, const double*, const double*, const double*,
                                                         double*, double*, double*, hipStream_t);
    template void DenseLauncher<double>::residual<float>(const Problem&, const float*, const double*, const double*,
                                                         double*, double*, double*, hipStream_t);
    template void DenseLauncher<double>::residual<double>(const Problem&, const double*, const double*,
                                                          const double*, double*, double*, double*, hipStream_t);
    template void DenseLauncher<float>::store<float>(const double*, int, float*, hipStream_t);
    template void DenseLauncher<float>::store<double>(const double*, int, double*, hipStream_t);
    template void DenseLauncher<double>::store<float>(const double*, int, float*, hipStream_t);
    template void DenseLauncher<double>::store<double>(const double*, int, double*, hipStream_t);
}  // namespace pbk
