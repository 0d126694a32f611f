// pb_cluster.hip - PCG with a cluster block-Jacobi preconditioner (plaMatrixPreconditionerClusterSize).
//
// PlaMatrix's GPU route preconditions its PCG with the inverses of the diagonal blocks of clusters of consecutive
// primary blocks (BundleAdjustPlaMatrix.cpp:330 passes the cluster size; options.cpp:393-394 bounds it to 1..16).
// The fused PCG iteration (pb_kernels.hip, k_pcg_iter) keeps one 6 x 6 block per unit, which is that
// preconditioner exactly when every primary block is one unit. Everywhere else (clusters of several primary
// blocks, or an intrinsic group of more than six refined slots, whose primary block spans two or three units)
// this file's PCG runs:
//   k_cl_factor   one workgroup per cluster: gather the d x d diagonal block (d = 6 units) from the block-sparse
//                 rows, Cholesky in place, L^-1 column by column, M = L^-T L^-1
//   per iteration four launches over the 6 nUnits rows (thread per row):
//     k_cl_spmv   q = S p                                       partials (p, q)
//     k_cl_step   alpha = (r, z) / (p, q); x += alpha p, r -= alpha q
//     k_cl_prec   z = M r (the row's cluster, dense)             partials (r, z), (r, r)
//     k_cl_dir    beta = (r, z)_new / (r, z)_old, p = z + beta p; the stop rules
// Every launch sums the previous launch's per-block partials in the same fixed order, so all blocks derive
// bitwise the same scalars, and repeated solves are bitwise identical.
#include "pb_common.h"
#include "pb_kernels.h"

#include <hip/hip_runtime.h>

#include <algorithm>

namespace pbk
{
    namespace
    {
        constexpr int kClThreads = 256;

        int clBlocks(long n) { return int((n + kClThreads - 1) / kClThreads); }

        // block total of v (every thread receives it)
        __device__ __forceinline__ double clBlockSum(double v, double* lds)
        {
            v = waveSum(v);
            const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
            __syncthreads();
            if (lane == 0) lds[wv] = v;
            __syncthreads();
            double s = 0.0;
            for (int i = 0; i < int(blockDim.x >> 6); ++i) s += lds[i];
            return s;
        }

        // sum of n per-block partials in a fixed order (wave 0, lane-strided, then the butterfly), broadcast
        __device__ __forceinline__ double clSumParts(const double* part, int n, double* slot)
        {
            __syncthreads();
            if ((threadIdx.x >> 6) == 0)
            {
                double v = 0.0;
                for (int i = threadIdx.x; i < n; i += kWave) v += part[i];
                v = waveSum(v);
                if (threadIdx.x == 0) *slot = v;
            }
            __syncthreads();
            return *slot;
        }
    }  // namespace

    // ---------------------------------------------------------------------------------------------
    template <typename T>
    __global__ void __launch_bounds__(kClThreads) k_cl_factor(Problem P, const T* S, ClusterDev C, int32_t* flags)
    {
        const int cl = blockIdx.x;
        const int u0 = C.unitBegin[cl], u1 = C.unitBegin[cl + 1];
        const int d = 6 * (u1 - u0);
        double* A = C.inv + C.matOff[cl];  // row-major d x d
        double* Y = C.work + C.matOff[cl];
        for (int i = threadIdx.x; i < d * d; i += blockDim.x) A[i] = 0.0;
        __syncthreads();
        for (int u = u0; u < u1; ++u)
        {
            const int kb = P.rowPtr[u], ke = P.rowPtr[u + 1];
            for (int idx = threadIdx.x; idx < (ke - kb) * 36; idx += blockDim.x)
            {
                const int k = kb + idx / 36, e = idx % 36;
                const int v = P.colIdx[k];
                if (v < u0 || v >= u1) continue;
                A[((u - u0) * 6 + e / 6) * d + (v - u0) * 6 + e % 6] = double(S[size_t(k) * 36 + e]);
            }
        }
        // Cholesky A = L L^T in place (lower triangle), column by column
        __shared__ double piv;
        __shared__ int bad;
        if (threadIdx.x == 0) bad = 0;
        for (int j = 0; j < d; ++j)
        {
            __syncthreads();
            if (threadIdx.x == 0)
            {
                double a = A[j * d + j];
                if (!(a > 0.0) || !pbdev::finite(a))
                {
                    bad = 1;
                    a = 1.0;
                }
                piv = sqrt(a);
            }
            __syncthreads();
            const double l = piv;
            for (int i = j + threadIdx.x; i < d; i += blockDim.x) A[i * d + j] = i == j ? l : A[i * d + j] / l;
            __syncthreads();
            const int m = d - j - 1;
            for (int idx = threadIdx.x; idx < m * m; idx += blockDim.x)
            {
                const int i = j + 1 + idx / m, k = j + 1 + idx % m;
                if (k <= i) A[i * d + k] -= A[i * d + j] * A[k * d + j];
            }
        }
        __syncthreads();
        if (threadIdx.x == 0 && bad) setFlag(flags, kFlagNotSpd);
        // Y = L^-1 (lower), one column per thread
        for (int col = threadIdx.x; col < d; col += blockDim.x)
            for (int i = 0; i < d; ++i)
            {
                if (i < col)
                {
                    Y[i * d + col] = 0.0;
                    continue;
                }
                double s = i == col ? 1.0 : 0.0;
                for (int k = col; k < i; ++k) s -= A[i * d + k] * Y[k * d + col];
                Y[i * d + col] = s / A[i * d + i];
            }
        __syncthreads();
        // M = Y^T Y (symmetric) over A
        for (int idx = threadIdx.x; idx < d * d; idx += blockDim.x)
        {
            const int i = idx / d, j = idx % d;
            double s = 0.0;
            for (int k = max(i, j); k < d; ++k) s += Y[k * d + i] * Y[k * d + j];
            A[idx] = s;
        }
    }

    // z = M r for the rows of one launch; partials (r, z) into rz and (r, r) into rr
    template <typename T>
    __device__ __forceinline__ void clPrecondition(const Problem& P, const Work<T>& W, const ClusterDev& C, double* rz,
                                                   double* rr)
    {
        __shared__ double lds[8];
        const int g = blockIdx.x * blockDim.x + threadIdx.x;
        const T* r = W.cgR[0];
        T* z = W.cgU[0];
        double a = 0.0, b = 0.0;
        if (g < 6 * P.nUnits)
        {
            const int cl = C.unitCluster[g / 6];
            const int u0 = C.unitBegin[cl], d = 6 * (C.unitBegin[cl + 1] - u0);
            const int i = g - 6 * u0;
            const double* M = C.inv + C.matOff[cl];
            const T* rc = r + 6 * u0;
            double s = 0.0;
            for (int j = 0; j < d; ++j) s += M[size_t(j) * d + i] * double(rc[j]);  // column i = row i
            z[g] = T(s);
            const double rg = double(r[g]);
            a = rg * double(T(s));
            b = rg * rg;
        }
        const double sa = clBlockSum(a, lds);
        const double sb = clBlockSum(b, lds);
        if (threadIdx.x == 0)
        {
            rz[blockIdx.x] = sa;
            rr[blockIdx.x] = sb;
        }
    }

    template <typename T>
    __global__ void __launch_bounds__(kClThreads) k_cl_init(Problem P, Work<T> W, ClusterDev C)
    {
        const int g = blockIdx.x * blockDim.x + threadIdx.x;
        if (g < 6 * P.nUnits)
        {
            W.x[g] = T(0);
            W.cgR[0][g] = T(W.rhs[g]);
        }
        if (g == 0)
        {
            for (int i = 0; i < kIflagCount; ++i) W.iflags[i] = 0;
            W.scalars[kScalarQ0] = 0.0;
            W.scalars[kScalarQ1] = 0.0;
        }
    }

    template <typename T>
    __global__ void __launch_bounds__(kClThreads) k_cl_init_prec(Problem P, Work<T> W, ClusterDev C)
    {
        const int nb = gridDim.x;
        clPrecondition(P, W, C, W.cgPart[1], W.cgPart[0] + nb);
    }

    // p = z; the threshold from |rhs| (= |r_0|) and the stop test at iteration 0
    template <typename T>
    __global__ void __launch_bounds__(kClThreads) k_cl_init_dir(Problem P, Work<T> W, double relTol, double absTol)
    {
        __shared__ double slot;
        const int nb = gridDim.x;
        const double bb = clSumParts(W.cgPart[0] + nb, nb, &slot);
        const int g = blockIdx.x * blockDim.x + threadIdx.x;
        if (g < 6 * P.nUnits) W.p[g] = W.cgU[0][g];
        if (g == 0)
        {
            const double bn = sqrt(bb);
            const double th = fmax(relTol * bn, absTol);
            W.scalars[kScalarBnorm] = bn;
            W.scalars[kScalarThreshold] = th;
            W.scalars[kScalarThreshold2] = th * th;
            W.scalars[kScalarRnorm] = bn;
            if (bb <= th * th) W.iflags[kIflagDone0] = 1;
        }
    }

    template <typename T>
    __global__ void __launch_bounds__(kClThreads) k_cl_spmv(Problem P, Work<T> W)
    {
        if (W.iflags[kIflagDone0]) return;
        __shared__ double lds[8];
        const int g = blockIdx.x * blockDim.x + threadIdx.x;
        double pq = 0.0;
        if (g < 6 * P.nUnits)
        {
            const int u = g / 6, a = g % 6;
            double s = 0.0;
            for (int k = P.rowPtr[u]; k < P.rowPtr[u + 1]; ++k)
            {
                const T* blk = W.S + size_t(k) * 36 + a * 6;
                const T* pv = W.p + 6 * P.colIdx[k];
#pragma unroll
                for (int b = 0; b < 6; ++b) s += double(blk[b]) * double(pv[b]);
            }
            W.cgW[0][g] = T(s);
            pq = double(W.p[g]) * double(T(s));
        }
        const double t = clBlockSum(pq, lds);
        if (threadIdx.x == 0) W.cgPart[0][blockIdx.x] = t;
    }

    template <typename T>
    __global__ void __launch_bounds__(kClThreads) k_cl_step(Problem P, Work<T> W, int it)
    {
        if (W.iflags[kIflagDone0]) return;
        __shared__ double slot;
        const int nb = gridDim.x;
        const double pq = clSumParts(W.cgPart[0], nb, &slot);
        const double rz = clSumParts(W.cgPart[1] + (it & 1) * nb, nb, &slot);
        if (!(pq > 0.0) || !pbdev::finite(pq))
        {
            if (blockIdx.x == 0 && threadIdx.x == 0)
            {
                W.iflags[kIflagBreakdown] = 1;
                W.iflags[kIflagDone0] = 1;
            }
            return;
        }
        const double alpha = rz / pq;
        const int g = blockIdx.x * blockDim.x + threadIdx.x;
        if (g < 6 * P.nUnits)
        {
            W.x[g] = T(double(W.x[g]) + alpha * double(W.p[g]));
            W.cgR[0][g] = T(double(W.cgR[0][g]) - alpha * double(W.cgW[0][g]));
        }
    }

    template <typename T>
    __global__ void __launch_bounds__(kClThreads) k_cl_prec(Problem P, Work<T> W, ClusterDev C, int it)
    {
        if (W.iflags[kIflagDone0]) return;
        const int nb = gridDim.x;
        clPrecondition(P, W, C, W.cgPart[1] + ((it + 1) & 1) * nb, W.cgPart[0] + nb);
    }

    template <typename T>
    __global__ void __launch_bounds__(kClThreads) k_cl_dir(Problem P, Work<T> W, int it, int maxIt, double modelEta)
    {
        if (W.iflags[kIflagDone0]) return;
        __shared__ double slot;
        const int nb = gridDim.x;
        const double rzNew = clSumParts(W.cgPart[1] + ((it + 1) & 1) * nb, nb, &slot);
        const double rzOld = clSumParts(W.cgPart[1] + (it & 1) * nb, nb, &slot);
        const double beta = rzNew / rzOld;
        const int g = blockIdx.x * blockDim.x + threadIdx.x;
        if (g < 6 * P.nUnits) W.p[g] = T(double(W.cgU[0][g]) + beta * double(W.p[g]));
        if (blockIdx.x != 0) return;
        const double rr = clSumParts(W.cgPart[0] + nb, nb, &slot);
        const double pq = clSumParts(W.cgPart[0], nb, &slot);
        if (threadIdx.x != 0) return;
        // quadratic model of CG from x = 0: each step lowers it by alpha (r, z) / 2
        const double qOld = W.scalars[kScalarQ0 + (it & 1)];
        const double qNew = qOld - 0.5 * (rzOld / pq) * rzOld;
        W.scalars[kScalarQ0 + ((it + 1) & 1)] = qNew;
        W.scalars[kScalarRnorm] = sqrt(rr);
        W.iflags[kIflagIters] = it + 1;
        bool stop = rr <= W.scalars[kScalarThreshold2] || it + 1 >= maxIt;
        if (modelEta > 0.0 && it + 1 >= kCgModelMinIterations && qNew < 0.0 &&
            double(it + 1) * (qNew - qOld) / qNew < modelEta)
        {
            W.iflags[kIflagModelStop] = 1;
            stop = true;
        }
        if (stop) W.iflags[kIflagDone0] = 1;
    }

    // ---------------------------------------------------------------------------------------------
    template <typename T>
    void ClusterLauncher<T>::factor(const Problem& P, const T* S, const ClusterDev& C, int32_t* flags, hipStream_t s)
    {
        if (C.n > 0) k_cl_factor<T><<<C.n, kClThreads, 0, s>>>(P, S, C, flags);
    }

    template <typename T>
    void ClusterLauncher<T>::init(const Problem& P, const Work<T>& W, const ClusterDev& C, double relTol,
                                  double absTol, hipStream_t s)
    {
        const int nb = std::max(1, clBlocks(6L * P.nUnits));
        k_cl_init<T><<<nb, kClThreads, 0, s>>>(P, W, C);
        k_cl_init_prec<T><<<nb, kClThreads, 0, s>>>(P, W, C);
        k_cl_init_dir<T><<<nb, kClThreads, 0, s>>>(P, W, relTol, absTol);
    }

    template <typename T>
    void ClusterLauncher<T>::iteration(const Problem& P, const Work<T>& W, const ClusterDev& C, int it, int maxIt,
                                       double modelEta, hipStream_t s)
    {
        const int nb = std::max(1, clBlocks(6L * P.nUnits));
        k_cl_spmv<T><<<nb, kClThreads, 0, s>>>(P, W);
        k_cl_step<T><<<nb, kClThreads, 0, s>>>(P, W, it);
        k_cl_prec<T><<<nb, kClThreads, 0, s>>>(P, W, C, it);
        k_cl_dir<T><<<nb, kClThreads, 0, s>>>(P, W, it, maxIt, modelEta);
    }

    template struct ClusterLauncher<float>;
    template struct ClusterLauncher<double>;
}  // namespace pbk
