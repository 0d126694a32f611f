// pb_tile.h - 64 x 64 tile routines in LDS for one 256-thread workgroup (gfx950, wave64), FP64 or FP32: tile
// Cholesky, triangular inverse and MFMA tile products, shared by the band route (pb_band.hip) and the batched
// window factorisation (pb_batch.hip).
#pragma once

#include "pb_common.h"
#include "pb_kernels.h"

#include <hip/hip_runtime.h>

namespace pbk
{
    namespace tile
    {
        constexpr int kT = kDenseTile;
        constexpr int kLd = kT + 1;  // LDS tile leading dimension: buf[column][row]

        template <typename T>
        struct BandMfma;
        // Result element e of lane l is row kLaneRow * (l >> 4) + kItemRow * e, column l & 15 of the 16 x 16 block
        // (the two instructions lay their results out differently; measured on gfx950).
        template <>
        struct BandMfma<double>
        {
            static constexpr int kLaneRow = 1, kItemRow = 4;
            typedef double V __attribute__((ext_vector_type(4)));
            static __device__ __forceinline__ V op(double a, double b, V c)
            {
                return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
            }
        };
        template <>
        struct BandMfma<float>
        {
            static constexpr int kLaneRow = 4, kItemRow = 1;
            typedef float V __attribute__((ext_vector_type(4)));
            static __device__ __forceinline__ V op(float a, float b, V c)
            {
                return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
            }
        };

        __device__ __forceinline__ double bandReadLane(double v, int j)
        {
            const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(v));
            const unsigned lo = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(b & 0xffffffffu), j));
            const unsigned hi = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(b >> 32), j));
            return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
        }
        __device__ __forceinline__ float bandReadLane(float v, int j)
        {
            return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
        }

        // 1/p from the hardware reciprocal refined by Newton's iteration (two steps in FP64, one in FP32)
        __device__ __forceinline__ double bandRecip(double p)
        {
            double r = __builtin_amdgcn_rcp(p);
            r = __builtin_fma(r, __builtin_fma(-p, r, 1.0), r);
            return __builtin_fma(r, __builtin_fma(-p, r, 1.0), r);
        }
        __device__ __forceinline__ float bandRecip(float p)
        {
            const float r = __builtin_amdgcn_rcpf(p);
            return __builtin_fmaf(r, __builtin_fmaf(-p, r, 1.0f), r);
        }

        // 1/sqrt(p) from the hardware estimate refined by Newton's iteration y <- y + y (1 - p y^2) / 2 (two steps in
        // FP64, one in FP32)
        __device__ __forceinline__ double bandRsqrt(double p)
        {
            double y = __builtin_amdgcn_rsq(p);
            y = __builtin_fma(0.5 * y, __builtin_fma(-(p * y), y, 1.0), y);
            return __builtin_fma(0.5 * y, __builtin_fma(-(p * y), y, 1.0), y);
        }
        __device__ __forceinline__ float bandRsqrt(float p)
        {
            const float y = __builtin_amdgcn_rsqf(p);
            return __builtin_fmaf(0.5f * y, __builtin_fmaf(-(p * y), y, 1.0f), y);
        }

        __device__ __forceinline__ double bandFma(double a, double b, double c) { return __builtin_fma(a, b, c); }
        __device__ __forceinline__ float bandFma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

        // Cholesky of the 64 x 64 tile buf (lower triangle read; L written with zeros above the diagonal) by the
        // whole workgroup, in four panels of 16 columns: wave 0 factors the panel with lane r holding row r's 16
        // entries in registers (column j published in LDS, every lane subtracts L(r, j) L(c, j) from its entries
        // c > j), then all waves apply the panel to the trailing columns (rank-16 update in LDS).
        constexpr int kPanel = 16;
        template <typename T>
        __device__ void cholTile(T (*buf)[kLd], T* col, bool& bad)
        {
            const int lane = threadIdx.x & 63;
            for (int c0 = 0; c0 < kT; c0 += kPanel)
            {
                if (threadIdx.x < 64)
                {
                    // the pivot of the next column is formed by its own lane from its own value, so the serial
                    // chain is readlane -> 1/sqrt -> scale -> fma -> readlane; the LDS exchange of the column
                    // feeds the other lanes' updates off that chain
                    const int r = lane;
                    T a[kPanel];
#pragma unroll
                    for (int q = 0; q < kPanel; ++q) a[q] = buf[c0 + q][r];
                    T p = bandReadLane(a[0], c0);
#pragma unroll
                    for (int j = 0; j < kPanel; ++j)
                    {
                        const int cj = c0 + j;
                        if (!(p > T(0) && p < T(INFINITY)))
                        {
                            bad = true;
                            p = T(1);
                        }
                        const T rs = bandRsqrt(p);
                        const T l = r > cj ? a[j] * rs : (r == cj ? p * rs : T(0));
                        a[j] = l;
                        col[r] = l;
                        if (j + 1 < kPanel) p = bandReadLane(bandFma(-l, l, a[j + 1]), cj + 1);
                        waveLdsFence();
#pragma unroll
                        for (int q = j + 1; q < kPanel; ++q) a[q] = bandFma(-l, col[c0 + q], a[q]);
                        waveLdsFence();
                    }
#pragma unroll
                    for (int q = 0; q < kPanel; ++q) buf[c0 + q][r] = r >= c0 + q ? a[q] : T(0);
                }
                __syncthreads();
                // trailing columns c >= c0 + 16, rows r >= c0 + 16 (both triangles of the trailing block)
                const int t0 = c0 + kPanel, nt = kT - t0;
                for (int idx = threadIdx.x; idx < nt * nt; idx += blockDim.x)
                {
                    const int r = t0 + idx % nt, c = t0 + idx / nt;
                    T s = T(0);
#pragma unroll
                    for (int q = 0; q < kPanel; ++q) s += buf[c0 + q][r] * buf[c0 + q][c];
                    buf[c][r] -= s;
                }
                __syncthreads();
            }
        }

        // inv = L^-1 of the lower-triangular tile L (zeros above the diagonal) by the whole workgroup, in four row
        // panels of 16: every thread subtracts the finished panels' contribution from its entries of the panel
        // rows (inv(i, j) = [i == j] - sum_{k < i0} L(i, k) inv(k, j)), then wave 0 substitutes down the panel with
        // lane j holding column j's 16 entries in registers.
        template <typename T>
        __device__ void invTile(T (*L)[kLd], T (*inv)[kLd], T* dinv)
        {
            if (threadIdx.x < 64) dinv[threadIdx.x] = bandRecip(L[threadIdx.x][threadIdx.x]);
            for (int i0 = 0; i0 < kT; i0 += kPanel)
            {
                {
                    // thread t: row i = i0 + t % 16 of the panel, columns j = t / 16 + 16 q (q < 4)
                    const int i = i0 + (threadIdx.x & (kPanel - 1)), j0 = threadIdx.x / kPanel;
                    T s[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) s[q] = i == j0 + 16 * q ? T(1) : T(0);
                    for (int k = 0; k < i0; k += 4)
                    {
                        T l[4];
#pragma unroll
                        for (int u = 0; u < 4; ++u) l[u] = L[k + u][i];
#pragma unroll
                        for (int q = 0; q < 4; ++q)
#pragma unroll
                            for (int u = 0; u < 4; ++u) s[q] -= l[u] * inv[j0 + 16 * q][k + u];
                    }
#pragma unroll
                    for (int q = 0; q < 4; ++q) inv[j0 + 16 * q][i] = s[q];
                }
                __syncthreads();
                if (threadIdx.x < 64)
                {
                    // right-looking substitution: once x_p is known it is removed from every later row, so the
                    // serial chain per row is one multiply and one fma
                    const int j = threadIdx.x;
                    T x[kPanel];
#pragma unroll
                    for (int q = 0; q < kPanel; ++q) x[q] = inv[j][i0 + q];
#pragma unroll
                    for (int p = 0; p < kPanel; ++p)
                    {
                        x[p] *= dinv[i0 + p];
#pragma unroll
                        for (int q = p + 1; q < kPanel; ++q) x[q] = bandFma(-L[i0 + p][i0 + q], x[p], x[q]);
                    }
#pragma unroll
                    for (int q = 0; q < kPanel; ++q) inv[j][i0 + q] = x[q];
                }
                __syncthreads();
            }
        }

        // O = (neg ? -1 : 1) A B, or O -= A B with sub, on 64 x 64 LDS tiles; A(r, k) = transA ? a[r][k] : a[k][r],
        // B(k, c) = transB ? b[k][c] : b[c][k]. Wave w forms the 32 x 32 quadrant (w >> 1, w & 1) as 2 x 2 MFMA
        // 16x16x4 blocks (operand A(i, kk) from lane i + 16 kk, B(kk, j) from lane j + 16 kk).
        template <typename T>
        __device__ void tileProduct(T (*o)[kLd], T (*a)[kLd], bool transA, T (*b)[kLd], bool transB, bool sub,
                                    bool neg)
        {
            using V = typename BandMfma<T>::V;
            constexpr int kLR = BandMfma<T>::kLaneRow, kIR = BandMfma<T>::kItemRow;
            const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
            const int qi = (w >> 1) * 32, qj = (w & 1) * 32;
            const int li = lane & 15, lk = lane >> 4;
            V acc[2][2];
#pragma unroll
            for (int bi = 0; bi < 2; ++bi)
#pragma unroll
                for (int bj = 0; bj < 2; ++bj)
#pragma unroll
                    for (int e = 0; e < 4; ++e) acc[bi][bj][e] = T(0);
#pragma unroll 4
            for (int k = 0; k < kT; k += 4)
            {
                const int kk = k + lk;
                T av[2], bv[2];
#pragma unroll
                for (int x = 0; x < 2; ++x)
                {
                    const int r = qi + 16 * x + li, c = qj + 16 * x + li;
                    av[x] = transA ? a[r][kk] : a[kk][r];
                    bv[x] = transB ? b[kk][c] : b[c][kk];
                }
#pragma unroll
                for (int bi = 0; bi < 2; ++bi)
#pragma unroll
                    for (int bj = 0; bj < 2; ++bj) acc[bi][bj] = BandMfma<T>::op(av[bi], bv[bj], acc[bi][bj]);
            }
            __syncthreads();  // every wave has read a and b before o (which may alias neither) is written
#pragma unroll
            for (int bi = 0; bi < 2; ++bi)
#pragma unroll
                for (int bj = 0; bj < 2; ++bj)
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                    {
                        const int r = qi + 16 * bi + kLR * lk + kIR * e, c = qj + 16 * bj + li;
                        T& d = o[c][r];
                        const T v = acc[bi][bj][e];
                        d = sub ? d - v : (neg ? -v : v);
                    }
        }

        // 64 x 64 tile of M (leading dimension ld, from row r0, column c0) into buf[column][row]; every thread
        // issues its 16 loads before it stores (the loads of a workgroup overlap instead of queueing one by one).
        template <typename T>
        __device__ __forceinline__ void loadTile(T (*buf)[kLd], const T* M, int ld, int r0, int c0)
        {
            constexpr int kPer = kT * kT / 256;
            const int r = threadIdx.x & 63, c = threadIdx.x >> 6;
            T v[kPer];
#pragma unroll
            for (int q = 0; q < kPer; ++q) v[q] = M[size_t(c0 + c + 4 * q) * ld + r0 + r];
#pragma unroll
            for (int q = 0; q < kPer; ++q) buf[c + 4 * q][r] = v[q];
        }

        // M(r0 + r, c0 + c) = buf[c][r] (the tile as loaded by loadTile)
        template <typename T>
        __device__ __forceinline__ void storeTile(T* M, int ld, int r0, int c0, T (*buf)[kLd])
        {
            constexpr int kPer = kT * kT / 256;
            const int r = threadIdx.x & 63, c = threadIdx.x >> 6;
#pragma unroll
            for (int q = 0; q < kPer; ++q) M[size_t(c0 + c + 4 * q) * ld + r0 + r] = buf[c + 4 * q][r];
        }

        // G = A B, or G -= A B with sub, for the 64 x 64 tile G of a column-major matrix in global memory (leading
        // dimension ld; G(r, c) at G[c * ld + r]); A and B as in tileProduct. The result goes from the MFMA
        // accumulators straight to memory (the accumulators of a subtraction start from G), so the two LDS operand
        // tiles are all the LDS it needs; callers synchronise before the operands are overwritten.
        template <typename T>
        __device__ void tileProductGlobal(T* G, int ld, T (*a)[kLd], bool transA, T (*b)[kLd], bool transB, bool sub)
        {
            using V = typename BandMfma<T>::V;
            constexpr int kLR = BandMfma<T>::kLaneRow, kIR = BandMfma<T>::kItemRow;
            const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
            const int qi = (w >> 1) * 32, qj = (w & 1) * 32;
            const int li = lane & 15, lk = lane >> 4;
            V acc[2][2];
#pragma unroll
            for (int bi = 0; bi < 2; ++bi)
#pragma unroll
                for (int bj = 0; bj < 2; ++bj)
#pragma unroll
                    for (int e = 0; e < 4; ++e) acc[bi][bj][e] = T(0);
#pragma unroll 4
            for (int k = 0; k < kT; k += 4)
            {
                const int kk = k + lk;
                T av[2], bv[2];
#pragma unroll
                for (int x = 0; x < 2; ++x)
                {
                    const int r = qi + 16 * x + li, c = qj + 16 * x + li;
                    av[x] = transA ? a[r][kk] : a[kk][r];
                    bv[x] = transB ? b[kk][c] : b[c][kk];
                }
#pragma unroll
                for (int bi = 0; bi < 2; ++bi)
#pragma unroll
                    for (int bj = 0; bj < 2; ++bj) acc[bi][bj] = BandMfma<T>::op(av[bi], bv[bj], acc[bi][bj]);
            }
#pragma unroll
            for (int bi = 0; bi < 2; ++bi)
#pragma unroll
                for (int bj = 0; bj < 2; ++bj)
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                    {
                        const int r = qi + 16 * bi + kLR * lk + kIR * e, c = qj + 16 * bj + li;
                        T& d = G[size_t(c) * ld + r];
                        d = sub ? d - acc[bi][bj][e] : acc[bi][bj][e];
                    }
        }

        // M(r0 + r, c0 + c) = buf(c, r) (the transpose of the LDS tile), or 0 when buf is null
        template <typename T>
        __device__ __forceinline__ void storeTransposed(T* M, int ld, int r0, int c0, T (*buf)[kLd])
        {
            constexpr int kPer = kT * kT / 256;
            const int r = threadIdx.x & 63, c = threadIdx.x >> 6;
#pragma unroll
            for (int q = 0; q < kPer; ++q) M[size_t(c0 + c + 4 * q) * ld + r0 + r] = buf ? buf[r][c + 4 * q] : T(0);
        }

    }  // namespace tile
}  // namespace pbk
