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
                // trailing block (rows and columns >= c0 + 16) minus the panel's rank-16 product, in 16 x 16 blocks
                // on or below the diagonal, one block per wave with MFMA: A(i, k) = L(r0 + i, c0 + k),
                // B(k, j) = L(s0 + j, c0 + k). Blocks above the diagonal are not needed: each panel reads its
                // columns on and below the diagonal and writes zeros above it.
                const int t0 = c0 + kPanel, nb = (kT - t0) / kPanel;
                const int li = lane & 15, lk = lane >> 4;
                for (int b = threadIdx.x >> 6; b < nb * (nb + 1) / 2; b += blockDim.x >> 6)
                {
                    int bi = 0;
                    while ((bi + 1) * (bi + 2) / 2 <= b) ++bi;
                    const int r0 = t0 + kPanel * bi, s0 = t0 + kPanel * (b - bi * (bi + 1) / 2);
                    typename BandMfma<T>::V acc = {T(0), T(0), T(0), T(0)};
#pragma unroll
                    for (int k = 0; k < kPanel; k += 4)
                        acc = BandMfma<T>::op(buf[c0 + k + lk][r0 + li], buf[c0 + k + lk][s0 + li], acc);
                    constexpr int kLR = BandMfma<T>::kLaneRow, kIR = BandMfma<T>::kItemRow;
#pragma unroll
                    for (int e = 0; e < 4; ++e) buf[s0 + li][r0 + kLR * lk + kIR * e] -= acc[e];
                }
                __syncthreads();
            }
        }

        // acc += A B for one 16 x 16 block with MFMA 16x16x4 steps over k = k0 .. k1 - 1 (a multiple of 4 apart);
        // A(i, k) = a[ac + k][ar + i], B(k, j) = b[bc + j][br + k] (LDS tiles stored buf[column][row])
        template <typename T>
        __device__ __forceinline__ void mfmaBlock(typename BandMfma<T>::V& acc, T (*a)[kLd], int ar, int ac,
                                                  T (*b)[kLd], int br, int bc, int k0, int k1)
        {
            const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
            for (int k = k0; k < k1; k += 4)
                acc = BandMfma<T>::op(a[ac + k + lk][ar + li], b[bc + li][br + k + lk], acc);
        }

        // o(r0 + i, c0 + j) = (neg ? -1 : 1) acc(i, j) for the 16 x 16 block held in MFMA accumulators
        template <typename T>
        __device__ __forceinline__ void putBlock(T (*o)[kLd], int r0, int c0, const typename BandMfma<T>::V& acc,
                                                 bool neg)
        {
            constexpr int kLR = BandMfma<T>::kLaneRow, kIR = BandMfma<T>::kItemRow;
            const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
#pragma unroll
            for (int e = 0; e < 4; ++e) o[c0 + li][r0 + kLR * lk + kIR * e] = neg ? -acc[e] : acc[e];
        }

        // inv = L^-1 of the lower-triangular tile L (zeros above the diagonal) by the whole workgroup, blockwise in
        // 16 x 16 blocks: wave w inverts diagonal block w by substitution (lane j holds column j; once x_p is known
        // it is removed from every later row, so the serial chain per row is one multiply and one fma); then the
        // off-diagonal block of each 32 x 32 half, X = -D1^-1 (B D0^-1), and the lower-left 32 x 32 block,
        // X = -C^-1 (B A^-1), are MFMA products (the recursive form of a triangular inverse). The blocks above the
        // diagonal are zero.
        template <typename T>
        __device__ void invTile(T (*L)[kLd], T (*inv)[kLd], T* dinv)
        {
            using V = typename BandMfma<T>::V;
            const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
            const int li = lane & 15, lk = lane >> 4;
            if (threadIdx.x < 64) dinv[threadIdx.x] = bandRecip(L[threadIdx.x][threadIdx.x]);
            for (int idx = threadIdx.x; idx < kT * kT; idx += blockDim.x)
            {
                const int r = idx & 63, c = idx >> 6;
                if ((c >> 4) > (r >> 4)) inv[c][r] = T(0);
            }
            __syncthreads();
            {
                const int b0 = 16 * w, j = li;
                T x[16];
#pragma unroll
                for (int q = 0; q < 16; ++q) x[q] = q == j ? T(1) : T(0);
#pragma unroll
                for (int p = 0; p < 16; ++p)
                {
                    x[p] *= dinv[b0 + p];
#pragma unroll
                    for (int q = p + 1; q < 16; ++q) x[q] = bandFma(-L[b0 + p][b0 + q], x[p], x[q]);
                }
                if (lk == 0)
                {
#pragma unroll
                    for (int q = 0; q < 16; ++q) inv[b0 + j][b0 + q] = x[q];
                }
            }
            __syncthreads();
            V acc;
            if (w < 2)
            {
                // wave w: rows r0 .. r0 + 15, columns c0 .. c0 + 15 of the 32 x 32 half w; the product B D0^-1
                // is parked in the block it becomes and read back by the same wave
                const int r0 = 32 * w + 16, c0 = 32 * w;
#pragma unroll
                for (int e = 0; e < 4; ++e) acc[e] = T(0);
                mfmaBlock(acc, L, r0, c0, inv, c0, c0, 0, 16);
                putBlock(inv, r0, c0, acc, false);
                waveLdsFence();
#pragma unroll
                for (int e = 0; e < 4; ++e) acc[e] = T(0);
                mfmaBlock(acc, inv, r0, r0, inv, r0, c0, 0, 16);
                waveLdsFence();
                putBlock(inv, r0, c0, acc, true);
            }
            __syncthreads();
            // wave w: block (bi, bj) of the lower-left 32 x 32 block; A^-1 is zero above its diagonal blocks and
            // C^-1 likewise, which bounds the k ranges
            const int bi = w >> 1, bj = w & 1;
            const int r0 = 32 + 16 * bi, c0 = 16 * bj;
#pragma unroll
            for (int e = 0; e < 4; ++e) acc[e] = T(0);
            mfmaBlock(acc, L, r0, 0, inv, 0, c0, 16 * bj, 32);
            putBlock(inv, r0, c0, acc, false);
            __syncthreads();
#pragma unroll
            for (int e = 0; e < 4; ++e) acc[e] = T(0);
            mfmaBlock(acc, inv, 32 + 16 * bi, 32, inv, 32, c0, 0, 16 * (bi + 1));
            __syncthreads();
            putBlock(inv, r0, c0, acc, true);
            __syncthreads();
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
