// pb_band.hip - block-tridiagonal reduced solve by odd-even cyclic reduction (gfx950, wave64), FP64 or FP32.
//
// The reduced system S of a strip or block of cameras couples each camera only with its neighbours in a
// bandwidth-reducing order (the camera graph of an aerial block has a bandwidth of ~17 cameras in 1000), so the
// pose rows cut into super-blocks of sp = 64 t rows form a block-tridiagonal matrix; the intrinsic-group rows form
// a dense border coupled to all of them. Instead of the n^3 / 3 of the dense Cholesky this takes O(m sp^3):
//   k_band_clear, k_band_gather   the super-blocks D_k, the first-level couplings K, the border columns BdT_k
//                                 and the border block C from the block-sparse rows of S (identity padding)
//   per level (every second remaining super-block e, all in parallel):
//     k_band_factor               D_e = L L^T in LDS (64 x 64 tiles: one-wave tile Cholesky and inverse, tile
//                                 products), W_e = L^-T written in place of D_e
//     k_band_gemm (X tasks)       X = W_e^T K for the two couplings of e, XB_e = W_e^T BdT_e      (MFMA 16x16x4)
//     k_band_gemm (U tasks)       D_s -= X^T X and BdT_s -= X^T XB for the survivors s, the new coupling
//                                 -X_b^T X_a between the survivors, Pc_e = XB_e^T XB_e            (MFMA 16x16x4)
//     k_band_csum                 C -= sum_e Pc_e in a fixed order
//   k_band_final + the dense route  the last super-block with the border (order sp + nbp)
// The solve runs the same levels forward (y_e = W_e^T (r_e - pending X^T y)) and backward
// (x_e = W_e (y_e - X x_neighbours - XB_e x_border)), one workgroup per eliminated block and level.
// Every reduction has a fixed order, so results are bitwise reproducible.
#include "pb_common.h"
#include "pb_kernels.h"
#include "pb_tile.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>

namespace pbk
{
    namespace
    {
        using namespace tile;

        int blocksOf(long n, int per) { return int((n + per - 1) / per); }
    }  // namespace

    // ---------------------------------------------------------------------------------------------
    // Zero the super-blocks, the first-level couplings, the border columns and the border block; identity on
    // the padding rows of each super-block (rows past its units) and of the border.
    template <typename TA>
    __global__ void k_band_clear(TA* Z, long long offD, long long offC, long long offEnd, int sp, int su, int nPose,
                                 int nb, int nbp)
    {
        const long long total = offEnd - offD;
        for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
             i += (long long)gridDim.x * blockDim.x)
        {
            const long long g = offD + i;
            TA v = TA(0);
            if (g < offD + (long long)(nPose + su - 1) / su * sp * sp)
            {
                const long long local = g - offD;
                const int k = int(local / ((long long)sp * sp));
                const int rc = int(local % ((long long)sp * sp));
                const int r = rc % sp, c = rc / sp;
                const int used = 6 * min(su, nPose - k * su);
                if (r == c && r >= used) v = TA(1);
            }
            else if (g >= offC && g < offC + (long long)nbp * nbp)
            {
                const int rc = int(g - offC);
                const int r = rc % nbp, c = rc / nbp;
                if (r == c && r >= nb) v = TA(1);
            }
            Z[g] = v;
        }
    }

    // Scatter of the stored 6 x 6 blocks (one wave per unit row): pose/pose in one super-block into D, pose rows of
    // an odd super-block against a neighbouring super-block into the first-level coupling (rows = the odd block,
    // eliminated first), pose rows against border columns into BdT, border against border into C. Blocks whose
    // transpose carries the same values are skipped.
    template <typename TW, typename TA>
    __global__ void __launch_bounds__(256) k_band_gather(Problem P, const TW* S, BandDev B, TA* Z)
    {
        const int u = blockIdx.x * 4 + (threadIdx.x >> 6);
        if (u >= P.nUnits) return;
        const int lane = threadIdx.x & 63;
        const int sp = B.sp, nbp = B.nbp;
        const int bu = B.unitBlk[u], ou = B.unitOff[u];
        for (int k = P.rowPtr[u] + lane; k < P.rowPtr[u + 1]; k += kWave)
        {
            const int v = P.colIdx[k];
            const int bv = B.unitBlk[v], ov = B.unitOff[v];
            TA* dst = nullptr;
            int ld = 0;
            if (bu >= 0 && bv >= 0)
            {
                if (bu == bv)
                {
                    dst = Z + B.offD + size_t(bu) * sp * sp;
                    ld = sp;
                }
                else if ((bu & 1) && (bu == bv + 1 || bu + 1 == bv))
                {
                    dst = Z + B.offK + size_t(B.k0slot[min(bu, bv)]) * sp * sp;
                    ld = sp;
                }
            }
            else if (bu >= 0)
            {
                dst = Z + B.offB + size_t(bu) * sp * nbp;
                ld = sp;
            }
            else if (bv < 0)
            {
                dst = Z + B.offC;
                ld = nbp;
            }
            if (!dst) continue;
            const TW* blk = S + size_t(k) * 36;
#pragma unroll
            for (int e = 0; e < 36; ++e) dst[size_t(ov + e % 6) * ld + ou + e / 6] = TA(blk[e]);
        }
    }

    // ---------------------------------------------------------------------------------------------
    // One workgroup per eliminated super-block of a level: D_e = L L^T, then W_e = L^-T written over D_e.
    // sp = 128: L00 = chol(A00), I0 = L00^-1, L10 = A10 I0^T, A11 -= L10 L10^T, L11 = chol(A11), I1 = L11^-1,
    // L^-1 = [I0, 0; -I1 L10 I0, I1].
    // Dynamic LDS: four 64 x 65 tiles and a 64-entry column (133 KiB in FP64 for sp = 128, above the default
    // 64 KiB limit; the launcher raises the kernel's limit once).
    template <typename T>
    constexpr size_t bandFactorLds(int sp)
    {
        return (size_t(sp > kT ? 4 : 2) * kT * kLd + kT) * sizeof(T);
    }

    template <typename T>
    __global__ void __launch_bounds__(256) k_band_factor(T* Z, const BandElim* elims, long long offD, int sp,
                                                         int32_t* words)
    {
        extern __shared__ __align__(16) unsigned char bandLds[];
        T(*b0)[kLd] = reinterpret_cast<T(*)[kLd]>(bandLds);
        T(*b1)[kLd] = b0 + kT;
        T(*b2)[kLd] = b1 + kT;  // only with sp > 64
        T(*b3)[kLd] = b2 + kT;
        T* col = reinterpret_cast<T*>(b0 + (sp > kT ? 4 : 2) * kT);
        const int e = elims[blockIdx.x].e;
        T* M = Z + offD + size_t(e) * sp * sp;
        bool bad = false;
        loadTile(b0, M, sp, 0, 0);
        if (sp > kT)
        {
            loadTile(b2, M, sp, kT, 0);
            loadTile(b3, M, sp, kT, kT);
        }
        __syncthreads();
        cholTile(b0, col, bad);
        invTile(b0, b1, col);
        if (sp > kT)
        {
            tileProduct(b0, b2, false, b1, true, false, false);  // L10 = A10 I0^T
            __syncthreads();
            tileProduct(b3, b0, false, b0, true, true, false);   // A11 -= L10 L10^T
            __syncthreads();
            cholTile(b3, col, bad);
            invTile(b3, b2, col);                                 // I1
            tileProduct(b3, b0, false, b1, false, false, false);  // L10 I0
            __syncthreads();
            tileProduct(b0, b2, false, b3, false, false, true);   // Linv10 = -I1 (L10 I0)
            __syncthreads();
        }
        // W(r, c) = Linv(c, r), tile by tile: W_00 = I0^T, W_01 = Linv10^T, W_11 = I1^T, W_10 = 0
        storeTransposed(M, sp, 0, 0, b1);
        if (sp > kT)
        {
            storeTransposed(M, sp, 0, kT, b0);
            storeTransposed(M, sp, kT, kT, b2);
            storeTransposed(M, sp, kT, 0, static_cast<T(*)[kLd]>(nullptr));
        }
        if (bad && threadIdx.x == 0) atomicOr(words + kDenseNotSpd, 1);
    }

    // ---------------------------------------------------------------------------------------------
    // Four workgroups per 64 x 64 output tile of a task, one per 32 x 32 quarter: out = beta out +- sum_terms
    // P^T Q (MFMA 16x16x4, operands staged in LDS 32 rows of k at a time). Wave w forms the 16 x 16 block
    // (w >> 1, w & 1) of the quarter. Quarters spread a level's products over four times as many CUs; every
    // output element is accumulated over k in the same order as with whole tiles.
    constexpr int kBandGemmQuarters = 4;
    template <typename T>
    __global__ void __launch_bounds__(256) k_band_gemm(T* Z, const BandTask* tasks, const BandTile* tiles, int sp)
    {
        using V = typename BandMfma<T>::V;
        constexpr int kStage = 32, kG = 32;
        __shared__ T sa[kStage][kG + 2], sb[kStage][kG + 2];  // [k][column of P / Q]
        const BandTile tl = tiles[blockIdx.x / kBandGemmQuarters];
        const int quarter = blockIdx.x % kBandGemmQuarters;
        const BandTask tk = tasks[tl.task];
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        const int qi = (w >> 1) * 16, qj = (w & 1) * 16;
        const int li = lane & 15, lk = lane >> 4;
        constexpr int kLR = BandMfma<T>::kLaneRow, kIR = BandMfma<T>::kItemRow;
        const int r0 = tl.ti * kT + (quarter >> 1) * kG, c0 = tl.tj * kT + (quarter & 1) * kG;
        T* out = Z + tk.out + size_t(c0) * tk.ldo + r0;
        V acc;
        {
            const T* src = out + size_t(qj + li) * tk.ldo + qi + kLR * lk;
#pragma unroll
            for (int e = 0; e < 4; ++e) acc[e] = tk.beta ? src[kIR * e] : T(0);
        }
        const int kEnd = tk.kTri ? min(sp, r0 + kG) : sp;
        const T sign = tk.alphaNeg ? T(-1) : T(1);
        // steps of kStage rows of k over every term; the next step's operands are loaded into registers while
        // the current one is multiplied
        constexpr int kPer = kStage * kG / 256;
        const int perTerm = kEnd / kStage, steps = perTerm * tk.terms;
        T pa[kPer], pb[kPer];
        auto load = [&](int st) {
            const int term = st / perTerm, k0 = (st % perTerm) * kStage;
            const T* Pm = Z + (term ? tk.p1 : tk.p0) + size_t(r0) * sp + k0;
            const T* Qm = Z + (term ? tk.q1 : tk.q0) + size_t(c0) * sp + k0;
#pragma unroll
            for (int q = 0; q < kPer; ++q)
            {
                const int idx = threadIdx.x + 256 * q;
                pa[q] = Pm[size_t(idx / kStage) * sp + idx % kStage];
                pb[q] = Qm[size_t(idx / kStage) * sp + idx % kStage];
            }
        };
        if (steps > 0) load(0);
        for (int st = 0; st < steps; ++st)
        {
            __syncthreads();  // every wave is done with the previous slice
#pragma unroll
            for (int q = 0; q < kPer; ++q)
            {
                const int idx = threadIdx.x + 256 * q;
                sa[idx % kStage][idx / kStage] = pa[q];
                sb[idx % kStage][idx / kStage] = pb[q];
            }
            __syncthreads();
            if (st + 1 < steps) load(st + 1);
#pragma unroll
            for (int kk = 0; kk < kStage; kk += 4)
                acc = BandMfma<T>::op(sign * sa[kk + lk][qi + li], sb[kk + lk][qj + li], acc);
        }
        T* dst = out + size_t(qj + li) * tk.ldo + qi + kLR * lk;
#pragma unroll
        for (int e = 0; e < 4; ++e) dst[kIR * e] = acc[e];
    }

    // C -= sum over the level's eliminated blocks e of Pc_e, in elimination order (one thread per entry).
    template <typename T>
    __global__ void k_band_csum(T* Z, const BandElim* elims, int count, long long offC, long long offPc, int nbp)
    {
        const int i = blockIdx.x * blockDim.x + threadIdx.x;
        if (i >= nbp * nbp) return;
        T s = T(0);
        // eight loads in flight per step, summed in elimination order
        int q = 0;
        for (; q + 8 <= count; q += 8)
        {
            T v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = Z[offPc + size_t(elims[q + u].e) * nbp * nbp + i];
#pragma unroll
            for (int u = 0; u < 8; ++u) s += v[u];
        }
        for (; q < count; ++q) s += Z[offPc + size_t(elims[q].e) * nbp * nbp + i];
        Z[offC + i] -= s;
    }

    // F = [D_0, BdT_0; BdT_0^T, C] (order nF = sp + nbp, leading dimension nF), from the lower triangle of D_0.
    template <typename T>
    __global__ void k_band_final(T* Z, BandDev B)
    {
        const int nF = B.nF, sp = B.sp;
        const long long total = (long long)nF * nF;
        for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
             idx += (long long)gridDim.x * blockDim.x)
        {
            const int r = int(idx % nF), c = int(idx / nF);
            T v;
            if (r < sp && c < sp)
                v = r >= c ? Z[B.offD + size_t(c) * sp + r] : Z[B.offD + size_t(r) * sp + c];
            else if (r < sp)
                v = Z[B.offB + size_t(c - sp) * sp + r];
            else if (c < sp)
                v = Z[B.offB + size_t(r - sp) * sp + c];
            else
                v = Z[B.offC + size_t(c - sp) * B.nbp + (r - sp)];
            Z[B.offF + idx] = v;
        }
    }

    // ---------------------------------------------------------------------------------------------
    // Solve helpers for one workgroup of 256 threads and 16 outputs of a matrix-vector product with a
    // column-major matrix M (leading dimension ld); a level's products are spread over sp / 16 workgroups per
    // eliminated block, so each workgroup issues its loads together and waits about one memory latency.
    // T-form: out_k = sum_i M(i, k) x_i for k = k0 .. k0 + 15 (k < nOut), i < K <= 128 (i <= k with tri, M upper
    // triangular): 16 lanes per output, each summing 8 consecutive i of column k (one 64-byte run), then a 16-lane
    // butterfly; store(k, out_k) runs on the first lane of each output.
    template <typename T, typename F>
    __device__ __forceinline__ void gemvT16(const T* M, int ld, int k0, int nOut, int K, bool tri, const T* x,
                                            F&& store)
    {
        const int g = threadIdx.x >> 4, h = threadIdx.x & 15;
        const int k = k0 + g;
        T s = T(0);
        if (k < nOut)
        {
            const int iEnd = tri ? min(K, k + 1) : K;
            const T* col = M + size_t(k) * ld + 8 * h;
#pragma unroll
            for (int j = 0; j < 8; ++j)
                if (8 * h + j < iEnd) s += col[j] * x[8 * h + j];
        }
#pragma unroll
        for (int off = 8; off > 0; off >>= 1) s += __shfl_xor(s, off, 16);
        if (h == 0 && k < nOut) store(k, s);
    }

    // N-form part: thread (r = t & 15, kq = t >> 4) sums M(i, k) x_k over k = kq, kq + 16, ..., kq + 16 (J - 1) in
    // order (K = 16 J, J <= 16); the caller adds the 16 parts of row i in kq order. Triangular factors are stored
    // with their zeros, so a triangular product is a full one.
    template <typename T>
    __device__ __forceinline__ T rowPart(const T* M, int ld, int i, int J, const T* x)
    {
        const int kq = threadIdx.x >> 4;
        T v[16];
#pragma unroll
        for (int j = 0; j < 16; ++j)
            if (j < J) v[j] = M[size_t(kq + 16 * j) * ld + i];
        T s = T(0);
#pragma unroll
        for (int j = 0; j < 16; ++j)
            if (j < J) s += v[j] * x[kq + 16 * j];
        return s;
    }

    // Forward level, first launch (sp / 16 workgroups per eliminated block): r_e = v_e minus its pending
    // contributions, and 16 entries of y_e = W_e^T r_e into Y.
    template <typename T>
    __global__ void __launch_bounds__(256) k_band_forward_y(T* Z, BandDev B, const BandElim* elims)
    {
        __shared__ T r[kBandMaxRows];
        const int sp = B.sp, parts = sp / 16;
        const BandElim el = elims[blockIdx.x / parts];
        const int p = blockIdx.x % parts;
        const int i = threadIdx.x;
        if (i < sp)
        {
            T s = Z[B.offV + size_t(el.e) * sp + i];
            for (int q = el.pendBegin; q < el.pendEnd; ++q) s -= Z[B.offPV + size_t(B.pend[q]) * sp + i];
            r[i] = s;
        }
        __syncthreads();
        T* y = Z + B.offY + size_t(el.e) * sp;
        gemvT16(Z + B.offD + size_t(el.e) * sp * sp, sp, 16 * p, sp, sp, true, r, [&](int k, T s) { y[k] = s; });
    }

    // Forward level, second launch (2 sp / 16 + nbp / 16 workgroups per eliminated block): v_e = y_e, and 16
    // entries of one contribution: X_a^T y_e or X_b^T y_e to a survivor's pending slot, or XB_e^T y_e to the
    // border.
    template <typename T>
    __global__ void __launch_bounds__(256) k_band_forward_c(T* Z, BandDev B, const BandElim* elims)
    {
        __shared__ T y[kBandMaxRows];
        const int sp = B.sp, nbp = B.nbp, ps = sp / 16, parts = 2 * ps + nbp / 16;
        const BandElim el = elims[blockIdx.x / parts];
        const int q = blockIdx.x % parts;
        if (q >= ps && q < 2 * ps && el.slotB < 0) return;  // the last block of a level has one survivor
        const int i = threadIdx.x;
        if (i < sp)
        {
            const T v = Z[B.offY + size_t(el.e) * sp + i];
            y[i] = v;
            if (q == 0) Z[B.offV + size_t(el.e) * sp + i] = v;
        }
        __syncthreads();
        if (q < 2 * ps)
        {
            const int slot = q < ps ? el.slotA : el.slotB;
            T* pv = Z + B.offPV + size_t(slot) * sp;
            gemvT16(Z + B.offX + size_t(slot) * sp * sp, sp, 16 * (q % ps), sp, sp, false, y,
                    [&](int k, T s) { pv[k] = s; });
        }
        else
        {
            T* pB = Z + B.offPB + size_t(el.e) * nbp;
            gemvT16(Z + B.offXB + size_t(el.e) * sp * nbp, sp, 16 * (q - 2 * ps), nbp, sp, false, y,
                    [&](int k, T s) { pB[k] = s; });
        }
    }

    // The last system's right-hand side: block 0 minus its pending contributions, the border minus every
    // eliminated block's contribution (in block order). One workgroup.
    template <typename T>
    __global__ void __launch_bounds__(256) k_band_final_rhs(T* Z, BandDev B)
    {
        const int sp = B.sp, nbp = B.nbp;
        for (int i = threadIdx.x; i < sp + nbp; i += blockDim.x)
        {
            T s;
            if (i < sp)
            {
                s = Z[B.offV + i];
                for (int p = B.lastPendBegin; p < B.lastPendEnd; ++p) s -= Z[B.offPV + size_t(B.pend[p]) * sp + i];
            }
            else
            {
                const int j = i - sp;
                s = Z[B.offV + size_t(B.m) * sp + j];
                int e = 1;
                for (; e + 8 <= B.m; e += 8)
                {
                    T v[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) v[u] = Z[B.offPB + size_t(e + u) * nbp + j];
#pragma unroll
                    for (int u = 0; u < 8; ++u) s -= v[u];
                }
                for (; e < B.m; ++e) s -= Z[B.offPB + size_t(e) * nbp + j];
            }
            Z[B.offFV + i] = s;
        }
    }

    // The last system's solution back into the band vector (block 0 and the border).
    template <typename T>
    __global__ void k_band_final_x(T* Z, BandDev B)
    {
        const int i = blockIdx.x * blockDim.x + threadIdx.x;
        if (i >= B.sp + B.nbp) return;
        const size_t dst = i < B.sp ? size_t(i) : size_t(B.m) * B.sp + (i - B.sp);
        Z[B.offV + dst] = Z[B.offFV + i];
    }

    // Backward level, first launch (sp / 16 workgroups per eliminated block): 16 rows of
    // t_e = ((y_e - X_a x_a) - X_b x_b) - XB_e x_border into Y (y_e is in v_e).
    template <typename T>
    __global__ void __launch_bounds__(256) k_band_backward_t(T* Z, BandDev B, const BandElim* elims)
    {
        __shared__ T xa[kBandMaxRows], xb[kBandMaxRows], xB[kBandMaxBorder];
        __shared__ T part[3][16][17];
        const int sp = B.sp, nbp = B.nbp, parts = sp / 16;
        const BandElim el = elims[blockIdx.x / parts];
        const int p = blockIdx.x % parts;
        const T* V = Z + B.offV;
        for (int k = threadIdx.x; k < sp; k += blockDim.x)
        {
            xa[k] = V[size_t(el.a) * sp + k];
            xb[k] = el.b >= 0 ? V[size_t(el.b) * sp + k] : T(0);
        }
        for (int k = threadIdx.x; k < nbp; k += blockDim.x) xB[k] = V[size_t(B.m) * sp + k];
        __syncthreads();
        const int r = threadIdx.x & 15, kq = threadIdx.x >> 4, i = 16 * p + r;
        part[0][kq][r] = rowPart(Z + B.offX + size_t(el.slotA) * sp * sp, sp, i, sp / 16, xa);
        part[1][kq][r] = el.slotB >= 0 ? rowPart(Z + B.offX + size_t(el.slotB) * sp * sp, sp, i, sp / 16, xb) : T(0);
        part[2][kq][r] = nbp > 0 ? rowPart(Z + B.offXB + size_t(el.e) * sp * nbp, sp, i, nbp / 16, xB) : T(0);
        __syncthreads();
        if (threadIdx.x < 16)
        {
            T ya = T(0), yb = T(0), yB = T(0);
            for (int q = 0; q < 16; ++q)
            {
                ya += part[0][q][r];
                yb += part[1][q][r];
                yB += part[2][q][r];
            }
            Z[B.offY + size_t(el.e) * sp + i] = ((V[size_t(el.e) * sp + i] - ya) - yb) - yB;
        }
    }

    // Backward level, second launch (sp / 16 workgroups per eliminated block): 16 rows of x_e = W_e t_e into v_e.
    template <typename T>
    __global__ void __launch_bounds__(256) k_band_backward_x(T* Z, BandDev B, const BandElim* elims)
    {
        __shared__ T t[kBandMaxRows];
        __shared__ T part[16][17];
        const int sp = B.sp, parts = sp / 16;
        const BandElim el = elims[blockIdx.x / parts];
        const int p = blockIdx.x % parts;
        for (int k = threadIdx.x; k < sp; k += blockDim.x) t[k] = Z[B.offY + size_t(el.e) * sp + k];
        __syncthreads();
        const int r = threadIdx.x & 15, kq = threadIdx.x >> 4, i = 16 * p + r;
        part[kq][r] = rowPart(Z + B.offD + size_t(el.e) * sp * sp, sp, i, sp / 16, t);
        __syncthreads();
        if (threadIdx.x < 16)
        {
            T x = T(0);
            for (int q = 0; q < 16; ++q) x += part[q][r];
            Z[B.offV + size_t(el.e) * sp + i] = x;
        }
    }

    template <typename TA>
    __global__ void k_band_to(BandDev B, const double* src, TA* Z)
    {
        const int i = blockIdx.x * blockDim.x + threadIdx.x;
        if (i >= B.vN) return;
        const int j = B.srcOf[i];
        Z[B.offV + i] = j >= 0 ? TA(src[j]) : TA(0);
    }

    template <typename TA>
    __global__ void k_band_from(BandDev B, double* x, const TA* Z, int overwrite)
    {
        const int j = blockIdx.x * blockDim.x + threadIdx.x;
        if (j >= B.n) return;
        const double v = double(Z[B.offV + B.rowOf[j]]);
        x[j] = overwrite ? v : x[j] + v;
    }

    // ---------------------------------------------------------------------------------------------
    template <typename TA>
    template <typename TW>
    void BandLauncher<TA>::gather(const Problem& P, const TW* S, const BandView& B, TA* Z, hipStream_t s)
    {
        const BandDev& D = B.d;
        const long long clearEnd = D.offC + (long long)D.nbp * D.nbp;  // D, K, B, C are contiguous
        k_band_clear<TA><<<2048, 256, 0, s>>>(Z, D.offD, D.offC, clearEnd, D.sp, D.su, D.nPose, D.nb, D.nbp);
        k_band_gather<TW, TA><<<blocksOf(P.nUnits, 4), 256, 0, s>>>(P, S, D, Z);
    }

    template <typename TA>
    void BandLauncher<TA>::factor(const BandView& B, TA* Z, const DenseState& st, hipStream_t s)
    {
        const BandDev& D = B.d;
        const size_t lds = bandFactorLds<TA>(D.sp);
        static const bool ldsRaised = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_band_factor<TA>),
                                                          hipFuncAttributeMaxDynamicSharedMemorySize,
                                                          int(bandFactorLds<TA>(kBandMaxRows))) == hipSuccess;
        if (!ldsRaised && lds > 64 * 1024) throw std::runtime_error("hipFuncSetAttribute (band factor LDS)");
        DenseLauncher<TA>::begin(st, s);
        for (int l = 0; l < B.levels; ++l)
        {
            const int e0 = B.elimStart[l], ne = B.elimStart[l + 1] - e0;
            k_band_factor<TA><<<ne, 256, lds, s>>>(Z, D.elims + e0, D.offD, D.sp, st.words);
            const int x0 = B.tileXStart[l], nx = B.tileXStart[l + 1] - x0;
            if (nx > 0) k_band_gemm<TA><<<nx * kBandGemmQuarters, 256, 0, s>>>(Z, D.tasks, D.tiles + x0, D.sp);
            const int u0 = B.tileUStart[l], nu = B.tileUStart[l + 1] - u0;
            if (nu > 0) k_band_gemm<TA><<<nu * kBandGemmQuarters, 256, 0, s>>>(Z, D.tasks, D.tiles + u0, D.sp);
            if (D.nbp > 0)
                k_band_csum<TA><<<blocksOf((long)D.nbp * D.nbp, 256), 256, 0, s>>>(Z, D.elims + e0, ne, D.offC,
                                                                                      D.offPc, D.nbp);
        }
        k_band_final<TA><<<blocksOf((long)D.nF * D.nF, 256), 256, 0, s>>>(Z, D);
        DenseLauncher<TA>::factor(Z + D.offF, D.nF / kT, st, s, false);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) throw std::runtime_error(std::string("band factorisation launch: ") + hipGetErrorString(e));
    }

    template <typename TA>
    void BandLauncher<TA>::solve(const BandView& B, TA* Z, const DenseState& st, hipStream_t s)
    {
        const BandDev& D = B.d;
        const int ps = D.sp / 16;
        for (int l = 0; l < B.levels; ++l)
        {
            const int e0 = B.elimStart[l], ne = B.elimStart[l + 1] - e0;
            k_band_forward_y<TA><<<ne * ps, 256, 0, s>>>(Z, D, D.elims + e0);
            k_band_forward_c<TA><<<ne * (2 * ps + D.nbp / 16), 256, 0, s>>>(Z, D, D.elims + e0);
        }
        k_band_final_rhs<TA><<<1, 256, 0, s>>>(Z, D);
        DenseLauncher<TA>::solve(Z + D.offF, D.nF / kT, Z + D.offFV, st, s);
        k_band_final_x<TA><<<blocksOf(D.nF, 256), 256, 0, s>>>(Z, D);
        for (int l = B.levels - 1; l >= 0; --l)
        {
            const int e0 = B.elimStart[l], ne = B.elimStart[l + 1] - e0;
            k_band_backward_t<TA><<<ne * ps, 256, 0, s>>>(Z, D, D.elims + e0);
            k_band_backward_x<TA><<<ne * ps, 256, 0, s>>>(Z, D, D.elims + e0);
        }
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) throw std::runtime_error(std::string("band solve launch: ") + hipGetErrorString(e));
    }

    template <typename TA>
    void BandLauncher<TA>::toBand(const BandView& B, const double* src, TA* Z, hipStream_t s)
    {
        k_band_to<TA><<<blocksOf(B.d.vN, 256), 256, 0, s>>>(B.d, src, Z);
    }

    template <typename TA>
    void BandLauncher<TA>::fromBand(const BandView& B, double* x, const TA* Z, bool overwrite, hipStream_t s)
    {
        k_band_from<TA><<<blocksOf(B.d.n, 256), 256, 0, s>>>(B.d, x, Z, overwrite ? 1 : 0);
    }

    template struct BandLauncher<float>;
    template struct BandLauncher<double>;
    template void BandLauncher<float>::gather<float>(const Problem&, const float*, const BandView&, float*,
                                                     hipStream_t);
    template void BandLauncher<float>::gather<double>(const Problem&, const double*, const BandView&, float*,
                                                      hipStream_t);
    template void BandLauncher<double>::gather<float>(const Problem&, const float*, const BandView&, double*,
                                                      hipStream_t);
    template void BandLauncher<double>::gather<double>(const Problem&, const double*, const BandView&, double*,
                                                       hipStream_t);
}  // namespace pbk
