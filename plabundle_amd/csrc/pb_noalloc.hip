// pb_noalloc.hip - pb_hip_dense_solve_noalloc: the dense FP64 factor-and-solve of pb_hip_dense_solve on caller
// device memory, with every intermediate in a static per-device scratch array, so the call allocates nothing.
//
// Scratch layout (bytes from the array's device address):
//   [0, kMatBytes)                padded column-major matrix, order npad = n rounded up to 64 (identity padding)
//   [kMatBytes, +kVecBytes)       padded right-hand side / solution
//   [kMatBytes + kVecBytes, ...)  solver words (kDenseWordCount int32) and the two sweep progress words
// Every call rewrites the matrix and vector parts entirely; the solver words and progress words are cleared once
// per device at set-up and advance by epoch after that, so nothing has to start at zero per call.
#include "pb_common.h"
#include "pb_kernels.h"
#include "../../include/pb_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

namespace
{
    constexpr size_t kMaxRows = PB_NOALLOC_MAX_ROWS;
    constexpr size_t kMatBytes = kMaxRows * kMaxRows * sizeof(double);
    constexpr size_t kVecBytes = kMaxRows * sizeof(double);
    constexpr size_t kWordBytes = 256;  // kDenseWordCount int32 + 2 uint64, rounded up
    static_assert(kMatBytes + kVecBytes + kWordBytes <= PB_NOALLOC_SCRATCH_BYTES, "scratch layout");
    static_assert(kMaxRows % pbk::kDenseTile == 0, "the largest system is whole tiles");

    static __device__ __align__(256) unsigned char scratch[PB_NOALLOC_SCRATCH_BYTES];

    // A (n x n, leading dimension n) into the padded matrix, the padding the identity; b into the padded vector.
    __global__ void k_noalloc_pad(const double* A, const double* b, int n, int npad, double* P, double* v)
    {
        const size_t total = size_t(npad) * npad;
        for (size_t e = size_t(blockIdx.x) * blockDim.x + threadIdx.x; e < total; e += size_t(gridDim.x) * blockDim.x)
        {
            const int r = int(e % npad), c = int(e / npad);
            P[e] = (r < n && c < n) ? A[size_t(c) * n + r] : (r == c ? 1.0 : 0.0);
        }
        const int i = blockIdx.x * blockDim.x + threadIdx.x;
        if (i < npad) v[i] = i < n ? b[i] : 0.0;
    }

    __global__ void k_noalloc_clear(int32_t* words, unsigned long long* progress)
    {
        if (threadIdx.x < pbk::kDenseWordCount) words[threadIdx.x] = 0;
        if (threadIdx.x < 2) progress[threadIdx.x] = 0;
    }

    __global__ void k_noalloc_out(const double* v, int n, double* x)
    {
        const int i = blockIdx.x * blockDim.x + threadIdx.x;
        if (i < n) x[i] = v[i];
    }

    // Device address of `scratch` per device, resolved once per device with that device current.
    std::mutex g_scratchMutex;
    std::vector<unsigned char*> g_scratch;
}  // namespace

// Defined in pb_host.cpp: the context's device, stream and pinned readback area, and its error string.
extern "C++" bool pbNoallocContext(pb_ctx* ctx, int* device, hipStream_t* stream, int32_t** pinnedWords);
extern "C++" pb_status pbNoallocFail(pb_ctx* ctx, pb_status st, const std::string& message);

extern "C" pb_status pb_hip_dense_solve_noalloc(pb_ctx* ctx, int32_t n, const double* A, double* b, int32_t* info)
{
    int device = 0;
    hipStream_t s = nullptr;
    int32_t* pinned = nullptr;
    if (!pbNoallocContext(ctx, &device, &s, &pinned)) return PB_ERR_INVALID;
    if (n <= 0 || !A || !b) return pbNoallocFail(ctx, PB_ERR_INVALID, "dense solve: n must be positive, A and b set");
    if (size_t(n) > kMaxRows)
        return pbNoallocFail(ctx, PB_ERR_INVALID,
                             "dense solve without allocation: n = " + std::to_string(n) + " exceeds the scratch limit of " +
                                 std::to_string(kMaxRows) + " rows");
    if (hipSetDevice(device) != hipSuccess) return pbNoallocFail(ctx, PB_ERR_DEVICE, "hipSetDevice failed");
    unsigned char* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratchMutex);
        if (g_scratch.size() <= size_t(device)) g_scratch.resize(size_t(device) + 1, nullptr);
        if (!g_scratch[device])
        {
            void* p = nullptr;
            const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p)
                return pbNoallocFail(ctx, PB_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            unsigned char* q = static_cast<unsigned char*>(p);
            k_noalloc_clear<<<1, 64, 0, s>>>(reinterpret_cast<int32_t*>(q + kMatBytes + kVecBytes),
                                             reinterpret_cast<unsigned long long*>(q + kMatBytes + kVecBytes + 128));
            if (hipStreamSynchronize(s) != hipSuccess) return pbNoallocFail(ctx, PB_ERR_DEVICE, "scratch set-up failed");
            g_scratch[device] = q;
        }
        base = g_scratch[device];
    }
    const int nt = (n + pbk::kDenseTile - 1) / pbk::kDenseTile;
    const int npad = nt * pbk::kDenseTile;
    double* P = reinterpret_cast<double*>(base);
    double* v = reinterpret_cast<double*>(base + kMatBytes);
    pbk::DenseState st;
    st.words = reinterpret_cast<int32_t*>(base + kMatBytes + kVecBytes);
    st.progress = reinterpret_cast<unsigned long long*>(base + kMatBytes + kVecBytes + 128);
    const int blocks = std::max((npad + 255) / 256, std::min(2048, (npad * npad + 255) / 256));
    k_noalloc_pad<<<blocks, 256, 0, s>>>(A, b, n, npad, P, v);
    pbk::DenseLauncher<double>::factor(P, nt, st, s);
    pbk::DenseLauncher<double>::solve(P, nt, v, st, s);
    k_noalloc_out<<<(n + 255) / 256, 256, 0, s>>>(v, n, b);
    if (hipMemcpyAsync(pinned, st.words, sizeof(int32_t) * pbk::kDenseWordCount, hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return pbNoallocFail(ctx, PB_ERR_DEVICE, "dense solve without allocation failed on the device");
    if (pinned[pbk::kDenseTimeout]) return pbNoallocFail(ctx, PB_ERR_DEVICE, "dense triangular solve stopped waiting");
    if (info) *info = pinned[pbk::kDenseNotSpd] ? 1 : 0;
    return PB_OK;
}
