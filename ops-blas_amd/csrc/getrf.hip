// aclblasSgetrfBatched: batched LU with partial pivoting, FP32, in place.
//
// Reference: blas/getrf_batched/arch35/sgetrf_batched_host.cpp:32-118 (argument
// checks, launch) and sgetrf_batched_kernel.cpp:38-221 (right-looking
// unblocked LU: pivot search, row swap, division multipliers, rank-1 update,
// stop at the first zero pivot).
//
// MI355X design.  The matrix is held in registers, distributed by rows, and
// rows are never physically swapped during the factorisation: every row keeps
// its current logical position in a register, the pivot is chosen over the
// rows whose position is >= k, and at the end each row is stored once at its
// final position.  The LAPACK-format pivot vector (ipiv[k] = position of the
// chosen row at step k, 1-based) falls out of the position bookkeeping.
//   n <= 64  : one segment of a wave per matrix, one or two rows per lane
//              (tuned per size bucket), the pivot row broadcast through a
//              per-segment LDS slot (no barriers).
//   n <= 256 : blocked right-looking LU, one workgroup per matrix
//              (getrf_blocked.hip, lu_blocked.h): 32-column panels factored in
//              registers, trailing updates on the FP32 matrix instruction.
//   n > 256  : one workgroup per matrix working in global memory (L2), with
//              explicit row swaps.
// The n <= 64 and n > 256 paths perform, element by element, the same
// operations in the same order (l = a_ik / p, a_ij = fma(-l, u_kj, a_ij)) and
// produce bit-identical factors and pivots; the blocked path has the same
// pivot rule and multipliers but sums the trailing update blockwise, as
// LAPACK's blocked sgetrf does.
#include "device_common.h"
#include "drivers.h"
#include "internal.h"
#include "lapack_kernels.h"
#include "tuning.h"

namespace aclblas {
namespace {

using dev::kWave;

// ---------------------------------------------------------------------------
// n > 256: one workgroup per matrix, in global memory
// ---------------------------------------------------------------------------
constexpr int kGlobalThreads = 1024;

template <bool PIVOT>
__global__ __launch_bounds__(kGlobalThreads) void GetrfGlobalKernel(
    float* const* __restrict__ Aarray, int n, int lda, int* __restrict__ ipiv, int* __restrict__ info, int batch)
{
    constexpr int kWaves = kGlobalThreads / kWave;
    __shared__ uint64_t pkey[kWaves];
    __shared__ float spivot;
    const int b = blockIdx.x;
    if (b >= batch) {
        return;
    }
    float* A = Aarray[b];
    const int t = threadIdx.x;
    const int lane = dev::LaneId();
    const int wv = t / kWave;
    int infoVal = 0;

    for (int k = 0; k < n; ++k) {
        float* colk = A + static_cast<size_t>(k) * lda;
        int p = k;
        if (PIVOT) {
            uint64_t key = 0;
            for (int i = k + t; i < n; i += kGlobalThreads) {
                uint64_t kk = dev::PivotKey(colk[i], i, i);
                key = kk > key ? kk : key;
            }
            key = dev::WaveMaxKey(key);
            if (lane == 0) {
                pkey[wv] = key;
            }
            __syncthreads();
            uint64_t best = pkey[0];
#pragma unroll
            for (int q = 1; q < kWaves; ++q) {
                best = pkey[q] > best ? pkey[q] : best;
            }
            p = dev::KeyRow(best);
            if (t == 0) {
                ipiv[static_cast<size_t>(b) * n + k] = p + 1;
                spivot = colk[p];
            }
        } else if (t == 0) {
            spivot = colk[k];
        }
        __syncthreads();
        const float pivot = spivot;
        if (pivot == 0.0f) {
            infoVal = k + 1;
            break;
        }
        if (p != k) {
            for (int j = t; j < n; j += kGlobalThreads) {
                float* cj = A + static_cast<size_t>(j) * lda;
                const float x = cj[k];
                cj[k] = cj[p];
                cj[p] = x;
            }
        }
        __syncthreads();
        for (int i = k + 1 + t; i < n; i += kGlobalThreads) {
            colk[i] = colk[i] / pivot;
        }
        __syncthreads();
        const uint32_t m = static_cast<uint32_t>(n - k - 1);
        const uint32_t total = m * m;  // n <= 65535 keeps this in range
        for (uint32_t e = t; e < total; e += kGlobalThreads) {
            const int jj = static_cast<int>(e / m);
            const int ii = static_cast<int>(e - static_cast<uint32_t>(jj) * m);
            float* cj = A + static_cast<size_t>(k + 1 + jj) * lda;
            cj[k + 1 + ii] = fmaf(-colk[k + 1 + ii], cj[k], cj[k + 1 + ii]);
        }
        __syncthreads();
    }
    if (info != nullptr && t == 0) {
        info[b] = infoVal;
    }
}

// Rows per lane of the n <= 64 LU: the tuned choice for (S, batch), else two
// rows per lane while that still leaves two waves per SIMD (they halve the
// per-matrix instruction count but also the waves).
int WaveRows(const _aclblas_handle* h, int n, int batch)
{
    const int cus = h->cu_count > 0 ? h->cu_count : 256;
    const int S = n <= 8 ? 8 : (n <= 16 ? 16 : (n <= 32 ? 32 : 64));
    const long long wavesTwoRows = static_cast<long long>(batch) * S / 128;
    const int builtIn = (S > 8 && wavesTwoRows >= 8ll * cus) ? 2 : 1;
    return tune::Get(h, tune::kOpLu, tune::LuBucket(S, batch, cus), tune::Choice{builtIn, 1}).a;
}

template <bool PIVOT>
void LaunchGetrf(const _aclblas_handle* h, float* const* A, int n, int lda, int* ipiv, int* info, int batch)
{
    const hipStream_t st = h->stream;
    if (n <= 64) {
        LaunchGetrfWave<PIVOT>(st, A, n, lda, ipiv, info, batch, WaveRows(h, n, batch));
    } else if (n <= 256) {
        const int cus = h->cu_count > 0 ? h->cu_count : 256;
        LaunchGetrfBlocked<PIVOT>(st, A, n, lda, ipiv, info, batch, batch < 2 * cus ? batch : 2 * cus);
    } else {
        hipLaunchKernelGGL((GetrfGlobalKernel<PIVOT>), dim3(batch), dim3(kGlobalThreads), 0, st, A, n, lda, ipiv,
                           info, batch);
    }
}

// Argument checks in the reference's order (sgetrf_batched_host.cpp:32-64):
// lda is checked before the n==0 / batch==0 no-op.
aclblasStatus_t ValidateGetrf(int n, float* const* A, int lda, const int* piv, const int* info, int batch)
{
    if (n < 0 || batch < 0) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (lda < MaxInt(1, n)) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (n == 0 || batch == 0) {
        return ACLBLAS_STATUS_SUCCESS;
    }
    if (A == nullptr) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (piv != nullptr && info == nullptr) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    return ACLBLAS_STATUS_SUCCESS;
}

}  // namespace

aclblasStatus_t GetrfOutOfPlaceLaunch(const _aclblas_handle* h, int n, const float* const* A, int lda, float* out,
                                      float** table, int* piv, int* info, int batch)
{
    if (n <= 0 || n > 32) {
        return ACLBLAS_STATUS_NOT_SUPPORTED;
    }
    LaunchGetrfWaveOutOfPlace(h->stream, A, n, lda, out, table, piv, info, batch, WaveRows(h, n, batch));
    return LaunchStatus("aclblasSgetrfBatched");
}

// Internal entry shared with the solve/refinement drivers.
aclblasStatus_t GetrfBatchedLaunch(const _aclblas_handle* h, int n, float* const* A, int lda, int* piv, int* info,
                                   int batch)
{
    if (piv != nullptr) {
        LaunchGetrf<true>(h, A, n, lda, piv, info, batch);
    } else {
        LaunchGetrf<false>(h, A, n, lda, piv, info, batch);
    }
    return LaunchStatus("aclblasSgetrfBatched");
}

}  // namespace aclblas

extern "C" aclblasStatus_t aclblasSgetrfBatched(
    aclblasHandle_t handle, int n, float* const Aarray[], int lda, int* PivotArray, int* infoArray, int batchSize)
{
    using namespace aclblas;
    if (handle == nullptr) {
        return ACLBLAS_STATUS_HANDLE_IS_NULLPTR;
    }
    aclblasStatus_t st = ValidateGetrf(n, Aarray, lda, PivotArray, infoArray, batchSize);
    if (st != ACLBLAS_STATUS_SUCCESS) {
        LogError("aclblasSgetrfBatched", "invalid argument (n=%d lda=%d batch=%d)", n, lda, batchSize);
        return st;
    }
    if (n == 0 || batchSize == 0) {
        return ACLBLAS_STATUS_SUCCESS;
    }
    if (n > kMaxLapackN) {
        LogError("aclblasSgetrfBatched", "n=%d exceeds the supported maximum %d", n, kMaxLapackN);
        return ACLBLAS_STATUS_NOT_SUPPORTED;
    }
    if (!HandleHasDevice(handle)) {
        return ACLBLAS_STATUS_EXECUTION_FAILED;
    }
    return GetrfBatchedLaunch(handle, n, Aarray, lda, PivotArray, infoArray, batchSize);
}
