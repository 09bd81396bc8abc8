// aclblasSgetrsBatchedAdjoint: reverse-mode derivative of X = A^-1 B for a
// batch of systems factored by aclblasSgetrfBatched.
//
// Given the forward solution X and the upstream gradient dX = dL/dX:
//   dB = A^-T dX            one transposed solve with the saved LU factors and
//                           pivots (aclblasSgetrsBatched, trans T), in place;
//   dA = -dB X^T            one FP32 batched GEMM (n x nrhs times nrhs x n).
// The inverse is never formed.  The reference library has no such entry; its
// two building blocks are the reference's getrs with trans T
// (blas/getrs_batched/arch35/sgetrs_batched_kernel.cpp:263-362) and a batched
// FP32 GEMM (blas/gemm_batched/arch35/gemm_batched_host.cpp:278-360).
#include "drivers.h"
#include "internal.h"

namespace aclblas {
namespace {

constexpr int kMaxNrhs = 256;

aclblasStatus_t ValidateAdjoint(int n, int nrhs, const float* const* LU, int lda, const float* const* X, int ldx,
                                float* const* dX, int lddx, float* const* dA, int ldda, int batch)
{
    if (n < 0 || nrhs < 0 || nrhs > kMaxNrhs || batch < 0) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (lda < MaxInt(1, n) || ldx < MaxInt(1, n) || lddx < MaxInt(1, n)) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (dA != nullptr && ldda < MaxInt(1, n)) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (n == 0 || batch == 0) {
        return ACLBLAS_STATUS_SUCCESS;
    }
    if (LU == nullptr || (nrhs > 0 && (dX == nullptr || (dA != nullptr && X == nullptr)))) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    return ACLBLAS_STATUS_SUCCESS;
}

}  // namespace
}  // namespace aclblas

extern "C" aclblasStatus_t aclblasSgetrsBatchedAdjoint(
    aclblasHandle_t handle, int n, int nrhs, const float* const LUarray[], int lda, const int* devIpiv,
    const float* const Xarray[], int ldx, float* const dXarray[], int lddx, float* const dAarray[], int ldda,
    int batchCount)
{
    using namespace aclblas;
    if (handle == nullptr) {
        return ACLBLAS_STATUS_HANDLE_IS_NULLPTR;
    }
    aclblasStatus_t st = ValidateAdjoint(n, nrhs, LUarray, lda, Xarray, ldx, dXarray, lddx, dAarray, ldda,
                                         batchCount);
    if (st != ACLBLAS_STATUS_SUCCESS) {
        LogError("aclblasSgetrsBatchedAdjoint", "invalid argument (n=%d nrhs=%d batch=%d)", n, nrhs, batchCount);
        return st;
    }
    if (n == 0 || batchCount == 0) {
        return ACLBLAS_STATUS_SUCCESS;
    }
    if (!HandleHasDevice(handle)) {
        return ACLBLAS_STATUS_EXECUTION_FAILED;
    }
    if (nrhs > 0) {
        st = GetrsBatchedLaunch(handle->stream, true, n, nrhs, LUarray, lda, devIpiv, dXarray, lddx, batchCount);
        if (st != ACLBLAS_STATUS_SUCCESS) {
            return st;
        }
    }
    if (dAarray == nullptr) {
        return ACLBLAS_STATUS_SUCCESS;
    }
    // dA = -dB X^T  (k = nrhs; nrhs == 0 gives dA = 0)
    return SgemmBatchedLaunch(handle, false, true, n, n, nrhs, -1.0f,
                              const_cast<const float* const*>(dXarray), lddx, Xarray, ldx, 0.0f, dAarray, ldda,
                              batchCount);
}
