// aclblasSmatinvBatched: Ainv[i] = A[i]^-1 for n <= 32 as getrf followed by
// getri.
//
// Reference: blas/matinv_batched/arch35/matinv_batched_host.cpp:42-257
// (argument checks and their order, n <= 32, the workspace of
// batch * 8 + n^2 * batch * 4 + n * batch * 4 bytes taken from the handle's
// active workspace without growing it, EXECUTION_FAILED when it is smaller,
// pivoted LU, then the out-of-place inverse whose info replaces the LU's).
//
// MI355X design: two launches on the handle's stream instead of five.  The
// register-resident LU reads each A[i] and writes its factors to a packed
// n x n block of the workspace, the block's address to the workspace's
// pointer table and all n pivots; the inverse kernel (one lane per column)
// reads the factors and writes Ainv[i] directly.  The reference copies A into
// Ainv, factors Ainv and copies the factors out again before inverting; the
// factors, pivots, inverse and info are the same, with three HBM passes fewer.
// The pivots after a zero pivot are written as 0, so the interchanges that a
// singular member's LU never reaches are skipped by the inverse (the reference
// leaves them as whatever the workspace held).  Nothing is allocated, so the
// call can be captured.
#include "drivers.h"
#include "internal.h"

namespace aclblas {
namespace {

constexpr int kMaxMatinvN = 32;

// reference matinv_batched_host.cpp:42-81 (same checks, same order)
aclblasStatus_t ValidateMatinv(int n, int lda, int ldaInv, int batch, const float* const* A, float* const* Ainv,
                               const int* info)
{
    if (n < 0 || n > kMaxMatinvN || batch < 0) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (lda < MaxInt(1, n) || ldaInv < MaxInt(1, n)) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (n == 0 || batch == 0) {
        return ACLBLAS_STATUS_SUCCESS;
    }
    if (A == nullptr || Ainv == nullptr || info == nullptr) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    return ACLBLAS_STATUS_SUCCESS;
}

}  // namespace
}  // namespace aclblas

extern "C" aclblasStatus_t aclblasSmatinvBatched(
    aclblasHandle_t handle, int n, const float* const A[], int lda, float* const Ainv[], int lda_inv, int* info,
    int batchSize)
{
    using namespace aclblas;
    if (handle == nullptr) {
        return ACLBLAS_STATUS_HANDLE_IS_NULLPTR;
    }
    aclblasStatus_t st = ValidateMatinv(n, lda, lda_inv, batchSize, A, Ainv, info);
    if (st != ACLBLAS_STATUS_SUCCESS) {
        LogError("aclblasSmatinvBatched", "invalid argument (n=%d lda=%d lda_inv=%d batch=%d)", n, lda, lda_inv,
                 batchSize);
        return st;
    }
    if (n == 0 || batchSize == 0) {
        return ACLBLAS_STATUS_SUCCESS;
    }
    if (!HandleHasDevice(handle)) {
        return ACLBLAS_STATUS_EXECUTION_FAILED;
    }
    // workspace: pointer table, packed n x n blocks, pivots (reference
    // matinv_batched_host.cpp:95-121)
    const size_t batch = static_cast<size_t>(batchSize);
    const size_t tableBytes = batch * sizeof(float*);
    const size_t blockBytes = static_cast<size_t>(n) * static_cast<size_t>(n) * batch * sizeof(float);
    const size_t pivBytes = static_cast<size_t>(n) * batch * sizeof(int);
    const size_t need = tableBytes + blockBytes + pivBytes;
    if (handle->workspace == nullptr || need > handle->workspace_size) {
        LogError("aclblasSmatinvBatched",
                 "workspace of %zu bytes needed, %zu available: call aclblasSetWorkspace with a larger buffer", need,
                 handle->workspace == nullptr ? static_cast<size_t>(0) : handle->workspace_size);
        return ACLBLAS_STATUS_EXECUTION_FAILED;
    }
    MarkWorkspaceUse(handle);
    auto* ws = static_cast<unsigned char*>(handle->workspace);
    float** table = reinterpret_cast<float**>(ws);
    float* blocks = reinterpret_cast<float*>(ws + tableBytes);
    int* piv = reinterpret_cast<int*>(ws + tableBytes + blockBytes);

    st = GetrfOutOfPlaceLaunch(handle, n, A, lda, blocks, table, piv, info, batchSize);
    if (st != ACLBLAS_STATUS_SUCCESS) {
        return st;
    }
    return GetriBatchedLaunch(handle->stream, n, table, n, piv, Ainv, lda_inv, info, batchSize);
}
