// aclblasSgemmGroupedBatched: FP32 GEMMs of several shapes in one launch.
//
// Reference: blas/gemm_grouped_batched/arch35/gemm_grouped_batched_host.cpp
// (argument checks :155-290 and :448-477 in that order; inactive groups
// groupSize == 0 || m == 0 || n == 0 skipped :491; k == 0 or alpha == 0 is a
// beta-only update :404-424; host pointer arrays :568-575) and the golden
// semantics of test/gemm_grouped_batched/gemm_grouped_batched_golden.h
// (alpha == 0: C = beta * C element by element; otherwise sgemm).
//
// MI355X design.  The host flattens every active (group, batch member) into a
// problem record with its own tile grid, orders the records by decreasing k
// (longest tiles first) and copies the table to the workspace through the
// handle's pinned staging buffer (one asynchronous copy, no synchronisation).
// One persistent launch of the FP32 MFMA tile kernel then strides over the
// concatenated tile lists; its tile edge (64 or 128) and workgroups per compute
// unit come from the tuning table for the call's total tile count (built-in:
// 128, two per compute unit).
#include <algorithm>
#include <vector>

#include "cann_ops_blas.h"
#include "gemm_kernel.h"
#include "internal.h"
#include "tuning.h"

namespace aclblas {
namespace {

using gemm::GroupedProblem;

constexpr int kDefaultTile = 128;
constexpr int kDefaultBlocksPerCu = 2;

bool ValidTrans(aclblasOperation_t t) { return t == ACLBLAS_OP_N || t == ACLBLAS_OP_T; }

aclblasStatus_t Validate(aclblasHandle_t handle, int groupCount, const aclblasOperation_t* ta,
                         const aclblasOperation_t* tb, const int* m, const int* n, const int* k, const float* alpha,
                         const float* const* A, const int* lda, const float* const* B, const int* ldb,
                         const float* beta, float* const* C, const int* ldc, const int* groupSize, long long& total)
{
    total = 0;
    if (handle == nullptr) {
        return ACLBLAS_STATUS_HANDLE_IS_NULLPTR;
    }
    if (groupCount < 0) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (groupCount == 0) {
        return ACLBLAS_STATUS_SUCCESS;
    }
    if (ta == nullptr || tb == nullptr || m == nullptr || n == nullptr || k == nullptr || alpha == nullptr ||
        beta == nullptr || lda == nullptr || ldb == nullptr || ldc == nullptr || groupSize == nullptr ||
        A == nullptr || B == nullptr || C == nullptr) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    for (int g = 0; g < groupCount; ++g) {
        if (m[g] < 0 || n[g] < 0 || k[g] < 0 || groupSize[g] < 0) {
            return ACLBLAS_STATUS_INVALID_VALUE;
        }
        if (groupSize[g] == 0) {
            continue;
        }
        if (!ValidTrans(ta[g]) || !ValidTrans(tb[g])) {
            return ACLBLAS_STATUS_INVALID_ENUM;
        }
        const int ldaMin = ta[g] == ACLBLAS_OP_T ? MaxInt(1, k[g]) : MaxInt(1, m[g]);
        const int ldbMin = tb[g] == ACLBLAS_OP_T ? MaxInt(1, n[g]) : MaxInt(1, k[g]);
        if (lda[g] < ldaMin || ldb[g] < ldbMin || ldc[g] < MaxInt(1, m[g])) {
            return ACLBLAS_STATUS_INVALID_VALUE;
        }
        total += groupSize[g];
    }
    return ACLBLAS_STATUS_SUCCESS;
}

bool Aligned16(const void* p, int ld) { return reinterpret_cast<uintptr_t>(p) % 16 == 0 && ld % 4 == 0; }

}  // namespace
}  // namespace aclblas

extern "C" aclblasStatus_t aclblasSgemmGroupedBatched(
    aclblasHandle_t handle, int groupCount, const aclblasOperation_t* transaArray,
    const aclblasOperation_t* transbArray, const int* mArray, const int* nArray, const int* kArray,
    const float* alphaArray, const float* const* Aarray, const int* ldaArray, const float* const* Barray,
    const int* ldbArray, const float* betaArray, float* const* Carray, const int* ldcArray,
    const int* groupSizeArray)
{
    using namespace aclblas;
    long long total = 0;
    aclblasStatus_t st = Validate(handle, groupCount, transaArray, transbArray, mArray, nArray, kArray, alphaArray,
                                  Aarray, ldaArray, Barray, ldbArray, betaArray, Carray, ldcArray, groupSizeArray,
                                  total);
    if (st != ACLBLAS_STATUS_SUCCESS) {
        LogError("aclblasSgemmGroupedBatched", "invalid argument (status %d)", static_cast<int>(st));
        return st;
    }
    if (total == 0) {
        return ACLBLAS_STATUS_SUCCESS;
    }

    // flatten the active problems (pointer arrays are host arrays, indexed by
    // the running batch position over all groups)
    std::vector<GroupedProblem> probs;
    probs.reserve(static_cast<size_t>(total));
    long long slot = 0;
    for (int g = 0; g < groupCount; ++g) {
        const int size = groupSizeArray[g];
        const bool inactive = size == 0 || mArray[g] == 0 || nArray[g] == 0;
        const float alpha = alphaArray[g];
        const float beta = betaArray[g];
        for (int i = 0; i < size; ++i, ++slot) {
            if (inactive || (alpha == 0.0f && beta == 1.0f)) {
                continue;
            }
            GroupedProblem p{};
            p.A = Aarray[slot];
            p.B = Barray[slot];
            p.C = Carray[slot];
            p.m = mArray[g];
            p.n = nArray[g];
            // alpha == 0: C = beta * C without touching A and B
            p.k = alpha == 0.0f ? 0 : kArray[g];
            p.lda = ldaArray[g];
            p.ldb = ldbArray[g];
            p.ldc = ldcArray[g];
            p.transA = transaArray[g] == ACLBLAS_OP_T ? 1 : 0;
            p.transB = transbArray[g] == ACLBLAS_OP_T ? 1 : 0;
            p.alpha = alpha;
            p.beta = beta;
            p.vecA = Aligned16(p.A, p.lda) ? 1 : 0;
            p.vecB = Aligned16(p.B, p.ldb) ? 1 : 0;
            p.vecC = Aligned16(p.C, p.ldc) ? 1 : 0;
            if (p.k > 0 && (p.A == nullptr || p.B == nullptr)) {
                LogError("aclblasSgemmGroupedBatched", "null A/B matrix in group %d", g);
                return ACLBLAS_STATUS_INVALID_VALUE;
            }
            if (p.C == nullptr) {
                LogError("aclblasSgemmGroupedBatched", "null C matrix in group %d", g);
                return ACLBLAS_STATUS_INVALID_VALUE;
            }
            probs.push_back(p);
        }
    }
    if (probs.empty()) {
        return ACLBLAS_STATUS_SUCCESS;
    }
    if (!HandleHasDevice(handle)) {
        return ACLBLAS_STATUS_EXECUTION_FAILED;
    }
    std::stable_sort(probs.begin(), probs.end(),
                     [](const GroupedProblem& a, const GroupedProblem& b) { return a.k > b.k; });
    const int cus = handle->cu_count > 0 ? handle->cu_count : 256;
    long long tiles128 = 0;
    for (const GroupedProblem& p : probs) {
        tiles128 += static_cast<long long>((p.m + 127) / 128) * ((p.n + 127) / 128);
    }
    const tune::Choice c = tune::Get(handle, tune::kOpGrouped, tune::GroupedBucket(tiles128, cus),
                                     tune::Choice{kDefaultTile, kDefaultBlocksPerCu});
    const int tile = c.a;
    long long tiles = 0;
    for (GroupedProblem& p : probs) {
        p.tilesM = (p.m + tile - 1) / tile;
        p.tilesN = (p.n + tile - 1) / tile;
        p.tileStart = static_cast<int>(tiles);
        tiles += static_cast<long long>(p.tilesM) * p.tilesN;
        if (tiles > 0x7FFFFFFFll) {
            LogError("aclblasSgemmGroupedBatched", "more than 2^31 tiles");
            return ACLBLAS_STATUS_NOT_SUPPORTED;
        }
    }
    void* dev = nullptr;
    st = StageToWorkspace(handle, probs.data(), probs.size() * sizeof(GroupedProblem), &dev);
    if (st != ACLBLAS_STATUS_SUCCESS) {
        return st;
    }
    const long long cap = static_cast<long long>(cus) * c.b;
    const unsigned grid = static_cast<unsigned>(tiles < cap ? tiles : cap);
    const auto* table = static_cast<const GroupedProblem*>(dev);
    if (tile == 64) {
        hipLaunchKernelGGL((gemm::GroupedGemmKernel<gemm::kF32, gemm::kF32, gemm::kOutF32, 64, 64>), dim3(grid),
                           dim3(gemm::kThreads), 0, handle->stream, table, static_cast<int>(probs.size()),
                           static_cast<int>(tiles));
    } else {
        hipLaunchKernelGGL((gemm::GroupedGemmKernel<gemm::kF32, gemm::kF32, gemm::kOutF32, 128, 128>), dim3(grid),
                           dim3(gemm::kThreads), 0, handle->stream, table, static_cast<int>(probs.size()),
                           static_cast<int>(tiles));
    }
    return LaunchStatus("aclblasSgemmGroupedBatched");
}
