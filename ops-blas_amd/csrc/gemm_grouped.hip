// aclblasSgemmGroupedBatched: FP32 GEMMs of several shapes in one launch;
// aclblasGemmGroupedBatchedEx (FP16 / BF16 / FP8, device pointer arrays) at the
// end of the file, on the same persistent kernel.
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
#include <cmath>
#include <cstring>
#include <limits>
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

bool ValidTrans(const aclblasOperation_t& t)
{
    const int v = EnumBits(t);
    return v == ACLBLAS_OP_N || v == ACLBLAS_OP_T;
}

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
    const tune::Choice c = tune::Get(handle, tune::kOpGrouped, tune::GroupedBucket(0, tiles128, cus),
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
                           static_cast<int>(tiles), nullptr, nullptr, nullptr);
    } else {
        hipLaunchKernelGGL((gemm::GroupedGemmKernel<gemm::kF32, gemm::kF32, gemm::kOutF32, 128, 128>), dim3(grid),
                           dim3(gemm::kThreads), 0, handle->stream, table, static_cast<int>(probs.size()),
                           static_cast<int>(tiles), nullptr, nullptr, nullptr);
    }
    return LaunchStatus("aclblasSgemmGroupedBatched");
}

// ------------------------------------------------------ 16-bit / FP8 groups
namespace aclblas {
namespace {

// IEEE binary16 bits -> float (exact), for alpha / beta given under COMPUTE_16F.
float HostHalfToFloat(uint16_t h)
{
    const uint32_t sign = static_cast<uint32_t>(h & 0x8000u) << 16;
    const int exp = (h >> 10) & 0x1F;
    const uint32_t frac = h & 0x3FFu;
    float mag;
    if (exp == 0) {
        mag = std::ldexp(static_cast<float>(frac), -24);
    } else if (exp == 31) {
        mag = frac == 0 ? std::numeric_limits<float>::infinity() : std::numeric_limits<float>::quiet_NaN();
    } else {
        mag = std::ldexp(static_cast<float>(frac | 0x400u), exp - 25);
    }
    uint32_t bits;
    std::memcpy(&bits, &mag, sizeof(bits));
    bits |= sign;
    float out;
    std::memcpy(&out, &bits, sizeof(out));
    return out;
}

bool IsFp8Type(int t) { return t == ACL_FLOAT8_E4M3FN || t == ACL_FLOAT8_E5M2; }

// reference gemm_grouped_batched_ex_host.cpp:59-72
bool GroupedExTypesSupported(int at, int bt, int ct, int ctype)
{
    if (at == ACL_FLOAT16 && bt == ACL_FLOAT16 && ct == ACL_FLOAT16) {
        return ctype == ACLBLAS_COMPUTE_16F || ctype == ACLBLAS_COMPUTE_32F;
    }
    if (at == ACL_BF16 && bt == ACL_BF16 && ct == ACL_BF16) {
        return ctype == ACLBLAS_COMPUTE_32F;
    }
    return ct == ACL_FLOAT16 && ctype == ACLBLAS_COMPUTE_32F && IsFp8Type(at) && IsFp8Type(bt);
}

// reference gemm_grouped_batched_ex_host.cpp:115-175 (same checks, same order)
aclblasStatus_t ValidateGroupedEx(const aclblasOperation_t* ta, const aclblasOperation_t* tb, const int* m,
                                  const int* n, const int* k, const void* alpha, const int* lda, const int* ldb,
                                  const void* beta, const int* ldc, int groupCount, const int* groupSize)
{
    if (groupCount < 0) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (groupCount == 0) {
        return ACLBLAS_STATUS_SUCCESS;
    }
    if (ta == nullptr || tb == nullptr || m == nullptr || n == nullptr || k == nullptr || alpha == nullptr ||
        lda == nullptr || ldb == nullptr || beta == nullptr || ldc == nullptr || groupSize == nullptr) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    for (int g = 0; g < groupCount; ++g) {
        if (!IsValidOp(ta[g]) || !IsValidOp(tb[g]) || m[g] < 0 || n[g] < 0 || k[g] < 0 || groupSize[g] < 0) {
            return ACLBLAS_STATUS_INVALID_VALUE;
        }
        const int minLda = ta[g] == ACLBLAS_OP_N ? MaxInt(1, m[g]) : MaxInt(1, k[g]);
        const int minLdb = tb[g] == ACLBLAS_OP_N ? MaxInt(1, k[g]) : MaxInt(1, n[g]);
        if (lda[g] < minLda || ldb[g] < minLdb || ldc[g] < MaxInt(1, m[g])) {
            return ACLBLAS_STATUS_INVALID_VALUE;
        }
    }
    return ACLBLAS_STATUS_SUCCESS;
}


int ElemBytesOf(aclDataType t) { return IsFp8Type(t) ? 1 : 2; }

template <int KA, int KB, int OK>
void LaunchGroupedEx(hipStream_t st, int tile, unsigned grid, const GroupedProblem* table, int count, int tiles,
                     const void* const* A, const void* const* B, void* const* C)
{
    if (tile == 64) {
        hipLaunchKernelGGL((gemm::GroupedGemmKernel<KA, KB, OK, 64, 64>), dim3(grid), dim3(gemm::kThreads), 0, st,
                           table, count, tiles, A, B, C);
    } else {
        hipLaunchKernelGGL((gemm::GroupedGemmKernel<KA, KB, OK, 128, 128>), dim3(grid), dim3(gemm::kThreads), 0,
                           st, table, count, tiles, A, B, C);
    }
}

}  // namespace
}  // namespace aclblas

// Reference blas/gemm_grouped_batched_ex/arch35/gemm_grouped_batched_ex_host.cpp
// :356-414 (checks in that order) and gemm_grouped_batched_ex_epilogue_kernel.cpp
// :88-148 (FP32 products, alpha, + beta C, BF16 rounded to nearest even, FP16 to
// nearest with ties away from zero, no FP8 non-finite rule).
//
// MI355X design: as aclblasSgemmGroupedBatched, one record per problem (its
// group's shape and scalars and its slot in the caller's device pointer
// arrays, which the host never reads), ordered by decreasing k, staged to the
// workspace, and one persistent launch of the MFMA tile kernel for the element
// types that reads A, B and C of each record from the pointer arrays.  A group
// without products (k == 0 or alpha == 0) becomes records with k = 0 and
// alpha = 0, which store 0 + beta C (nothing for beta == 1).  SUCCESS returns
// before any device work when no record remains.  Asynchronous on the
// handle's stream (the reference synchronises before returning).
extern "C" aclblasStatus_t aclblasGemmGroupedBatchedEx(
    aclblasHandle_t handle, const aclblasOperation_t transaArray[], const aclblasOperation_t transbArray[],
    const int mArray[], const int nArray[], const int kArray[], const void* alphaArray, const void* const Aarray[],
    aclDataType Atype, const int ldaArray[], const void* const Barray[], aclDataType Btype, const int ldbArray[],
    const void* betaArray, void* const Carray[], aclDataType Ctype, const int ldcArray[], int groupCount,
    const int groupSize[], aclblasComputeType_t computeType)
{
    using namespace aclblas;
    if (handle == nullptr) {
        return ACLBLAS_STATUS_HANDLE_IS_NULLPTR;
    }
    if (!GroupedExTypesSupported(EnumBits(Atype), EnumBits(Btype), EnumBits(Ctype), EnumBits(computeType))) {
        LogError("aclblasGemmGroupedBatchedEx", "unsupported type combination A=%d B=%d C=%d compute=%d",
                 static_cast<int>(Atype), static_cast<int>(Btype), static_cast<int>(Ctype),
                 static_cast<int>(computeType));
        return ACLBLAS_STATUS_NOT_SUPPORTED;
    }
    aclblasStatus_t st = ValidateGroupedEx(transaArray, transbArray, mArray, nArray, kArray, alphaArray, ldaArray,
                                           ldbArray, betaArray, ldcArray, groupCount, groupSize);
    if (st != ACLBLAS_STATUS_SUCCESS) {
        LogError("aclblasGemmGroupedBatchedEx", "invalid argument");
        return st;
    }
    if (groupCount == 0) {
        return ACLBLAS_STATUS_SUCCESS;
    }
    long long problems = 0;
    for (int g = 0; g < groupCount; ++g) {
        problems += groupSize[g];
    }
    if (problems > 0 && (Aarray == nullptr || Barray == nullptr || Carray == nullptr)) {
        LogError("aclblasGemmGroupedBatchedEx", "matrix pointer arrays must not be NULL");
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (problems > 0x7FFFFFFFll) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    const bool half = computeType == ACLBLAS_COMPUTE_16F;
    const int ea = ElemBytesOf(Atype);
    const int eb = ElemBytesOf(Btype);
    std::vector<GroupedProblem> probs;
    long long slot = 0;
    for (int g = 0; g < groupCount; ++g) {
        const int size = groupSize[g];
        if (size > 0 && mArray[g] > 0 && nArray[g] > 0) {
            const float alpha = half ? HostHalfToFloat(static_cast<const uint16_t*>(alphaArray)[g])
                                     : static_cast<const float*>(alphaArray)[g];
            const float beta = half ? HostHalfToFloat(static_cast<const uint16_t*>(betaArray)[g])
                                    : static_cast<const float*>(betaArray)[g];
            const bool products = kArray[g] > 0 && alpha != 0.0f;
            if (products || beta != 1.0f) {
                for (int i = 0; i < size; ++i) {
                    GroupedProblem p{};
                    p.slot = static_cast<int>(slot + i);
                    p.m = mArray[g];
                    p.n = nArray[g];
                    p.k = products ? kArray[g] : 0;
                    p.lda = ldaArray[g];
                    p.ldb = ldbArray[g];
                    p.ldc = ldcArray[g];
                    p.transA = transaArray[g] == ACLBLAS_OP_N ? 0 : 1;
                    p.transB = transbArray[g] == ACLBLAS_OP_N ? 0 : 1;
                    p.alpha = products ? alpha : 0.0f;
                    p.beta = beta;
                    // leading-dimension part of the alignment test (pointers on the device)
                    p.vecA = (static_cast<long long>(p.lda) * ea) % 16 == 0 ? 1 : 0;
                    p.vecB = (static_cast<long long>(p.ldb) * eb) % 16 == 0 ? 1 : 0;
                    p.vecC = (static_cast<long long>(p.ldc) * 2) % 16 == 0 ? 1 : 0;
                    probs.push_back(p);
                }
            }
        }
        slot += size;
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
    const tune::Choice c = tune::Get(handle, tune::kOpGrouped, tune::GroupedBucket(1, tiles128, cus),
                                     tune::Choice{kDefaultTile, kDefaultBlocksPerCu});
    const int tile = c.a;
    long long tiles = 0;
    for (GroupedProblem& p : probs) {
        p.tilesM = (p.m + tile - 1) / tile;
        p.tilesN = (p.n + tile - 1) / tile;
        p.tileStart = static_cast<int>(tiles);
        tiles += static_cast<long long>(p.tilesM) * p.tilesN;
        if (tiles > 0x7FFFFFFFll) {
            LogError("aclblasGemmGroupedBatchedEx", "more than 2^31 tiles");
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
    const int count = static_cast<int>(probs.size());
    const int nt = static_cast<int>(tiles);
    const hipStream_t s = handle->stream;
    using namespace gemm;
    const bool a4 = Atype == ACL_FLOAT8_E4M3FN;
    const bool b4 = Btype == ACL_FLOAT8_E4M3FN;
    if (Atype == ACL_FLOAT16) {
        LaunchGroupedEx<kF16, kF16, kOutF16A>(s, tile, grid, table, count, nt, Aarray, Barray, Carray);
    } else if (Atype == ACL_BF16) {
        LaunchGroupedEx<kBF16, kBF16, kOutBF16>(s, tile, grid, table, count, nt, Aarray, Barray, Carray);
    } else if (a4 && b4) {
        LaunchGroupedEx<kE4M3, kE4M3, kOutF16A>(s, tile, grid, table, count, nt, Aarray, Barray, Carray);
    } else if (a4) {
        LaunchGroupedEx<kE4M3, kE5M2, kOutF16A>(s, tile, grid, table, count, nt, Aarray, Barray, Carray);
    } else if (b4) {
        LaunchGroupedEx<kE5M2, kE4M3, kOutF16A>(s, tile, grid, table, count, nt, Aarray, Barray, Carray);
    } else {
        LaunchGroupedEx<kE5M2, kE5M2, kOutF16A>(s, tile, grid, table, count, nt, Aarray, Barray, Carray);
    }
    return LaunchStatus("aclblasGemmGroupedBatchedEx");
}
