// aclblasSgemm, aclblasSgemmEx, aclblasGemmEx: Level-3 GEMM on MFMA.
//
// Reference: blas/gemm_ex/arch35/sgemm_ex_host.cpp:33-137 (SgemmEx argument
// checks), :227-265 (k == 0 / alpha == 0 cases), :404-426 (entry: m == 0 or
// n == 0 returns before any check); blas/gemm/arch35/gemm_ex_host.cpp:45-185
// (GemmEx dtype table and checks), :295-355 (k == 0 and alpha == 0 cases),
// :450-494 (entry).
//
// MI355X design: one fused kernel per call (gemm_kernel.h).  The reference
// writes op(A)op(B) to a workspace temporary and runs a second alpha/beta pass
// whenever alpha != 1 or beta != 0; here alpha and beta are applied in the
// MFMA kernel's epilogue, so no workspace is needed and C is read at most once.
#include "gemm_kernel.h"
#include "drivers.h"
#include "internal.h"
#include "tuning.h"

namespace aclblas {
namespace {

using namespace gemm;

// C[:m, :n] = beta * C (beta == 0 writes zeros without reading C).
template <int OK>
__global__ __launch_bounds__(256) void ScaleCKernel(void* C, int m, int n, int ldc, float beta)
{
    const long long total = static_cast<long long>(m) * n;
    for (long long e = blockIdx.x * 256ll + threadIdx.x; e < total; e += 256ll * gridDim.x) {
        const int j = static_cast<int>(e / m);
        const int i = static_cast<int>(e - static_cast<long long>(j) * m);
        const size_t idx = static_cast<size_t>(j) * ldc + i;
        const float v = beta == 0.0f ? 0.0f : beta * LoadC<OK>(C, idx);
        StoreC<OK>(C, idx, v);
    }
}

template <int OK>
void LaunchScaleC(hipStream_t st, void* C, int m, int n, int ldc, float beta)
{
    const long long total = static_cast<long long>(m) * n;
    long long blocks = (total + 255) / 256;
    blocks = blocks > 4096 ? 4096 : blocks;
    hipLaunchKernelGGL((ScaleCKernel<OK>), dim3(static_cast<unsigned>(blocks)), dim3(256), 0, st, C, m, n, ldc,
                       beta);
}

void LaunchScale(hipStream_t st, int outKind, void* C, int m, int n, int ldc, float beta)
{
    switch (outKind) {
        case kOutF32: LaunchScaleC<kOutF32>(st, C, m, n, ldc, beta); break;
        case kOutBF16: LaunchScaleC<kOutBF16>(st, C, m, n, ldc, beta); break;
        default: LaunchScaleC<kOutF16>(st, C, m, n, ldc, beta); break;
    }
}

// C_b[:m, :n] = beta * C_b for every problem of a batch (pointer array).
// Grid: perProblem blocks for each problem, problems along x (no 65535 limit).
template <int OK>
__global__ __launch_bounds__(256) void ScaleCBatchedKernel(void* const* Carr, int m, int n, int ldc, float beta,
                                                            int perProblem)
{
    const int prob = blockIdx.x / perProblem;
    const int part = blockIdx.x - prob * perProblem;
    void* C = Carr[prob];
    const long long total = static_cast<long long>(m) * n;
    for (long long e = part * 256ll + threadIdx.x; e < total; e += 256ll * perProblem) {
        const int j = static_cast<int>(e / m);
        const int i = static_cast<int>(e - static_cast<long long>(j) * m);
        const size_t idx = static_cast<size_t>(j) * ldc + i;
        const float v = beta == 0.0f ? 0.0f : beta * LoadC<OK>(C, idx);
        StoreC<OK>(C, idx, v);
    }
}

void LaunchScaleBatched(hipStream_t st, int outKind, void* const* Carr, int m, int n, int ldc, float beta, int batch)
{
    const long long total = static_cast<long long>(m) * n;
    long long bx = (total + 255) / 256;
    bx = bx > 64 ? 64 : bx;
    const long long maxPer = 0x7FFFFFFFll / (batch > 0 ? batch : 1);
    bx = bx > maxPer ? (maxPer > 0 ? maxPer : 1) : bx;
    const int per = static_cast<int>(bx);
    const dim3 grid(static_cast<unsigned>(bx * batch));
    switch (outKind) {
        case kOutF32:
            hipLaunchKernelGGL((ScaleCBatchedKernel<kOutF32>), grid, dim3(256), 0, st, Carr, m, n, ldc, beta, per);
            break;
        case kOutBF16:
            hipLaunchKernelGGL((ScaleCBatchedKernel<kOutBF16>), grid, dim3(256), 0, st, Carr, m, n, ldc, beta, per);
            break;
        default:
            hipLaunchKernelGGL((ScaleCBatchedKernel<kOutF16>), grid, dim3(256), 0, st, Carr, m, n, ldc, beta, per);
            break;
    }
}

bool Aligned16(const void* p, int ld, int eb)
{
    return (reinterpret_cast<uintptr_t>(p) % 16 == 0) && ((static_cast<long long>(ld) * eb) % 16 == 0);
}

int KindOf(aclDataType t)
{
    switch (t) {
        case ACL_FLOAT: return kF32;
        case ACL_BF16: return kBF16;
        case ACL_FLOAT16: return kF16;
        case ACL_FLOAT8_E4M3FN: return kE4M3;
        case ACL_FLOAT8_E5M2: return kE5M2;
        default: return -1;
    }
}

int OutKindOf(aclDataType t)
{
    switch (t) {
        case ACL_FLOAT: return kOutF32;
        case ACL_BF16: return kOutBF16;
        default: return kOutF16;
    }
}

int ElemBytes(int kind) { return kind == kF32 ? 4 : (kind == kBF16 || kind == kF16) ? 2 : 1; }

// Launches the fused GEMM for a validated, non-degenerate problem (one
// problem when Aarr is null, otherwise `batch` problems through device
// pointer arrays).
aclblasStatus_t RunGemm(const _aclblas_handle* h, int ka, int kb, int ok, bool transA, bool transB, int m, int n,
                        int k, float alpha, const void* A, int lda, const void* B, int ldb, float beta, void* C,
                        int ldc, const float* scaleA, const float* scaleB, const void* const* Aarr = nullptr,
                        const void* const* Barr = nullptr, void* const* Carr = nullptr, int batch = 1)
{
    GemmArgs p{};
    p.A = A;
    p.B = B;
    p.C = C;
    p.Aarr = Aarr;
    p.Barr = Barr;
    p.Carr = Carr;
    p.m = m;
    p.n = n;
    p.k = k;
    p.lda = lda;
    p.ldb = ldb;
    p.ldc = ldc;
    p.transA = transA ? 1 : 0;
    p.transB = transB ? 1 : 0;
    p.alpha = alpha;
    p.beta = beta;
    p.scaleA = scaleA;
    p.scaleB = scaleB;
    // alignment of batched pointers is re-checked per problem on the device
    const void* a0 = Aarr != nullptr ? nullptr : A;
    const void* b0 = Barr != nullptr ? nullptr : B;
    const void* c0 = Carr != nullptr ? nullptr : C;
    p.vecA = Aligned16(a0, lda, ElemBytes(ka)) ? 1 : 0;
    p.vecB = Aligned16(b0, ldb, ElemBytes(kb)) ? 1 : 0;
    {
        const int cb = ok == kOutF32 ? 4 : 2;  // bytes per C element; four move together
        p.vecC = (reinterpret_cast<uintptr_t>(c0) % (4 * cb) == 0 && (static_cast<long long>(ldc) * cb) % (4 * cb) == 0)
                     ? 1 : 0;
    }
    const hipStream_t st = h->stream;
    const int cus = h->cu_count > 0 ? h->cu_count : 256;
    const int eb = ElemBytes(ka);
    const int cls = eb == 4 ? tune::kClassF32 : (eb == 2 ? tune::kClass16 : tune::kClassF8);
    const long long t128 = static_cast<long long>((m + 127) / 128) * ((n + 127) / 128) * batch;
    const tune::Choice c = tune::Get(h, tune::kOpGemm, tune::GemmBucket(cls, t128, k, cus),
                                     tune::Choice{DefaultGemmTile(eb, m, n, batch, cus), 8});
    const int tile = c.a;
    const int group = c.b;
    if (ka == kF32 && kb == kF32 && ok == kOutF32) {
        LaunchGemmTyped<kF32, kF32, kOutF32>(st, p, batch, tile, group);
    } else if (ka == kBF16 && kb == kBF16 && ok == kOutBF16) {
        LaunchGemmTyped<kBF16, kBF16, kOutBF16>(st, p, batch, tile, group);
    } else if (ka == kBF16 && kb == kBF16 && ok == kOutF32) {
        LaunchGemmTyped<kBF16, kBF16, kOutF32>(st, p, batch, tile, group);
    } else if (ka == kF16 && kb == kF16 && ok == kOutF16) {
        LaunchGemmTyped<kF16, kF16, kOutF16>(st, p, batch, tile, group);
    } else if (ka == kF16 && kb == kF16 && ok == kOutF32) {
        LaunchGemmTyped<kF16, kF16, kOutF32>(st, p, batch, tile, group);
    } else if (ka == kE4M3 && kb == kE4M3) {
        LaunchGemmTyped<kE4M3, kE4M3, kOutF16>(st, p, batch, tile, group);
    } else if (ka == kE4M3 && kb == kE5M2) {
        LaunchGemmTyped<kE4M3, kE5M2, kOutF16>(st, p, batch, tile, group);
    } else if (ka == kE5M2 && kb == kE4M3) {
        LaunchGemmTyped<kE5M2, kE4M3, kOutF16>(st, p, batch, tile, group);
    } else if (ka == kE5M2 && kb == kE5M2) {
        LaunchGemmTyped<kE5M2, kE5M2, kOutF16>(st, p, batch, tile, group);
    } else {
        return ACLBLAS_STATUS_NOT_SUPPORTED;
    }
    return LaunchStatus("aclblasGemm");
}

// ---------------------------------------------------------------- SgemmEx
aclblasStatus_t ValidateSgemmEx(aclblasHandle_t handle, const aclblasOperation_t& ta, const aclblasOperation_t& tb,
                                int m, int n, int k, int lda, int ldb, int ldc, const float* alpha, const float* beta,
                                const aclblasGemmAlgo_t& algo, const float* A, const float* B, const float* C)
{
    if (handle == nullptr) {
        return ACLBLAS_STATUS_HANDLE_IS_NULLPTR;
    }
    if (m < 0 || n < 0 || k < 0 || !IsValidOp(ta) || !IsValidOp(tb)) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    const bool tA = ta != ACLBLAS_OP_N;
    const bool tB = tb != ACLBLAS_OP_N;
    if (lda < (tA ? MaxInt(1, k) : MaxInt(1, m)) || ldb < (tB ? MaxInt(1, n) : MaxInt(1, k)) ||
        ldc < MaxInt(1, m)) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (alpha == nullptr || beta == nullptr) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (EnumBits(algo) != ACLBLAS_GEMM_DEFAULT) {
        return ACLBLAS_STATUS_NOT_SUPPORTED;
    }
    if (k > 0 && (A == nullptr || B == nullptr)) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (*beta != 0.0f && C == nullptr) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    return ACLBLAS_STATUS_SUCCESS;
}

// ---------------------------------------------------------------- GemmEx
// Accepted (A, B, C, compute) combinations: the reference's table
// (gemm_ex_host.cpp:45-69) plus BF16 x BF16 -> FP32 and FP16 x FP16 -> FP32.
bool ValidDtypes(int at, int bt, int ct, int ctype)
{
    const bool fp8a = at == ACL_FLOAT8_E4M3FN || at == ACL_FLOAT8_E5M2;
    const bool fp8b = bt == ACL_FLOAT8_E4M3FN || bt == ACL_FLOAT8_E5M2;
    switch (ctype) {
        case ACLBLAS_COMPUTE_16F:
            return at == ACL_FLOAT16 && bt == ACL_FLOAT16 && ct == ACL_FLOAT16;
        case ACLBLAS_COMPUTE_32F:
            if (at == ACL_FLOAT16) return bt == ACL_FLOAT16 && (ct == ACL_FLOAT16 || ct == ACL_FLOAT);
            if (at == ACL_BF16) return bt == ACL_BF16 && (ct == ACL_BF16 || ct == ACL_FLOAT);
            if (at == ACL_FLOAT) return bt == ACL_FLOAT && ct == ACL_FLOAT;
            if (fp8a) return fp8b && ct == ACL_FLOAT16;
            return false;
        default:
            return false;
    }
}

aclblasStatus_t ValidateGemmEx(aclblasHandle_t handle, const aclblasOperation_t& ta, const aclblasOperation_t& tb,
                               int m, int n, int k, int lda, int ldb, int ldc, const void* alpha, const void* beta,
                               const aclDataType& at, const aclDataType& bt, const aclDataType& ct,
                               const aclblasComputeType_t& ctype, const aclblasGemmAlgo_t& algo, const void* A,
                               const void* B, const void* C)
{
    if (handle == nullptr) {
        return ACLBLAS_STATUS_HANDLE_IS_NULLPTR;
    }
    if (m < 0 || n < 0 || k < 0 || !IsValidOp(ta) || !IsValidOp(tb)) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    const bool tA = ta != ACLBLAS_OP_N;
    const bool tB = tb != ACLBLAS_OP_N;
    if (lda < (tA ? MaxInt(1, k) : MaxInt(1, m)) || ldb < (tB ? MaxInt(1, n) : MaxInt(1, k)) ||
        ldc < MaxInt(1, m)) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (alpha == nullptr || beta == nullptr) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (EnumBits(algo) != ACLBLAS_GEMM_DEFAULT) {
        return ACLBLAS_STATUS_NOT_SUPPORTED;
    }
    if (!ValidDtypes(EnumBits(at), EnumBits(bt), EnumBits(ct), EnumBits(ctype))) {
        return ACLBLAS_STATUS_NOT_SUPPORTED;
    }
    if (m == 0 || n == 0) {
        return ACLBLAS_STATUS_SUCCESS;
    }
    if (k > 0 && (A == nullptr || B == nullptr)) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (C == nullptr && *static_cast<const float*>(beta) != 0.0f) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    return ACLBLAS_STATUS_SUCCESS;
}

}  // namespace

// Internal entry shared with the solve drivers: validated FP32/typed GEMM on a
// stream, alpha/beta by value.
aclblasStatus_t GemmLaunch(const _aclblas_handle* h, aclDataType at, aclDataType bt, aclDataType ct, bool transA,
                           bool transB, int m, int n, int k, float alpha, const void* A, int lda, const void* B,
                           int ldb, float beta, void* C, int ldc, const float* scaleA, const float* scaleB)
{
    if (m == 0 || n == 0) {
        return ACLBLAS_STATUS_SUCCESS;
    }
    const int ok = OutKindOf(ct);
    if (k == 0 || alpha == 0.0f) {
        if (beta != 1.0f) {
            LaunchScale(h->stream, ok, C, m, n, ldc, beta);
            return LaunchStatus("aclblasGemm");
        }
        return ACLBLAS_STATUS_SUCCESS;
    }
    return RunGemm(h, KindOf(at), KindOf(bt), ok, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C,
                   ldc, scaleA, scaleB);
}

}  // namespace aclblas

using namespace aclblas;

namespace {

// aclblasSgemmEx and aclblasSgemm; enum arguments by reference (unvalidated).
aclblasStatus_t SgemmExImpl(aclblasHandle_t handle, const aclblasOperation_t& transA,
                            const aclblasOperation_t& transB, int m, int n, int k, const float* alpha, const float* A,
                            int lda, const float* B, int ldb, const float* beta, float* C, int ldc,
                            const aclblasGemmAlgo_t& algo)
{
    // reference sgemm_ex_host.cpp:415-417: quick return before any check
    if (m == 0 || n == 0) {
        return ACLBLAS_STATUS_SUCCESS;
    }
    aclblasStatus_t st =
        ValidateSgemmEx(handle, transA, transB, m, n, k, lda, ldb, ldc, alpha, beta, algo, A, B, C);
    if (st != ACLBLAS_STATUS_SUCCESS) {
        LogError("aclblasSgemmEx", "invalid argument (status %d)", static_cast<int>(st));
        return st;
    }
    const float a = *alpha;
    const float b = *beta;
    if (C == nullptr) {
        return ACLBLAS_STATUS_SUCCESS;  // beta == 0 here (checked above): nothing to write
    }
    if ((k == 0 || a == 0.0f) && b == 1.0f) {
        return ACLBLAS_STATUS_SUCCESS;
    }
    if (!HandleHasDevice(handle)) {
        return ACLBLAS_STATUS_EXECUTION_FAILED;
    }
    return GemmLaunch(handle, ACL_FLOAT, ACL_FLOAT, ACL_FLOAT, transA != ACLBLAS_OP_N,
                      transB != ACLBLAS_OP_N, m, n, k, a, A, lda, B, ldb, b, C, ldc, nullptr, nullptr);
}

}  // namespace

extern "C" aclblasStatus_t aclblasSgemmEx(
    aclblasHandle_t handle, aclblasOperation_t transA, aclblasOperation_t transB, int m, int n, int k,
    const float* alpha, const float* A, int lda, const float* B, int ldb, const float* beta, float* C, int ldc,
    aclblasGemmAlgo_t algo)
{
    return SgemmExImpl(handle, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, algo);
}

extern "C" aclblasStatus_t aclblasSgemm(
    aclblasHandle_t handle, aclblasOperation_t transA, aclblasOperation_t transB, int m, int n, int k,
    const float* alpha, const float* A, int lda, const float* B, int ldb, const float* beta, float* C, int ldc)
{
    const aclblasGemmAlgo_t algo = ACLBLAS_GEMM_DEFAULT;
    return SgemmExImpl(handle, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, algo);
}

extern "C" aclblasStatus_t aclblasGemmEx(
    aclblasHandle_t handle, aclblasOperation_t transA, aclblasOperation_t transB, int m, int n, int k,
    const void* alpha, const void* A, aclDataType Atype, int lda, const void* B, aclDataType Btype, int ldb,
    const void* beta, void* C, aclDataType Ctype, int ldc, aclblasComputeType_t computeType, aclblasGemmAlgo_t algo)
{
    aclblasStatus_t st = ValidateGemmEx(handle, transA, transB, m, n, k, lda, ldb, ldc, alpha, beta, Atype, Btype,
                                        Ctype, computeType, algo, A, B, C);
    if (st != ACLBLAS_STATUS_SUCCESS) {
        LogError("aclblasGemmEx", "invalid argument (status %d)", static_cast<int>(st));
        return st;
    }
    if (m == 0 || n == 0 || C == nullptr) {
        return ACLBLAS_STATUS_SUCCESS;
    }
    const float a = *static_cast<const float*>(alpha);
    const float b = *static_cast<const float*>(beta);
    if ((k == 0 || a == 0.0f) && b == 1.0f) {
        return ACLBLAS_STATUS_SUCCESS;
    }
    if (!HandleHasDevice(handle)) {
        return ACLBLAS_STATUS_EXECUTION_FAILED;
    }
    return GemmLaunch(handle, Atype, Btype, Ctype, transA != ACLBLAS_OP_N,
                      transB != ACLBLAS_OP_N, m, n, k,
                      a, A, lda, B, ldb, b, C, ldc, nullptr, nullptr);
}

namespace aclblas {
namespace {

// reference blas/gemm_batched_ex/arch35/gemm_batched_ex_host.cpp:87-118
bool ValidBatchedDtypes(int at, int bt, int ct, int ctype)
{
    const bool fp8a = at == ACL_FLOAT8_E4M3FN || at == ACL_FLOAT8_E5M2;
    const bool fp8b = bt == ACL_FLOAT8_E4M3FN || bt == ACL_FLOAT8_E5M2;
    if (at == ACL_FLOAT16 && bt == ACL_FLOAT16 && ct == ACL_FLOAT16) {
        return ctype == ACLBLAS_COMPUTE_16F || ctype == ACLBLAS_COMPUTE_32F;
    }
    if (ctype != ACLBLAS_COMPUTE_32F) {
        return false;
    }
    if (at == ACL_BF16) {
        return bt == ACL_BF16 && ct == ACL_BF16;
    }
    return fp8a && fp8b && ct == ACL_FLOAT16;
}

// reference gemm_batched_ex_host.cpp:155-244 (order: handle, alpha, beta,
// dimensions and batch count, leading dimensions, algo, dtypes, FP8 compute
// type, pointer arrays when batchCount > 0)
aclblasStatus_t ValidateGemmBatchedEx(aclblasHandle_t handle, const aclblasOperation_t& ta,
                                      const aclblasOperation_t& tb, int m, int n, int k, int lda, int ldb, int ldc,
                                      int batch, const void* alpha, const void* beta, const aclDataType& at,
                                      const aclDataType& bt, const aclDataType& ct, const aclblasComputeType_t& ctype,
                                      const aclblasGemmAlgo_t& algo, const void* A, const void* B, const void* C)
{
    if (handle == nullptr) {
        return ACLBLAS_STATUS_HANDLE_IS_NULLPTR;
    }
    if (alpha == nullptr || beta == nullptr) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (m < 0 || n < 0 || k < 0 || !IsValidOp(ta) || !IsValidOp(tb) || batch < 0) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    const bool tA = ta != ACLBLAS_OP_N;
    const bool tB = tb != ACLBLAS_OP_N;
    if (lda < (tA ? MaxInt(1, k) : MaxInt(1, m)) || ldb < (tB ? MaxInt(1, n) : MaxInt(1, k)) ||
        ldc < MaxInt(1, m)) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (EnumBits(algo) != ACLBLAS_GEMM_DEFAULT) {
        return ACLBLAS_STATUS_NOT_SUPPORTED;
    }
    if (!ValidBatchedDtypes(EnumBits(at), EnumBits(bt), EnumBits(ct), EnumBits(ctype))) {
        return ACLBLAS_STATUS_NOT_SUPPORTED;
    }
    if (batch > 0 && (A == nullptr || B == nullptr || C == nullptr)) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    return ACLBLAS_STATUS_SUCCESS;
}

// reference blas/gemm_batched/arch35/gemm_batched_host.cpp:53-97
aclblasStatus_t ValidateSgemmBatched(aclblasHandle_t handle, const aclblasOperation_t& ta,
                                     const aclblasOperation_t& tb, int m, int n, int k, int lda, int ldb, int ldc,
                                     int batch, const float* alpha, const float* beta, const void* A, const void* B,
                                     const void* C)
{
    if (handle == nullptr) {
        return ACLBLAS_STATUS_HANDLE_IS_NULLPTR;
    }
    if (alpha == nullptr || beta == nullptr) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (m < 0 || n < 0 || k < 0 || batch < 0 || !IsValidOp(ta) || !IsValidOp(tb)) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    const bool tA = ta != ACLBLAS_OP_N;
    const bool tB = tb != ACLBLAS_OP_N;
    if (lda < (tA ? MaxInt(1, k) : MaxInt(1, m)) || ldb < (tB ? MaxInt(1, n) : MaxInt(1, k)) ||
        ldc < MaxInt(1, m)) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (batch > 0 && (A == nullptr || B == nullptr || C == nullptr)) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    return ACLBLAS_STATUS_SUCCESS;
}

aclblasStatus_t RunBatched(aclblasHandle_t handle, aclDataType at, aclDataType bt, aclDataType ct, bool tA, bool tB,
                           int m, int n, int k, float alpha, const void* const* A, int lda, const void* const* B,
                           int ldb, float beta, void* const* C, int ldc, int batch)
{
    if (!HandleHasDevice(handle)) {
        return ACLBLAS_STATUS_EXECUTION_FAILED;
    }
    const int ok = OutKindOf(ct);
    if (k == 0 || alpha == 0.0f) {
        if (beta == 1.0f) {
            return ACLBLAS_STATUS_SUCCESS;
        }
        LaunchScaleBatched(handle->stream, ok, C, m, n, ldc, beta, batch);
        return LaunchStatus("aclblasGemmBatched");
    }
    return RunGemm(handle, KindOf(at), KindOf(bt), ok, tA, tB, m, n, k, alpha, nullptr,
                   lda, nullptr, ldb, beta, nullptr, ldc, nullptr, nullptr, A, B, C, batch);
}

}  // namespace

aclblasStatus_t SgemmBatchedLaunch(const _aclblas_handle* h, bool transA, bool transB, int m, int n, int k,
                                   float alpha, const float* const* A, int lda, const float* const* B, int ldb,
                                   float beta, float* const* C, int ldc, int batch)
{
    if (k == 0 || alpha == 0.0f) {
        if (beta == 1.0f) {
            return ACLBLAS_STATUS_SUCCESS;
        }
        LaunchScaleBatched(h->stream, kOutF32, reinterpret_cast<void* const*>(C), m, n, ldc, beta, batch);
        return LaunchStatus("aclblasSgemmBatched");
    }
    return RunGemm(h, kF32, kF32, kOutF32, transA, transB, m, n, k, alpha, nullptr, lda, nullptr, ldb, beta,
                   nullptr, ldc, nullptr, nullptr, reinterpret_cast<const void* const*>(A),
                   reinterpret_cast<const void* const*>(B), reinterpret_cast<void* const*>(C), batch);
}

}  // namespace aclblas

extern "C" aclblasStatus_t aclblasGemmBatchedEx(
    aclblasHandle_t handle, aclblasOperation_t transa, aclblasOperation_t transb, int m, int n, int k,
    const void* alpha, const void* const Aarray[], aclDataType Atype, int lda, const void* const Barray[],
    aclDataType Btype, int ldb, const void* beta, void* const Carray[], aclDataType Ctype, int ldc, int batchCount,
    aclblasComputeType_t computeType, aclblasGemmAlgo_t algo)
{
    aclblasStatus_t st = ValidateGemmBatchedEx(handle, transa, transb, m, n, k, lda, ldb, ldc, batchCount, alpha,
                                               beta, Atype, Btype, Ctype, computeType, algo, Aarray, Barray, Carray);
    if (st != ACLBLAS_STATUS_SUCCESS) {
        LogError("aclblasGemmBatchedEx", "invalid argument (status %d)", static_cast<int>(st));
        return st;
    }
    if (m == 0 || n == 0 || batchCount == 0) {
        return ACLBLAS_STATUS_SUCCESS;
    }
    return RunBatched(handle, Atype, Btype, Ctype, transa != ACLBLAS_OP_N, transb != ACLBLAS_OP_N, m, n, k,
                      *static_cast<const float*>(alpha), Aarray, lda, Barray, ldb, *static_cast<const float*>(beta),
                      Carray, ldc, batchCount);
}

extern "C" aclblasStatus_t aclblasSgemmBatched(
    aclblasHandle_t handle, aclblasOperation_t transa, aclblasOperation_t transb, int m, int n, int k,
    const float* alpha, const float* const Aarray[], int lda, const float* const Barray[], int ldb,
    const float* beta, float* const Carray[], int ldc, int batchCount)
{
    aclblasStatus_t st = ValidateSgemmBatched(handle, transa, transb, m, n, k, lda, ldb, ldc, batchCount, alpha, beta,
                                              Aarray, Barray, Carray);
    if (st != ACLBLAS_STATUS_SUCCESS) {
        LogError("aclblasSgemmBatched", "invalid argument (status %d)", static_cast<int>(st));
        return st;
    }
    if (m == 0 || n == 0 || batchCount == 0) {
        return ACLBLAS_STATUS_SUCCESS;
    }
    return RunBatched(handle, ACL_FLOAT, ACL_FLOAT, ACL_FLOAT, transa != ACLBLAS_OP_N, transb != ACLBLAS_OP_N, m, n,
                      k, *alpha, reinterpret_cast<const void* const*>(Aarray), lda,
                      reinterpret_cast<const void* const*>(Barray), ldb, *beta,
                      reinterpret_cast<void* const*>(Carray), ldc, batchCount);
}
