// aclblasSgetrsBatched and aclblasSgetriBatched: triangular solves with the LU
// factors of aclblasSgetrfBatched.
//
// Reference: blas/getrs_batched/arch35/sgetrs_batched_host.cpp:29-185 (argument
// checks with LAPACK-style host info, nrhs <= 256), sgetrs_batched_kernel.cpp:
// 83-410 (N: P, L^-1, U^-1; T/C: U^-T, L^-T, P^T; out-of-range pivots skipped);
// blas/getri_batched/arch35/getri_batched_host.cpp:27-127 and
// getri_batched_kernel.cpp:39-308 (C = I, P, L^-1, U^-1; info = j+1 for the zero
// u_jj with the largest j).
//
// MI355X design.  One workgroup per (system, chunk of right-hand sides).  The
// chunk of B lives in LDS for the whole solve, already permuted.  Each
// triangular solve runs in 64-row blocks: wave 0 solves the 64x64 diagonal
// block with one lane per row (the solved value broadcast by v_readlane, no
// barriers inside the block), then all waves apply the block to the remaining
// rows with coalesced column reads of the factor.  Element by element the
// operations and their order are those of column-oriented substitution
// (b_i = fma(-t_ij, x_j, b_i) for j in solve order, x_j = b_j / u_jj), so the
// result does not depend on the blocking.
#include "device_common.h"
#include "drivers.h"
#include "internal.h"
#include "lapack_kernels.h"

namespace aclblas {
namespace {

using dev::kWave;

constexpr int kBlk = 64;            // rows per diagonal block
constexpr int kChunk = 8;           // entries of a trailing row loaded together
constexpr int kMaxRc = 8;           // right-hand sides per workgroup
constexpr size_t kLdsBudget = 64 * 1024;

enum SolveMode { kModeN = 0, kModeT = 1, kModeInv = 2 };

// Sequential application of the LAPACK swap list to the identity, giving perm
// with (P b)[i] = b[perm[i]].  Pivot entries outside [1, n] are skipped, as in
// reference sgetrs_batched_kernel.cpp:95-98.  scratch holds n ints.
__device__ void BuildPerm(int* perm, int* scratch, const int* piv, int n)
{
    const int t = threadIdx.x;
    if (n <= kWave) {
        if (t < kWave) {
            const int myp = t < n ? piv[t] - 1 : -1;
            int src = t;
            for (int k = 0; k < n; ++k) {
                const int p = __builtin_amdgcn_readlane(myp, k);
                if (p >= 0 && p < n && p != k) {
                    const int sk = __builtin_amdgcn_readlane(src, k);
                    const int sp = __builtin_amdgcn_readlane(src, p);
                    if (t == k) {
                        src = sp;
                    }
                    if (t == p) {
                        src = sk;
                    }
                }
            }
            if (t < n) {
                perm[t] = src;
            }
        }
    } else {
        for (int i = t; i < n; i += blockDim.x) {
            perm[i] = i;
            scratch[i] = piv[i] - 1;
        }
        __syncthreads();
        if (t == 0) {
            for (int k = 0; k < n; ++k) {
                const int p = scratch[k];
                if (p >= 0 && p < n && p != k) {
                    const int x = perm[k];
                    perm[k] = perm[p];
                    perm[p] = x;
                }
            }
        }
    }
    __syncthreads();
}

// Forward substitution with a lower-triangular factor T (unit diagonal when
// UNIT), in 64-row blocks.  Tel(i, j) returns T[i][j] for i > j (and i == j
// when !UNIT).  bs is the LDS chunk, row-major [n][RC].
template <int RC, bool UNIT, typename Tel>
__device__ __forceinline__ void LowerSolve(float* bs, int n, Tel tel)
{
    const int t = threadIdx.x;
    const int lane = dev::LaneId();
    for (int j0 = 0; j0 < n; j0 += kBlk) {
        const int nb = min(kBlk, n - j0);
        if (t < kWave) {
            float tr[kBlk];
            float x[RC];
            const bool valid = lane < nb;
            // unconditional loads (the block's first diagonal element where
            // the entry is not needed), padding afterwards: the loads issue
            // together instead of one branch-and-wait each
#pragma unroll
            for (int jj = 0; jj < kBlk; ++jj) {
                const bool need = valid && jj <= lane && jj < nb;
                tr[jj] = tel(need ? j0 + lane : j0, need ? j0 + jj : j0);
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int jj = 0; jj < kBlk; ++jj) {
                tr[jj] = (valid && jj <= lane && jj < nb) ? tr[jj] : 0.0f;
            }
#pragma unroll
            for (int r = 0; r < RC; ++r) {
                x[r] = valid ? bs[(j0 + lane) * RC + r] : 0.0f;
            }
#pragma unroll
            for (int jj = 0; jj < kBlk; ++jj) {
                if (jj < nb) {
#pragma unroll
                    for (int r = 0; r < RC; ++r) {
                        if (!UNIT && lane == jj) {
                            x[r] = x[r] / tr[jj];
                        }
                        const float xv = dev::ReadLane(x[r], jj);
                        if (lane > jj) {
                            x[r] = fmaf(-tr[jj], xv, x[r]);
                        }
                    }
                }
            }
            if (valid) {
#pragma unroll
                for (int r = 0; r < RC; ++r) {
                    bs[(j0 + lane) * RC + r] = x[r];
                }
            }
        }
        __syncthreads();
        for (int i = j0 + nb + t; i < n; i += blockDim.x) {
            float acc[RC];
#pragma unroll
            for (int r = 0; r < RC; ++r) {
                acc[r] = bs[i * RC + r];
            }
            // eight entries of the row per batch of loads, FMAs in column order
            for (int jc = 0; jc < nb; jc += kChunk) {
                float tv[kChunk];
#pragma unroll
                for (int e = 0; e < kChunk; ++e) {
                    tv[e] = tel(i, j0 + (jc + e < nb ? jc + e : 0));
                }
#pragma unroll
                for (int e = 0; e < kChunk; ++e) {
                    if (jc + e < nb) {
#pragma unroll
                        for (int r = 0; r < RC; ++r) {
                            acc[r] = fmaf(-tv[e], bs[(j0 + jc + e) * RC + r], acc[r]);
                        }
                    }
                }
            }
#pragma unroll
            for (int r = 0; r < RC; ++r) {
                bs[i * RC + r] = acc[r];
            }
        }
        __syncthreads();
    }
}

// Backward substitution with an upper-triangular factor T (unit diagonal when
// UNIT): rows processed from the bottom, x_j applied in decreasing j.
template <int RC, bool UNIT, typename Tel>
__device__ __forceinline__ void UpperSolve(float* bs, int n, Tel tel)
{
    const int t = threadIdx.x;
    const int lane = dev::LaneId();
    for (int j1 = n; j1 > 0; j1 -= kBlk) {
        const int j0 = max(0, j1 - kBlk);
        const int nb = j1 - j0;
        if (t < kWave) {
            float tr[kBlk];
            float x[RC];
            const bool valid = lane < nb;
#pragma unroll
            for (int jj = 0; jj < kBlk; ++jj) {
                const bool need = valid && jj >= lane && jj < nb;
                tr[jj] = tel(need ? j0 + lane : j0, need ? j0 + jj : j0);
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int jj = 0; jj < kBlk; ++jj) {
                tr[jj] = (valid && jj >= lane && jj < nb) ? tr[jj] : 0.0f;
            }
#pragma unroll
            for (int r = 0; r < RC; ++r) {
                x[r] = valid ? bs[(j0 + lane) * RC + r] : 0.0f;
            }
#pragma unroll
            for (int jj = kBlk - 1; jj >= 0; --jj) {
                if (jj < nb) {
#pragma unroll
                    for (int r = 0; r < RC; ++r) {
                        if (!UNIT && lane == jj) {
                            x[r] = x[r] / tr[jj];
                        }
                        const float xv = dev::ReadLane(x[r], jj);
                        if (lane < jj) {
                            x[r] = fmaf(-tr[jj], xv, x[r]);
                        }
                    }
                }
            }
            if (valid) {
#pragma unroll
                for (int r = 0; r < RC; ++r) {
                    bs[(j0 + lane) * RC + r] = x[r];
                }
            }
        }
        __syncthreads();
        for (int i = t; i < j0; i += blockDim.x) {
            float acc[RC];
#pragma unroll
            for (int r = 0; r < RC; ++r) {
                acc[r] = bs[i * RC + r];
            }
            // eight entries per batch of loads, FMAs in decreasing column order
            for (int jc = nb - 1; jc >= 0; jc -= kChunk) {
                float tv[kChunk];
#pragma unroll
                for (int e = 0; e < kChunk; ++e) {
                    tv[e] = tel(i, j0 + (jc - e >= 0 ? jc - e : 0));
                }
#pragma unroll
                for (int e = 0; e < kChunk; ++e) {
                    if (jc - e >= 0) {
#pragma unroll
                        for (int r = 0; r < RC; ++r) {
                            acc[r] = fmaf(-tv[e], bs[(j0 + jc - e) * RC + r], acc[r]);
                        }
                    }
                }
            }
#pragma unroll
            for (int r = 0; r < RC; ++r) {
                bs[i * RC + r] = acc[r];
            }
        }
        __syncthreads();
    }
}

// grid.x = batch * chunks; chunk c covers right-hand sides [c*RC, c*RC+RC).
// MODE kModeN / kModeT: B in place.  kModeInv: C = A^-1 out of place (B = C),
// nrhs = n, info written per matrix.
template <int RC, int MODE>
__global__ __launch_bounds__(256) void SolveKernel(
    const float* const* __restrict__ Aarray, int n, int lda, const int* __restrict__ ipiv,
    float* const* __restrict__ Barray, int ldb, int nrhs, int chunks, int* __restrict__ info)
{
    extern __shared__ float smem[];
    int* perm = reinterpret_cast<int*>(smem);
    float* bs = smem + n;
    __shared__ int zeroDiag;

    const int b = blockIdx.x / chunks;
    const int chunk = blockIdx.x - b * chunks;
    const int r0 = chunk * RC;
    const int rc = min(RC, nrhs - r0);
    const auto A = dev::G(Aarray[b]);
    const auto B = dev::G(Barray[b]);
    const int t = threadIdx.x;

    if (ipiv != nullptr) {
        BuildPerm(perm, reinterpret_cast<int*>(bs), ipiv + static_cast<size_t>(b) * n, n);
    } else {
        for (int i = t; i < n; i += blockDim.x) {
            perm[i] = i;
        }
        __syncthreads();
    }

    // load the chunk; rows permuted for the N and inverse solves
    for (int r = 0; r < RC; ++r) {
        for (int i = t; i < n; i += blockDim.x) {
            float v = 0.0f;
            if (r < rc) {
                if (MODE == kModeN) {
                    v = B[perm[i] + static_cast<size_t>(r0 + r) * ldb];
                } else if (MODE == kModeT) {
                    v = B[i + static_cast<size_t>(r0 + r) * ldb];
                } else {
                    v = (perm[i] == r0 + r) ? 1.0f : 0.0f;
                }
            }
            bs[i * RC + r] = v;
        }
    }
    if (MODE == kModeInv && t == 0) {
        zeroDiag = 0;
    }
    __syncthreads();

    auto lowerUnitN = [A, lda](int i, int j) { return A[i + static_cast<size_t>(j) * lda]; };
    if (MODE == kModeT) {
        // U^T y = b (lower, non-unit), then L^T z = y (upper, unit)
        auto ut = [A, lda](int i, int j) { return A[j + static_cast<size_t>(i) * lda]; };
        LowerSolve<RC, false>(bs, n, ut);
        UpperSolve<RC, true>(bs, n, ut);
    } else {
        LowerSolve<RC, true>(bs, n, lowerUnitN);
        bool doUpper = true;
        if (MODE == kModeInv) {
            for (int j = t; j < n; j += blockDim.x) {
                if (A[j + static_cast<size_t>(j) * lda] == 0.0f) {
                    atomicMax(&zeroDiag, j + 1);
                }
            }
            __syncthreads();
            doUpper = zeroDiag == 0;
        }
        if (doUpper) {
            UpperSolve<RC, false>(bs, n, lowerUnitN);
        }
    }

    for (int r = 0; r < rc; ++r) {
        for (int i = t; i < n; i += blockDim.x) {
            const int row = (MODE == kModeT) ? perm[i] : i;
            B[row + static_cast<size_t>(r0 + r) * ldb] = bs[i * RC + r];
        }
    }
    if (MODE == kModeInv && chunk == 0 && t == 0) {
        info[b] = zeroDiag;
    }
}

// Right-hand sides per workgroup: up to 8, fewer when the LDS chunk would not
// fit in 64 KiB.  0 means n is too large.
int ChooseRc(int n, int nrhs)
{
    const size_t perRow = static_cast<size_t>(n) * sizeof(float);
    if (perRow * 2 > kLdsBudget) {
        return 0;
    }
    int rc = static_cast<int>((kLdsBudget - perRow) / perRow);
    rc = rc > kMaxRc ? kMaxRc : rc;
    return rc < nrhs ? rc : nrhs;
}

template <int MODE, int RC>
void LaunchSolveRc(hipStream_t st, const float* const* A, int n, int lda, const int* piv, float* const* B, int ldb,
                   int nrhs, int* info, int batch)
{
    const int chunks = (nrhs + RC - 1) / RC;
    const size_t lds = static_cast<size_t>(n) * sizeof(int) + static_cast<size_t>(n) * RC * sizeof(float);
    const int threads = n <= kWave ? kWave : 256;
    const dim3 grid(static_cast<unsigned>(static_cast<long long>(batch) * chunks));
    hipLaunchKernelGGL((SolveKernel<RC, MODE>), grid, dim3(threads), lds, st, A, n, lda, piv, B, ldb, nrhs, chunks,
                       info);
}

template <int MODE>
void LaunchSolve(hipStream_t st, int rc, const float* const* A, int n, int lda, const int* piv, float* const* B,
                 int ldb, int nrhs, int* info, int batch)
{
    switch (rc) {
        case 1: LaunchSolveRc<MODE, 1>(st, A, n, lda, piv, B, ldb, nrhs, info, batch); break;
        case 2: LaunchSolveRc<MODE, 2>(st, A, n, lda, piv, B, ldb, nrhs, info, batch); break;
        case 3: LaunchSolveRc<MODE, 3>(st, A, n, lda, piv, B, ldb, nrhs, info, batch); break;
        case 4: LaunchSolveRc<MODE, 4>(st, A, n, lda, piv, B, ldb, nrhs, info, batch); break;
        default: LaunchSolveRc<MODE, 8>(st, A, n, lda, piv, B, ldb, nrhs, info, batch); break;
    }
}

int RoundRc(int rc)
{
    if (rc >= 8) return 8;
    if (rc >= 4) return 4;
    return rc;
}

// reference sgetrs_batched_host.cpp:29-40: LAPACK argument positions
constexpr int kPosTrans = 2;
constexpr int kPosN = 3;
constexpr int kPosNrhs = 4;
constexpr int kPosA = 5;
constexpr int kPosLda = 6;
constexpr int kPosB = 8;
constexpr int kPosLdb = 9;
constexpr int kPosBatch = 11;
constexpr int kMaxNrhs = 256;

void SetInfo(int* info, int v)
{
    if (info != nullptr) {
        *info = v;
    }
}

// Argument checks in the reference's order (sgetrs_batched_host.cpp:49-95).
aclblasStatus_t ValidateGetrs(const aclblasOperation_t& trans, int n, int nrhs, const float* const* A, int lda,
                              float* const* B, int ldb, int batch, int* info)
{
    if (!IsValidOp(trans)) {
        SetInfo(info, -kPosTrans);
        return ACLBLAS_STATUS_INVALID_ENUM;
    }
    if (n < 0) {
        SetInfo(info, -kPosN);
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (nrhs < 0 || nrhs > kMaxNrhs) {
        SetInfo(info, -kPosNrhs);
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (n > 0 && nrhs > 0 && batch > 0 && A == nullptr) {
        SetInfo(info, -kPosA);
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (lda < MaxInt(1, n)) {
        SetInfo(info, -kPosLda);
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (n > 0 && nrhs > 0 && batch > 0 && B == nullptr) {
        SetInfo(info, -kPosB);
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (ldb < MaxInt(1, n)) {
        SetInfo(info, -kPosLdb);
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (batch < 0) {
        SetInfo(info, -kPosBatch);
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    return ACLBLAS_STATUS_SUCCESS;
}

// reference getri_batched_host.cpp:27-64
aclblasStatus_t ValidateGetri(int n, int lda, int ldc, int batch, const float* const* A, float* const* C,
                              const int* info)
{
    if (n < 0 || batch < 0) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (lda < MaxInt(1, n) || ldc < MaxInt(1, n)) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (n == 0 || batch == 0) {
        return ACLBLAS_STATUS_SUCCESS;
    }
    if (A == nullptr || C == nullptr || info == nullptr) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    return ACLBLAS_STATUS_SUCCESS;
}

}  // namespace

// Internal entry shared with the solve/refinement drivers.
aclblasStatus_t GetrsBatchedLaunch(hipStream_t st, bool trans, int n, int nrhs, const float* const* A, int lda,
                                   const int* piv, float* const* B, int ldb, int batch)
{
    if (LaunchGetrsSmall(st, trans, n, nrhs, A, lda, piv, B, ldb, batch)) {
        return LaunchStatus("aclblasSgetrsBatched");
    }
    const int rc = ChooseRc(n, nrhs);
    if (rc <= 0) {
        return ACLBLAS_STATUS_NOT_SUPPORTED;
    }
    if (trans) {
        LaunchSolve<kModeT>(st, RoundRc(rc), A, n, lda, piv, B, ldb, nrhs, nullptr, batch);
    } else {
        LaunchSolve<kModeN>(st, RoundRc(rc), A, n, lda, piv, B, ldb, nrhs, nullptr, batch);
    }
    return LaunchStatus("aclblasSgetrsBatched");
}

aclblasStatus_t GetriBatchedLaunch(hipStream_t st, int n, const float* const* A, int lda, const int* piv,
                                   float* const* C, int ldc, int* info, int batch)
{
    if (LaunchGetriSmall(st, n, A, lda, piv, C, ldc, info, batch)) {
        return LaunchStatus("aclblasSgetriBatched");
    }
    const int rc = ChooseRc(n, n);
    if (rc <= 0) {
        return ACLBLAS_STATUS_NOT_SUPPORTED;
    }
    LaunchSolve<kModeInv>(st, RoundRc(rc), A, n, lda, piv, C, ldc, n, info, batch);
    return LaunchStatus("aclblasSgetriBatched");
}

}  // namespace aclblas

extern "C" aclblasStatus_t aclblasSgetrsBatched(
    aclblasHandle_t handle, aclblasOperation_t trans, int n, int nrhs, const float* const Aarray[], int lda,
    const int* devIpiv, float* const Barray[], int ldb, int* info, int batchCount)
{
    using namespace aclblas;
    if (handle == nullptr) {
        return ACLBLAS_STATUS_HANDLE_IS_NULLPTR;
    }
    aclblasStatus_t st = ValidateGetrs(trans, n, nrhs, Aarray, lda, Barray, ldb, batchCount, info);
    if (st != ACLBLAS_STATUS_SUCCESS) {
        LogError("aclblasSgetrsBatched", "invalid argument (status %d)", static_cast<int>(st));
        return st;
    }
    SetInfo(info, 0);
    if (n == 0 || nrhs == 0 || batchCount == 0) {
        return ACLBLAS_STATUS_SUCCESS;
    }
    if (!HandleHasDevice(handle)) {
        return ACLBLAS_STATUS_EXECUTION_FAILED;
    }
    return GetrsBatchedLaunch(handle->stream, trans != ACLBLAS_OP_N, n, nrhs, Aarray, lda, devIpiv, Barray, ldb,
                              batchCount);
}

extern "C" aclblasStatus_t aclblasSgetriBatched(
    aclblasHandle_t handle, int n, const float* const Aarray[], int lda, const int* PivotArray,
    float* const Carray[], int ldc, int* infoArray, int batchSize)
{
    using namespace aclblas;
    if (handle == nullptr) {
        return ACLBLAS_STATUS_HANDLE_IS_NULLPTR;
    }
    aclblasStatus_t st = ValidateGetri(n, lda, ldc, batchSize, Aarray, Carray, infoArray);
    if (st != ACLBLAS_STATUS_SUCCESS) {
        LogError("aclblasSgetriBatched", "invalid argument (n=%d lda=%d ldc=%d batch=%d)", n, lda, ldc, batchSize);
        return st;
    }
    if (n == 0 || batchSize == 0) {
        return ACLBLAS_STATUS_SUCCESS;
    }
    if (ChooseRc(n, n) <= 0) {
        LogError("aclblasSgetriBatched", "n=%d is too large", n);
        return ACLBLAS_STATUS_NOT_SUPPORTED;
    }
    if (!HandleHasDevice(handle)) {
        return ACLBLAS_STATUS_EXECUTION_FAILED;
    }
    return GetriBatchedLaunch(handle->stream, n, Aarray, lda, PivotArray, Carray, ldc, infoArray, batchSize);
}
