// aclblasStrsm: triangular solve with multiple right-hand sides, FP32,
// column-major: B := alpha op(A)^-1 B (left) or alpha B op(A)^-1 (right).
//
// Reference: blas/trsm/arch35/strsm_host.cpp:406-483 (entry, argument checks
// and their order, alpha == 0 writes zeros), :143-153 (zero path) and its
// blocked left path :216-404 (diagonal-panel solve, then a GEMM update of the
// rows still to be solved).
//
// MI355X design: one blocked algorithm for every side / uplo / trans case.
// op(A) is walked in 128 x 128 diagonal blocks in the order its triangle
// dictates (forward when the effective factor is lower, backward otherwise):
//   1. the diagonal block is solved against all right-hand sides by
//      TrsmDiagKernel: the block goes to LDS (diagonal taken as 1 for a unit
//      triangle), each lane owns one column (left side) or one row (right
//      side) of the block of B and runs the substitution in the operation
//      order of netlib strsm.f, so an order up to 128 (one block) rounds
//      exactly as the reference's golden;
//   2. the rest of B is updated by the library's FP32 MFMA GEMM
//      (aclblasSgemm, alpha -1, beta 1; operands addressed in place in A and
//      B, transposition through the GEMM's op flags).
// alpha != 1 is applied by one elementwise pass, before the solve or, for
// B op(A)^-T, after it (where strsm.f applies it).  No workspace,
// no allocation, no host synchronisation: the call captures into a graph.
#include "cann_ops_blas.h"
#include "device_common.h"
#include "internal.h"

namespace aclblas {
namespace {

constexpr int kTb = 128;         // diagonal block order
constexpr int kDiagThreads = 64;  // right-hand-side vectors per workgroup

// B[i + j ldb] *= alpha (or = 0 when zero), i < m, j < n.
__global__ __launch_bounds__(256) void TrsmScaleKernel(float* __restrict__ B, int m, int n, int ldb, float alpha,
                                                      int zero)
{
    const long long total = static_cast<long long>(m) * n;
    const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
    for (long long idx = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; idx < total; idx += stride) {
        const int j = static_cast<int>(idx / m);
        const int i = static_cast<int>(idx - static_cast<long long>(j) * m);
        float& v = B[i + static_cast<size_t>(j) * ldb];
        v = zero != 0 ? 0.0f : alpha * v;
    }
}

// Solves one diagonal block M y = b for `vectors` right-hand-side vectors.  M
// (kb x kb) is the block of op(A) for the left side and of op(A)^T for the
// right side: M(i, j) = A(k0 + i, k0 + j) when readDirect, else
// A(k0 + j, k0 + i); lower (forward) or upper.  Left side: vector = column c
// of B, b_i = B(k0 + i, c); right side: vector = row r, b_i = B(r, k0 + i).
// The substitution follows netlib strsm.f operation for operation: forward
// and descending-backward as column sweeps (each finished y_k subtracted
// from the rows still open), ascending-backward (asc) as inner products
// accumulated from index i + 1 upward; every multiply-subtract rounded
// separately; the diagonal divides (left) or multiplies by its reciprocal
// (right); a unit diagonal is not read.  One lane per vector, the vectors'
// values in LDS columns (conflict-free), the block in LDS.
template <bool LEFT>
__global__ __launch_bounds__(kDiagThreads) void TrsmDiagKernel(const float* __restrict__ A, int lda,
                                                             float* __restrict__ B, int ldb, int k0, int kb,
                                                             int vectors, int readDirect, int unit, int lowerM,
                                                             int asc)
{
    __shared__ float T[kTb][kTb + 1];
    __shared__ float ys[kTb][kDiagThreads];
    const int tid = threadIdx.x;
    for (int e = tid; e < kb * kb; e += kDiagThreads) {
        const int i = e % kb;
        const int j = e / kb;
        float v = 0.0f;
        if (i == j) {
            v = unit == 0 ? A[(k0 + i) + static_cast<size_t>(k0 + i) * lda] : 1.0f;
        } else if (lowerM != 0 ? i > j : i < j) {
            v = readDirect != 0 ? A[(k0 + i) + static_cast<size_t>(k0 + j) * lda]
                                : A[(k0 + j) + static_cast<size_t>(k0 + i) * lda];
        }
        T[i][j] = v;
    }
    const int vec = blockIdx.x * kDiagThreads + tid;
    const bool valid = vec < vectors;
    for (int i = 0; i < kb; ++i) {
        ys[i][tid] = !valid ? 0.0f
                            : (LEFT ? B[(k0 + i) + static_cast<size_t>(vec) * ldb]
                                    : B[vec + static_cast<size_t>(k0 + i) * ldb]);
    }
    __syncthreads();
    auto diag = [&](float y, int k) -> float {
        if (LEFT) {
            return y / T[k][k];
        }
        const float r = 1.0f / T[k][k];
        return r * y;
    };
    if (lowerM != 0) {
        for (int k = 0; k < kb; ++k) {
            const float yk = diag(ys[k][tid], k);
            ys[k][tid] = yk;
            for (int i = k + 1; i < kb; ++i) {
                ys[i][tid] = ys[i][tid] - T[i][k] * yk;
            }
        }
    } else if (asc != 0) {
        for (int i = kb - 1; i >= 0; --i) {
            float t = ys[i][tid];
            for (int k = i + 1; k < kb; ++k) {
                t = t - T[i][k] * ys[k][tid];
            }
            ys[i][tid] = diag(t, i);
        }
    } else {
        for (int k = kb - 1; k >= 0; --k) {
            const float yk = diag(ys[k][tid], k);
            ys[k][tid] = yk;
            for (int i = 0; i < k; ++i) {
                ys[i][tid] = ys[i][tid] - T[i][k] * yk;
            }
        }
    }
    if (valid) {
        for (int i = 0; i < kb; ++i) {
            if (LEFT) {
                B[(k0 + i) + static_cast<size_t>(vec) * ldb] = ys[i][tid];
            } else {
                B[vec + static_cast<size_t>(k0 + i) * ldb] = ys[i][tid];
            }
        }
    }
}

void LaunchDiag(hipStream_t st, bool left, const float* A, int lda, float* B, int ldb, int k0, int kb, int vectors,
                int readDirect, int unit, int lowerM, int asc)
{
    const dim3 grid(static_cast<unsigned>((vectors + kDiagThreads - 1) / kDiagThreads));
    const dim3 block(kDiagThreads);
    if (left) {
        hipLaunchKernelGGL(TrsmDiagKernel<true>, grid, block, 0, st, A, lda, B, ldb, k0, kb, vectors, readDirect, unit,
                           lowerM, asc);
    } else {
        hipLaunchKernelGGL(TrsmDiagKernel<false>, grid, block, 0, st, A, lda, B, ldb, k0, kb, vectors, readDirect,
                           unit, lowerM, asc);
    }
}

void LaunchScale(hipStream_t st, float* B, int m, int n, int ldb, float alpha, bool zero)
{
    const long long total = static_cast<long long>(m) * n;
    long long blocks = (total + 255) / 256;
    blocks = blocks > 4096 ? 4096 : (blocks < 1 ? 1 : blocks);
    hipLaunchKernelGGL(TrsmScaleKernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, st, B, m, n, ldb, alpha,
                       zero ? 1 : 0);
}

}  // namespace
}  // namespace aclblas

extern "C" aclblasStatus_t aclblasStrsm(aclblasHandle_t handle, aclblasSideMode_t side, aclblasFillMode_t uplo,
                                        aclblasOperation_t trans, aclblasDiagType_t diag, int m, int n,
                                        const float* alpha, const float* A, int lda, float* B, int ldb)
{
    using namespace aclblas;
    // argument checks in the reference's order (strsm_host.cpp:406-452)
    if (handle == nullptr) {
        return ACLBLAS_STATUS_HANDLE_IS_NULLPTR;
    }
    if (EnumBits(side) != ACLBLAS_SIDE_LEFT && EnumBits(side) != ACLBLAS_SIDE_RIGHT) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (m < 0 || n < 0) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if ((EnumBits(uplo) != ACLBLAS_UPPER && EnumBits(uplo) != ACLBLAS_LOWER) || !IsValidOp(trans) ||
        (EnumBits(diag) != ACLBLAS_NON_UNIT && EnumBits(diag) != ACLBLAS_UNIT) || alpha == nullptr) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (m == 0 || n == 0) {
        return ACLBLAS_STATUS_SUCCESS;
    }
    if (B == nullptr) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    const float a = *alpha;
    if (a != 0.0f && A == nullptr) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    const int k = side == ACLBLAS_SIDE_LEFT ? m : n;
    if (lda < MaxInt(1, k) || ldb < MaxInt(1, m)) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (!HandleHasDevice(handle)) {
        return ACLBLAS_STATUS_EXECUTION_FAILED;
    }
    hipStream_t st = handle->stream;
    if (a == 0.0f) {
        LaunchScale(st, B, m, n, ldb, 0.0f, true);
        return LaunchStatus("aclblasStrsm");
    }
    const bool upper = uplo == ACLBLAS_UPPER;
    const bool tr = trans != ACLBLAS_OP_N;
    const bool left = side == ACLBLAS_SIDE_LEFT;
    // strsm.f scales B by alpha before the solve, except for B op(A)^-T where
    // each finished column is scaled after it has been used
    const bool scaleAfter = !left && tr;
    if (a != 1.0f && !scaleAfter) {
        LaunchScale(st, B, m, n, ldb, a, false);
    }
    const int unit = diag == ACLBLAS_UNIT ? 1 : 0;
    // Left: op(A) lower <=> forward.  Right: op(A)^T lower <=> forward.
    const bool opLower = upper == tr;
    const bool forward = left ? opLower : !opLower;
    // M(i, j) read directly from A when M is A itself: left N, right T.
    const int readDirect = (left != tr) ? 1 : 0;
    // backward solves whose strsm.f loop accumulates inner products upward
    const int asc = (!forward && left == tr) ? 1 : 0;
    const int vectors = left ? n : m;
    const int blocks = (k + kTb - 1) / kTb;
    const float minus1 = -1.0f;
    const float one = 1.0f;
    for (int s = 0; s < blocks; ++s) {
        const int bi = forward ? s : blocks - 1 - s;
        const int k0 = bi * kTb;
        const int kb = k - k0 < kTb ? k - k0 : kTb;
        LaunchDiag(st, left, A, lda, B, ldb, k0, kb, vectors, readDirect, unit, forward ? 1 : 0, asc);
        aclblasStatus_t rs = LaunchStatus("aclblasStrsm");
        if (rs != ACLBLAS_STATUS_SUCCESS) {
            return rs;
        }
        const size_t k0s = static_cast<size_t>(k0);
        const size_t ldas = static_cast<size_t>(lda);
        const size_t ldbs = static_cast<size_t>(ldb);
        if (forward) {
            const int r1 = k0 + kb;
            const int rest = k - r1;
            if (rest <= 0) {
                continue;
            }
            const size_t r1s = static_cast<size_t>(r1);
            if (left) {
                // B[r1:, :] -= op(A)[r1:, k0:r1] Y
                rs = tr ? aclblasSgemm(handle, ACLBLAS_OP_T, ACLBLAS_OP_N, rest, n, kb, &minus1, A + k0s + r1s * ldas,
                                       lda, B + k0s, ldb, &one, B + r1s, ldb)
                        : aclblasSgemm(handle, ACLBLAS_OP_N, ACLBLAS_OP_N, rest, n, kb, &minus1, A + r1s + k0s * ldas,
                                       lda, B + k0s, ldb, &one, B + r1s, ldb);
            } else {
                // B[:, r1:] -= Y op(A)[k0:r1, r1:]
                rs = tr ? aclblasSgemm(handle, ACLBLAS_OP_N, ACLBLAS_OP_T, m, rest, kb, &minus1, B + k0s * ldbs, ldb,
                                       A + r1s + k0s * ldas, lda, &one, B + r1s * ldbs, ldb)
                        : aclblasSgemm(handle, ACLBLAS_OP_N, ACLBLAS_OP_N, m, rest, kb, &minus1, B + k0s * ldbs, ldb,
                                       A + k0s + r1s * ldas, lda, &one, B + r1s * ldbs, ldb);
            }
        } else {
            if (k0 == 0) {
                continue;
            }
            if (left) {
                // B[0:k0, :] -= op(A)[0:k0, k0:k0+kb] Y
                rs = tr ? aclblasSgemm(handle, ACLBLAS_OP_T, ACLBLAS_OP_N, k0, n, kb, &minus1, A + k0s, lda, B + k0s,
                                       ldb, &one, B, ldb)
                        : aclblasSgemm(handle, ACLBLAS_OP_N, ACLBLAS_OP_N, k0, n, kb, &minus1, A + k0s * ldas, lda,
                                       B + k0s, ldb, &one, B, ldb);
            } else {
                // B[:, 0:k0] -= Y op(A)[k0:k0+kb, 0:k0]
                rs = tr ? aclblasSgemm(handle, ACLBLAS_OP_N, ACLBLAS_OP_T, m, k0, kb, &minus1, B + k0s * ldbs, ldb,
                                       A + k0s * ldas, lda, &one, B, ldb)
                        : aclblasSgemm(handle, ACLBLAS_OP_N, ACLBLAS_OP_N, m, k0, kb, &minus1, B + k0s * ldbs, ldb,
                                       A + k0s, lda, &one, B, ldb);
            }
        }
        if (rs != ACLBLAS_STATUS_SUCCESS) {
            return rs;
        }
    }
    if (a != 1.0f && scaleAfter) {
        LaunchScale(st, B, m, n, ldb, a, false);
        return LaunchStatus("aclblasStrsm");
    }
    return ACLBLAS_STATUS_SUCCESS;
}
