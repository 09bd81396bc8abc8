// Internal launchers shared between the C entry points (no argument checks:
// callers validate first).  All enqueue on `st` and return the launch status.
#pragma once

#include <hip/hip_runtime.h>

#include "cann_ops_blas.h"

struct _aclblas_handle;

namespace aclblas {

// LU with partial pivoting (piv == nullptr: no pivoting) on the handle's
// stream with its tuned launch shape, getrf.hip.
aclblasStatus_t GetrfBatchedLaunch(const _aclblas_handle* h, int n, float* const* A, int lda, int* piv, int* info,
                                   int batch);

// Pivoted LU of n <= 32 matrices, out of place: A read, the factors written
// to out + b n^2 (leading dimension n), &out[b n^2] to table[b], all n pivots
// of each matrix written (0 after a zero pivot), getrf.hip.
aclblasStatus_t GetrfOutOfPlaceLaunch(const _aclblas_handle* h, int n, const float* const* A, int lda, float* out,
                                      float** table, int* piv, int* info, int batch);

// op(A) X = B with LU factors, B overwritten (trans: A^T), getrs.hip.
aclblasStatus_t GetrsBatchedLaunch(hipStream_t st, bool trans, int n, int nrhs, const float* const* A, int lda,
                                   const int* piv, float* const* B, int ldb, int batch);

// C = A^-1 from LU factors (out of place; C must not overlap A), info[b] = j+1
// for the zero u_jj with the largest j, getrs.hip.  NOT_SUPPORTED when n is too
// large for one LDS chunk.
aclblasStatus_t GetriBatchedLaunch(hipStream_t st, int n, const float* const* A, int lda, const int* piv,
                                   float* const* C, int ldc, int* info, int batch);

// C[i] = alpha op(A[i]) op(B[i]) + beta C[i], FP32, device pointer arrays, on the
// handle's stream with its tuned tile choice, gemm.hip.
aclblasStatus_t SgemmBatchedLaunch(const _aclblas_handle* h, bool transA, bool transB, int m, int n, int k,
                                   float alpha, const float* const* A, int lda, const float* const* B, int ldb,
                                   float beta, float* const* C, int ldc, int batch);

// C = alpha op(A) op(B) + beta C for a validated problem (any m, n, k; k == 0
// or alpha == 0 scales C), element types as aclblasGemmEx, on the handle's
// stream with its tuned tile choice, gemm.hip.
aclblasStatus_t GemmLaunch(const _aclblas_handle* h, aclDataType at, aclDataType bt, aclDataType ct, bool transA,
                           bool transB, int m, int n, int k, float alpha, const void* A, int lda, const void* B,
                           int ldb, float beta, void* C, int ldc, const float* scaleA, const float* scaleB);

// Row-major MX GEMM D = alpha op(A) op(B) + beta C (FP32 C, FP32 or BF16 D):
// MXFP8 E4M3FN (fp4 false) or MXFP4 E2M1 operands with E8M0 block scales in
// the layouts aclblasLtMatmul states; k a multiple of 32; lda / ldb are the
// bytes per stored row, lt_mx.hip.
aclblasStatus_t LtMxLaunch(hipStream_t st, bool fp4, bool transA, bool transB, int m, int n, int k, float alpha,
                           const void* A, long long lda, const void* B, long long ldb, const void* scaleA,
                           const void* scaleB, float beta, const float* C, int ldc, void* D, int ldd, bool bf16Out);

}  // namespace aclblas
