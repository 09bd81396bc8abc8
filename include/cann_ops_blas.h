/*
 * aclblas C API — Level-3 GEMM and batched-LU hot path, MI355X (gfx950) build.
 *
 * Every entry point keeps the reference's name, argument order, argument
 * meaning, status codes and host-pointer alpha/beta semantics, so a program
 * linked against the reference's libops_blas.so links against this library
 * unchanged.  Each declaration cites the reference prototype it replaces.
 *
 * Conventions shared by every call:
 *  - Matrices are column-major.  Device pointers for matrices, device arrays of
 *    device pointers for the batched LAPACK calls, host pointers for alpha/beta
 *    (read once, on the host, during the call).
 *  - Work is enqueued asynchronously on the handle's stream (aclblasSetStream);
 *    results are visible after the caller synchronises that stream.  No call on
 *    the hot path allocates device memory or synchronises, so sequences of them
 *    can be captured into a HIP graph.
 *  - A handle is not thread-safe: it owns one workspace.
 */
#ifndef ACLBLAS_AMD_CANN_OPS_BLAS_H
#define ACLBLAS_AMD_CANN_OPS_BLAS_H

#include <stddef.h>
#include <stdint.h>
#include <acl/acl.h>
#include "cann_ops_blas_common.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct _aclblas_handle* aclblasHandle_t;
typedef void (*aclblasLogCallback)(const char*);

/* ------------------------------------------------------------------------ */
/* Handle, stream, workspace, version, logging                               */
/* ------------------------------------------------------------------------ */

/* reference include/cann_ops_blas.h:41.  *handle must be NULL on entry
 * (INVALID_VALUE otherwise).  Pre-allocates a 32 MiB library workspace on the
 * current HIP device; ALLOC_FAILED if that allocation fails.  On a host with no
 * visible GPU the handle is created without a workspace so that argument
 * checking can be exercised; any call that would launch work then returns
 * EXECUTION_FAILED. */
aclblasStatus_t aclblasCreate(aclblasHandle_t* handle);

/* reference include/cann_ops_blas.h:53.  Synchronises the handle's stream,
 * frees the library workspace (never a user workspace), frees the handle. */
aclblasStatus_t aclblasDestroy(aclblasHandle_t handle);

/* reference include/cann_ops_blas.h:66.  Synchronises the old stream, binds
 * the new one (NULL = default stream) and reverts to the library workspace. */
aclblasStatus_t aclblasSetStream(aclblasHandle_t handle, aclrtStream stream);

/* reference include/cann_ops_blas.h:76 */
aclblasStatus_t aclblasGetStream(aclblasHandle_t handle, aclrtStream* stream);

/* reference include/cann_ops_blas.h:93.  Borrows caller device memory; grow
 * only (a smaller user workspace than the current one is ignored). */
aclblasStatus_t aclblasSetWorkspace(aclblasHandle_t handle, void* workspace, size_t workspaceSize);

/* reference include/cann_ops_blas.h:105.  MAJOR*10000 + MINOR*100 + PATCH. */
aclblasStatus_t aclblasGetVersion(aclblasHandle_t handle, int* version);

/* reference include/cann_ops_blas.h:116-133 (declared there, defined here with
 * C linkage).  Messages are emitted only on the host, never on a launch path. */
aclblasStatus_t aclblasLoggerConfigure(int logIsOn, int logToStdOut, int logToStdErr, const char* logFile);
aclblasStatus_t aclblasSetLoggerCallback(aclblasLogCallback userCallback);
aclblasStatus_t aclblasGetLoggerCallback(aclblasLogCallback* userCallback);

/* ------------------------------------------------------------------------ */
/* Level-3 GEMM                                                              */
/* ------------------------------------------------------------------------ */

/* C = alpha*op(A)*op(B) + beta*C, FP32.  Same semantics as aclblasSgemmEx with
 * algo = ACLBLAS_GEMM_DEFAULT (reference blas/gemm_ex/arch35/sgemm_ex_host.cpp:404);
 * the reference's header has no plain Sgemm, the signature is cuBLAS's. */
aclblasStatus_t aclblasSgemm(
    aclblasHandle_t handle, aclblasOperation_t transA, aclblasOperation_t transB, int m, int n, int k,
    const float* alpha, const float* A, int lda, const float* B, int ldb,
    const float* beta, float* C, int ldc);

/* reference include/cann_ops_blas.h:443-446 */
aclblasStatus_t aclblasSgemmEx(
    aclblasHandle_t handle, aclblasOperation_t transA, aclblasOperation_t transB, int m, int n, int k,
    const float* alpha, const float* A, int lda, const float* B, int ldb,
    const float* beta, float* C, int ldc, aclblasGemmAlgo_t algo);

/* reference include/cann_ops_blas.h:438-441.  Accepted (Atype, Btype, Ctype,
 * computeType): the reference's table (blas/gemm/arch35/gemm_ex_host.cpp:45-69)
 * plus BF16xBF16->FP32 and FP16xFP16->FP32 under COMPUTE_32F.  alpha and beta
 * are host floats for every combination. */
aclblasStatus_t aclblasGemmEx(
    aclblasHandle_t handle, aclblasOperation_t transA, aclblasOperation_t transB, int m, int n, int k,
    const void* alpha, const void* A, aclDataType Atype, int lda, const void* B, aclDataType Btype, int ldb,
    const void* beta, void* C, aclDataType Ctype, int ldc, aclblasComputeType_t computeType, aclblasGemmAlgo_t algo);

/* FP8 GEMM with per-tensor dequantisation scales (no counterpart in the
 * reference; the scale semantics are those of its Lt attributes
 * ACLBLASLT_MATMUL_DESC_A_SCALE_POINTER / B_SCALE_POINTER, reference
 * include/cann_ops_blasLt.h:160-200):
 *   C = alpha * scaleA[0] * scaleB[0] * op(A) op(B) + beta * C.
 * Atype/Btype: ACL_FLOAT8_E4M3FN or ACL_FLOAT8_E5M2 (any mix); Ctype: ACL_FLOAT16,
 * ACL_BF16 or ACL_FLOAT (else NOT_SUPPORTED).  alpha/beta: host floats;
 * scaleA/scaleB: device floats, NULL meaning 1.  Runs on the native FP8 matrix
 * instruction: each block of 128 products is summed with a relative error of
 * about 2^-14 of its largest product (FP32 accumulation across blocks).  Bad
 * enums or dimensions: INVALID_VALUE. */
aclblasStatus_t aclblasGemmExScaled(
    aclblasHandle_t handle, aclblasOperation_t transA, aclblasOperation_t transB, int m, int n, int k,
    const float* alpha, const void* A, aclDataType Atype, int lda, const float* scaleA, const void* B,
    aclDataType Btype, int ldb, const float* scaleB, const float* beta, void* C, aclDataType Ctype, int ldc);

/* Per-tensor FP8 quantisation (no counterpart in the reference; formats and
 * rounding of reference test/frame/dtype_utils.h:39-160): *scale = amax(|x|) /
 * fp8_max (1 when x is all zero; fp8_max = 448 for E4M3FN, 57344 for E5M2) and
 * q[i] = fp8(x[i] * (fp8_max / amax)), round to nearest even, saturating.  x:
 * count device floats; q: count device bytes; scale: one device float, written
 * on the handle's stream (so it can feed aclblasGemmExScaled directly). */
aclblasStatus_t aclblasQuantizeFp8(aclblasHandle_t handle, long long count, const float* x, aclDataType qtype,
                                   void* q, float* scale);

/* No counterpart in the reference.  aclblasQuantizeFp8 for callers that cannot
 * allocate (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): the per-block maxima live in a 4 KiB device array of the library
 * (one per GPU, held until the process ends) instead of the handle's workspace,
 * so the call allocates nothing and never grows the workspace.  Same arguments,
 * checks, results and stream semantics as aclblasQuantizeFp8; any count.  All
 * calls on one GPU share that array: calls that overlap on one GPU (several
 * streams or host threads) race. */
aclblasStatus_t aclblasQuantizeFp8NoAlloc(aclblasHandle_t handle, long long count, const float* x,
                                          aclDataType qtype, void* q, float* scale);

/* reference include/cann_ops_blas.h:409-414.  C[i] = alpha*op(A[i])*op(B[i]) +
 * beta*C[i] over device arrays of device pointers, one shape for the batch.
 * Accepted types: FP16 (COMPUTE_16F or 32F), BF16 and FP8 E4M3/E5M2 (any mix,
 * FP16 C) under COMPUTE_32F; FP32 returns NOT_SUPPORTED as in the reference
 * (use aclblasSgemmBatched). */
aclblasStatus_t aclblasGemmBatchedEx(
    aclblasHandle_t handle, aclblasOperation_t transa, aclblasOperation_t transb, int m, int n, int k,
    const void* alpha, const void* const Aarray[], aclDataType Atype, int lda,
    const void* const Barray[], aclDataType Btype, int ldb,
    const void* beta, void* const Carray[], aclDataType Ctype, int ldc,
    int batchCount, aclblasComputeType_t computeType, aclblasGemmAlgo_t algo);

/* reference include/cann_ops_blas.h:416-420.  FP32 batched GEMM over device
 * pointer arrays (the MFMA FP32 path). */
aclblasStatus_t aclblasSgemmBatched(
    aclblasHandle_t handle, aclblasOperation_t transa, aclblasOperation_t transb, int m, int n, int k,
    const float* alpha, const float* const Aarray[], int lda,
    const float* const Barray[], int ldb,
    const float* beta, float* const Carray[], int ldc,
    int batchCount);

/* reference include/cann_ops_blas.h:448-455.  FP32 GEMMs grouped by shape:
 * group g holds groupSizeArray[g] problems of shape (mArray[g], nArray[g],
 * kArray[g]) with transaArray[g]/transbArray[g] (N or T; C -> INVALID_ENUM),
 * alphaArray[g]/betaArray[g] (host floats) and ldaArray[g].. (host arrays).
 * Aarray/Barray/Carray are HOST arrays of device pointers, one per problem,
 * in group order.  Groups with groupSize, m or n equal to 0 are skipped;
 * alpha == 0 scales C by beta.  One persistent launch; the problem table is
 * copied to the handle's workspace asynchronously. */
aclblasStatus_t aclblasSgemmGroupedBatched(
    aclblasHandle_t handle, int groupCount, const aclblasOperation_t* transaArray,
    const aclblasOperation_t* transbArray, const int* mArray, const int* nArray, const int* kArray,
    const float* alphaArray, const float* const* Aarray, const int* ldaArray, const float* const* Barray,
    const int* ldbArray, const float* betaArray, float* const* Carray, const int* ldcArray,
    const int* groupSizeArray);

/* reference include/cann_ops_blas.h:461-467.  Grouped GEMM over FP16 x FP16 ->
 * FP16 (COMPUTE_16F or _32F), BF16 x BF16 -> BF16 (_32F) and FP8 E4M3/E5M2 in
 * any mix -> FP16 (_32F); other combinations NOT_SUPPORTED before any other
 * check.  trans/m/n/k/lda/ldb/ldc/groupSize and alphaArray/betaArray are HOST
 * arrays of groupCount entries (alpha/beta FP16 under COMPUTE_16F, float
 * otherwise); Aarray/Barray/Carray are DEVICE arrays of device pointers, the
 * problems of group 0 first.  Products accumulate in FP32; FP16 results round
 * to nearest with ties away from zero, BF16 results to nearest even. */
aclblasStatus_t aclblasGemmGroupedBatchedEx(
    aclblasHandle_t handle, const aclblasOperation_t transaArray[], const aclblasOperation_t transbArray[],
    const int mArray[], const int nArray[], const int kArray[], const void* alphaArray,
    const void* const Aarray[], aclDataType Atype, const int ldaArray[],
    const void* const Barray[], aclDataType Btype, const int ldbArray[],
    const void* betaArray, void* const Carray[], aclDataType Ctype, const int ldcArray[],
    int groupCount, const int groupSize[], aclblasComputeType_t computeType);

/* ------------------------------------------------------------------------ */
/* Batched LAPACK (FP32)                                                     */
/* ------------------------------------------------------------------------ */

/* reference include/cann_ops_blas.h:403-404.  In-place LU with partial
 * pivoting of batchSize n x n matrices.  Aarray: device array of device
 * pointers.  PivotArray: device int[n*batchSize] (1-based) or NULL for LU
 * without pivoting.  infoArray: device int[batchSize]; may be NULL only when
 * PivotArray is NULL.  info[b] = k+1 for the first zero pivot found at column
 * k, where factorisation of that matrix stops (pivots after k are not written). */
aclblasStatus_t aclblasSgetrfBatched(
    aclblasHandle_t handle, int n, float* const Aarray[], int lda, int* PivotArray, int* infoArray, int batchSize);

/* reference include/cann_ops_blas.h:457-459.  Solves op(A) X = B with the
 * factors of aclblasSgetrfBatched, overwriting B.  info is a HOST int: 0, or
 * -i for an illegal i-th argument.  nrhs is limited to [0, 256]. */
aclblasStatus_t aclblasSgetrsBatched(
    aclblasHandle_t handle, aclblasOperation_t trans, int n, int nrhs, const float* const Aarray[], int lda,
    const int* devIpiv, float* const Barray[], int ldb, int* info, int batchCount);

/* reference include/cann_ops_blas.h:430-432.  Out-of-place inverse C = A^-1
 * from LU factors.  infoArray (device, required): j+1 for the zero diagonal
 * element u_jj with the largest j, 0 otherwise. */
aclblasStatus_t aclblasSgetriBatched(
    aclblasHandle_t handle, int n, const float* const Aarray[], int lda, const int* PivotArray,
    float* const Carray[], int ldc, int* infoArray, int batchSize);

/* reference include/cann_ops_blas.h:339-341.  B := alpha op(A)^-1 B (side
 * LEFT, A m x m) or alpha B op(A)^-1 (side RIGHT, A n x n), A triangular
 * (uplo; diag UNIT: the diagonal is taken as 1 and not read), op = N, T or C
 * (= T), column-major FP32, B m x n (ldb).  Argument checks and their order as
 * the reference's strsm_host.cpp:406-452 (every bad value, including an
 * invalid side, uplo, trans or diag, gives INVALID_VALUE); m == 0 or n == 0
 * returns before B and A are checked; alpha == 0 writes zeros to B without
 * reading A.  No workspace. */
aclblasStatus_t aclblasStrsm(
    aclblasHandle_t handle, aclblasSideMode_t side, aclblasFillMode_t uplo, aclblasOperation_t trans,
    aclblasDiagType_t diag, int m, int n, const float* alpha, const float* A, int lda, float* B, int ldb);

/* reference include/cann_ops_blas.h:434-436.  Ainv[i] = A[i]^-1 for
 * 0 <= n <= 32 (pivoted LU, then the out-of-place inverse; A and Ainv are
 * device arrays of device pointers, column-major, Ainv[i] must not overlap
 * A[i]).  info (device, batchSize entries): the inverse's rule, j+1 for the
 * zero u_jj with the largest j, 0 otherwise.  Uses batchSize * 8 +
 * n^2 * batchSize * 4 + n * batchSize * 4 bytes of the handle's workspace and
 * returns EXECUTION_FAILED when the workspace is smaller (it is not grown). */
aclblasStatus_t aclblasSmatinvBatched(
    aclblasHandle_t handle, int n, const float* const A[], int lda, float* const Ainv[], int lda_inv, int* info,
    int batchSize);

/* ------------------------------------------------------------------------ */
/* Differentiation                                                           */
/* ------------------------------------------------------------------------ */

/* Reverse-mode derivative of X = A^-1 B (no counterpart in the reference; it
 * is built from the reference's transposed getrs,
 * blas/getrs_batched/arch35/sgetrs_batched_kernel.cpp:263-362, and an FP32
 * batched GEMM, blas/gemm_batched/arch35/gemm_batched_host.cpp:278-360).
 * LUarray/devIpiv: the factors and pivots from aclblasSgetrfBatched (devIpiv
 * NULL for the unpivoted factorisation).  Xarray: the forward solution (n x
 * nrhs).  dXarray: on entry the upstream gradient dL/dX, on exit dL/dB =
 * A^-T dX.  dAarray: on exit dL/dA = -dL/dB X^T (n x n, overwritten); NULL
 * skips it.  All arrays are device arrays of device pointers.  Status rules
 * as for aclblasSgetrsBatched, with every bad argument INVALID_VALUE. */
aclblasStatus_t aclblasSgetrsBatchedAdjoint(
    aclblasHandle_t handle, int n, int nrhs, const float* const LUarray[], int lda, const int* devIpiv,
    const float* const Xarray[], int ldx, float* const dXarray[], int lddx, float* const dAarray[], int ldda,
    int batchCount);

/* ------------------------------------------------------------------------ */
/* Mixed-precision batched solve                                             */
/* ------------------------------------------------------------------------ */

/* Precision of the LU's trailing updates A22 += (-L21) U12 (no reference
 * counterpart).  Panels, pivots, the diagonal blocks and every accumulation are
 * FP32 in all modes. */
typedef enum {
    ACLBLAS_MIXED_UPDATE_FP32 = 0,   /* FP32 matrix instruction */
    ACLBLAS_MIXED_UPDATE_BF16 = 1,   /* operands rounded to BF16: one product */
    ACLBLAS_MIXED_UPDATE_BF16X2 = 2, /* operands split into two BF16 planes: three products */
    ACLBLAS_MIXED_UPDATE_BF16X3 = 3  /* three BF16 planes (the FP32 value): six products */
} aclblasMixedUpdate_t;

/* Solves A_i X_i = B_i for a batch of systems of different orders (no
 * counterpart in the reference; it follows the reference's LU semantics of
 * blas/getrf_batched/arch35/sgetrf_batched_kernel.cpp:52-150 - partial
 * pivoting, stop at an exact zero pivot - with a blocked algorithm).  Per
 * system: n = nArray[i] <= 256, nrhs = nrhsArray[i] <= 16 (larger:
 * NOT_SUPPORTED), column-major FP32 A (lda), B (ldb), X (ldx); A and B are
 * not modified.  The LU runs in FP32 with its trailing updates on matrix cores
 * in the precision `update`; the solution is refined with FP64 residuals until
 * every right-hand side has normwise backward error
 * |b - A x|_inf / (|A|_inf |x|_inf + |b|_inf) <= tol, the backward error fails
 * to halve in a correction, or maxIter corrections have been applied.  With a
 * BF16 mode, a system that does not reach tol (or meets a zero pivot) is
 * factored again with FP32 updates and solved again (fallbackArray[i] = 1).
 * nArray .. ldxArray and the pointer arrays are HOST arrays (device matrices);
 * iterArray (corrections applied by the final factorisation), berrArray (final
 * worst backward error), infoArray (0, or k+1 for an exact zero pivot at
 * column k of the final factorisation, X = NaN) and fallbackArray are device
 * arrays of batchCount entries, each may be NULL.  INVALID_ENUM for an unknown
 * `update`.  Works in the handle's workspace (about 4.1 n^2 + 128 n bytes per
 * system plus an 80-byte record). */
aclblasStatus_t aclblasSgesvBatchedMixedEx(
    aclblasHandle_t handle, int batchCount, const int* nArray, const int* nrhsArray, const float* const Aarray[],
    const int* ldaArray, const float* const Barray[], const int* ldbArray, float* const Xarray[],
    const int* ldxArray, double tol, int maxIter, aclblasMixedUpdate_t update, int* iterArray, float* berrArray,
    int* infoArray, int* fallbackArray);

/* aclblasSgesvBatchedMixedEx with update = ACLBLAS_MIXED_UPDATE_BF16X2 and no
 * fallback array. */
aclblasStatus_t aclblasSgesvBatchedMixed(
    aclblasHandle_t handle, int batchCount, const int* nArray, const int* nrhsArray, const float* const Aarray[],
    const int* ldaArray, const float* const Barray[], const int* ldbArray, float* const Xarray[],
    const int* ldxArray, double tol, int maxIter, int* iterArray, float* berrArray, int* infoArray);

/* Registers caller-created HIP events (hipEvent_t, passed as void*) that the
 * calls made of several kernel launches record on the handle's stream: event
 * 2i just before and 2i+1 just after their i-th launch (no counterpart in the
 * reference: a profiling aid, since stream events placed around a call cannot
 * separate its kernels).  Honoured by aclblasSgesvBatchedMixed(Ex): launch 0
 * factors every system, 1 solves and refines, 2 and 3 re-factor and re-solve
 * the systems that fell back to FP32 updates (BF16 modes only).  Slots past
 * `count` or NULL entries are not recorded; count = 0 clears the list.  The
 * events must outlive their use. */
aclblasStatus_t aclblasSetLaunchEvents(aclblasHandle_t handle, void* const events[], int count);

/* ---- launch tuning (no reference counterpart: the reference picks its tiles
 * with fixed host rules, e.g. blas/gemm_grouped_batched/arch35/
 * gemm_grouped_batched_host.cpp:98-124).  Calls with several launch shapes map
 * their problem to a size bucket and take the shape from a table measured by
 * ops-blas_amd/autotune.py.  Operations and their two parameters:
 *   0 GEMM (GemmEx, Sgemm, SgemmEx, GemmBatchedEx, GemmExScaled):
 *     a = output tile edge 64 / 128 / 256 (256: 1- and 2-byte operands only),
 *     or for FP32 32 (32 x 32 tiles, eight k-blocks of loads in flight) and
 *     65 (64 x 64 tiles, four k-blocks in flight),
 *     b = tile rows walked together per L2 group, 1..64;
 *   1 grouped GEMM: a = tile edge 64 / 128, b = persistent workgroups per CU 1..4
 *     (separate buckets for aclblasSgemmGroupedBatched and
 *     aclblasGemmGroupedBatchedEx);
 *   2 mixed solve: a = factor workgroups per CU 1..4, b = solve workgroups per CU 1..4;
 *   3 batched LU (aclblasSgetrfBatched, n <= 64): a = matrix rows per lane 1 or 2
 *     (2 needs n > 8), b ignored.
 * Results do not depend on the choice (every output element sums its products
 * in the same order under every tile shape and grid). */

/* Forces (a, b) for every bucket of `op` on this handle; a < 0 removes the
 * override.  INVALID_VALUE when (a, b) is legal for no bucket of `op`; a forced
 * choice that a bucket cannot take (a 256 tile for FP32) leaves that bucket on
 * the table. */
aclblasStatus_t aclblasSetTuningOverride(aclblasHandle_t handle, int op, int a, int b);

/* The choice the handle uses for (op, bucket): override, table or built-in. */
aclblasStatus_t aclblasGetTuningChoice(aclblasHandle_t handle, int op, int bucket, int* a, int* b);

/* Bucket of a problem: op 0 takes cls (0 FP32, 1 BF16/FP16, 2 FP8 under
 * GemmEx, 3 FP8 under GemmExScaled), units = 128 x 128 output tiles over the
 * call and k; op 1 takes cls (0 FP32 under SgemmGroupedBatched, 1 16-bit or
 * FP8 under GemmGroupedBatchedEx) and units = 128 x 128 tiles over all
 * problems; op 2 takes
 * cls = the trailing-update class (0: BF16 and BF16X2, 1: FP32 and BF16X3) and
 * units = number of systems; op 3 takes cls = the order n (1..64) and units =
 * the batch count.
 * k is ignored for ops 1, 2 and 3. */
aclblasStatus_t aclblasTuningBucket(aclblasHandle_t handle, int op, int cls, long long units, int k, int* bucket);

#ifdef __cplusplus
}
#endif

#endif /* ACLBLAS_AMD_CANN_OPS_BLAS_H */
