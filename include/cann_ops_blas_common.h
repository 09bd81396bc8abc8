/*
 * Status codes and enums of the aclblas C API.
 *
 * Names and numeric values are part of the ABI and are identical to the
 * reference's include/cann_ops_blas_common.h:18-132 so that binaries built
 * against either header interoperate.
 */
#ifndef ACLBLAS_AMD_CANN_OPS_BLAS_COMMON_H
#define ACLBLAS_AMD_CANN_OPS_BLAS_COMMON_H

#ifdef __cplusplus
extern "C" {
#endif

/* reference include/cann_ops_blas_common.h:18-31 */
typedef enum aclblasStatus {
    ACLBLAS_STATUS_SUCCESS = 0,
    ACLBLAS_STATUS_NOT_INITIALIZED = 1,
    ACLBLAS_STATUS_ALLOC_FAILED = 2,
    ACLBLAS_STATUS_INVALID_VALUE = 3,
    ACLBLAS_STATUS_MAPPING_ERROR = 4,
    ACLBLAS_STATUS_EXECUTION_FAILED = 5,
    ACLBLAS_STATUS_INTERNAL_ERROR = 6,
    ACLBLAS_STATUS_NOT_SUPPORTED = 7,
    ACLBLAS_STATUS_ARCH_MISMATCH = 8,
    ACLBLAS_STATUS_HANDLE_IS_NULLPTR = 9,
    ACLBLAS_STATUS_INVALID_ENUM = 10,
    ACLBLAS_STATUS_UNKNOWN = 11
} aclblasStatus_t;

/* reference include/cann_ops_blas_common.h:35-38 */
typedef enum aclblasFillMode {
    ACLBLAS_UPPER = 121,
    ACLBLAS_LOWER = 122
} aclblasFillMode_t;

/* reference include/cann_ops_blas_common.h:41-44 */
typedef enum aclblasDiagType {
    ACLBLAS_NON_UNIT = 131,
    ACLBLAS_UNIT = 132
} aclblasDiagType_t;

/* reference include/cann_ops_blas_common.h:47-50 */
typedef enum aclblasSideMode {
    ACLBLAS_SIDE_LEFT = 141,
    ACLBLAS_SIDE_RIGHT = 142
} aclblasSideMode_t;

/* reference include/cann_ops_blas_common.h:52-65 */
#ifndef ACLBLAS_OPERATION_DECLARED
#define ACLBLAS_OPERATION_DECLARED
typedef enum aclblasOperation {
    ACLBLAS_OP_N = 111,
    ACLBLAS_OP_T = 112,
    ACLBLAS_OP_C = 113
} aclblasOperation_t;
#endif

typedef struct aclblasComplex {
    float real;
    float imag;
} aclblasComplex;

typedef struct aclblasDoubleComplex {
    double real;
    double imag;
} aclblasDoubleComplex;

/* reference include/cann_ops_blas_common.h:81-93 */
typedef enum aclblasComputeType {
    ACLBLAS_COMPUTE_16F = 0,
    ACLBLAS_COMPUTE_16F_PEDANTIC = 1,
    ACLBLAS_COMPUTE_32F = 2,
    ACLBLAS_COMPUTE_32F_PEDANTIC = 3,
    ACLBLAS_COMPUTE_32F_FAST_16F = 4,
    ACLBLAS_COMPUTE_32F_FAST_16BF = 5,
    ACLBLAS_COMPUTE_32F_FAST_TF32 = 6,
    ACLBLAS_COMPUTE_64F = 7,
    ACLBLAS_COMPUTE_64F_PEDANTIC = 8,
    ACLBLAS_COMPUTE_32I = 9,
    ACLBLAS_COMPUTE_32I_PEDANTIC = 10
} aclblasComputeType_t;

/* reference include/cann_ops_blas_common.h:97-107 */
typedef enum aclblasGemmAlgo {
    ACLBLAS_GEMM_DEFAULT = 0,
    ACLBLAS_GEMM_ALGO0 = 1,
    ACLBLAS_GEMM_ALGO1 = 2,
    ACLBLAS_GEMM_ALGO2 = 3,
    ACLBLAS_GEMM_ALGO3 = 4,
    ACLBLAS_GEMM_ALGO4 = 5,
    ACLBLAS_GEMM_ALGO5 = 6,
    ACLBLAS_GEMM_ALGO6 = 7,
    ACLBLAS_GEMM_ALGO7 = 8
} aclblasGemmAlgo_t;

/* LAPACK-style argument reporting through an `int* info` argument: 0 on
 * success, -i when the i-th argument (handle and info not counted) is illegal.
 * reference include/cann_ops_blas_common.h:122-132 */
typedef enum aclblasLapackInfo {
    ACLBLAS_LAPACK_INFO_OK = 0,
    ACLBLAS_LAPACK_INFO_ARG_1 = -1,
    ACLBLAS_LAPACK_INFO_ARG_2 = -2,
    ACLBLAS_LAPACK_INFO_ARG_3 = -3,
    ACLBLAS_LAPACK_INFO_ARG_4 = -4,
    ACLBLAS_LAPACK_INFO_ARG_5 = -5,
    ACLBLAS_LAPACK_INFO_ARG_6 = -6,
    ACLBLAS_LAPACK_INFO_ARG_7 = -7,
    ACLBLAS_LAPACK_INFO_ARG_8 = -8
} aclblasLapackInfo_t;

#ifdef __cplusplus
}
#endif

#endif /* ACLBLAS_AMD_CANN_OPS_BLAS_COMMON_H */
