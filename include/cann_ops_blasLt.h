/*
 * aclblasLt: the lightweight matmul API of the aclblas MI355X build.
 *
 * Drop-in for the reference's BLASLt matmul entry points (reference
 * include/cann_ops_blasLt.h: types :17-99, enums :100-257, functions
 * :301-924).  Type names, enum values, structure sizes and argument lists are
 * the reference's, so a caller compiled against it links against
 * libops_blas.so unchanged.  Not built: the matrix-transform, logger and
 * algo-config-attribute entries (reference :689-1085 beyond what is declared
 * below).
 *
 * aclblasLtMatmul computes D = alpha op(A) op(B) + beta C for
 *   FP32 x FP32 -> FP32 (C FP32)
 *   MXFP8 E4M3FN x MXFP8 E4M3FN -> FP32 or BF16 (C FP32)
 *   MXFP4 E2M1 x MXFP4 E2M1 -> FP32 or BF16 (C FP32)
 * with every matrix row-major (the reference's only supported order; the
 * layout ORDER attribute is stored and returned, and as in the reference it
 * does not change the computation).  MX operands carry one E8M0 scale byte
 * (2^(s - 127)) per 32 elements along K through A/B_SCALE_POINTER, in the
 * reference's two layouts (see aclblasLtMatmul below).  K must be a multiple
 * of 32 for MX types.
 */
#ifndef ACLBLAS_AMD_CANN_OPS_BLASLT_H
#define ACLBLAS_AMD_CANN_OPS_BLASLT_H

#include "cann_ops_blas_common.h"
#include "acl/acl.h"

#ifdef __cplusplus
#include <cstddef>
#include <cstdint>
extern "C" {
#else
#include <stddef.h>
#include <stdint.h>
#endif

typedef struct _aclblaslt_handle* aclblasLtHandle_t;

/* Opaque descriptor storage (the reference's sizes, :24-47). */
typedef struct {
    uint64_t data[8];
} aclblasLtMatrixLayoutOpaque_t;
typedef aclblasLtMatrixLayoutOpaque_t* aclblasLtMatrixLayout_t;

typedef struct {
    uint64_t data[24];
} aclblasLtMatmulDescOpaque_t;
typedef aclblasLtMatmulDescOpaque_t* aclblasLtMatmulDesc_t;

typedef struct {
    uint64_t data[8];
} aclblasLtMatmulPreferenceOpaque_t;
typedef aclblasLtMatmulPreferenceOpaque_t* aclblasLtMatmulPreference_t;

/* Algorithm blob (reference :59-67). */
typedef struct _aclblasLtMatmulAlgo_t {
    uint8_t data[16];
    size_t max_workspace_bytes;
} aclblasLtMatmulAlgo_t;

/* One heuristic result (reference :70-77). */
typedef struct _aclblasLtMatmulHeuristicResult_t {
    aclblasLtMatmulAlgo_t algo;
    size_t workspaceSize;
    aclblasStatus_t state;
    float wavesCount;
    int reserved[4];
} aclblasLtMatmulHeuristicResult_t;

typedef enum aclblasLtOrder {
    ACLBLASLT_ORDER_COL = 0,
    ACLBLASLT_ORDER_ROW = 1,
    ACLBLASLT_ORDER_COL32 = 2,
    ACLBLASLT_ORDER_COL4_4R2_8C = 3,
    ACLBLASLT_ORDER_COL32_2R_4R4 = 4
} aclblasLtOrder_t;

typedef enum aclblasLtPropertyType {
    ACLBLASLT_PROPERTY_MAJOR_VERSION = 0,
    ACLBLASLT_PROPERTY_MINOR_VERSION = 1,
    ACLBLASLT_PROPERTY_PATCH_LEVEL = 2
} aclblasLtPropertyType_t;

/* Epilogue values (reference :117-145); aclblasLtMatmul runs DEFAULT only. */
typedef enum aclblasLtEpilogue {
    ACLBLASLT_EPILOGUE_DEFAULT = 1,
    ACLBLASLT_EPILOGUE_RELU = 2,
    ACLBLASLT_EPILOGUE_BIAS = 4,
    ACLBLASLT_EPILOGUE_RELU_BIAS = 6,
    ACLBLASLT_EPILOGUE_GELU = 32,
    ACLBLASLT_EPILOGUE_GELU_BIAS = 36
} aclblasLtEpilogue_t;

typedef enum aclblasLtMatrixLayoutAttribute {
    ACLBLASLT_MATRIX_LAYOUT_BATCH_COUNT = 0,          /* int32_t, default 1 */
    ACLBLASLT_MATRIX_LAYOUT_STRIDED_BATCH_OFFSET = 1, /* int64_t, default 0 */
    ACLBLASLT_MATRIX_LAYOUT_TYPE = 2,                 /* uint32_t (aclDataType) */
    ACLBLASLT_MATRIX_LAYOUT_ORDER = 3,                /* int32_t (aclblasLtOrder_t), default COL */
    ACLBLASLT_MATRIX_LAYOUT_ROWS = 4,                 /* uint64_t */
    ACLBLASLT_MATRIX_LAYOUT_COLS = 5,                 /* uint64_t */
    ACLBLASLT_MATRIX_LAYOUT_LD = 6                    /* int64_t, elements */
} aclblasLtMatrixLayoutAttribute_t;

typedef enum aclblasLtMatmulDescAttribute {
    ACLBLASLT_MATMUL_DESC_TRANSA = 0,           /* int32_t */
    ACLBLASLT_MATMUL_DESC_TRANSB = 1,           /* int32_t */
    ACLBLASLT_MATMUL_DESC_EPILOGUE = 2,         /* uint32_t */
    ACLBLASLT_MATMUL_DESC_BIAS_POINTER = 3,     /* void* */
    ACLBLASLT_MATMUL_DESC_BIAS_DATA_TYPE = 4,   /* int32_t (aclDataType) */
    ACLBLASLT_MATMUL_DESC_A_SCALE_POINTER = 5,  /* void*: the E8M0 block scales of an MX A */
    ACLBLASLT_MATMUL_DESC_B_SCALE_POINTER = 6,  /* void*: the same for B */
    ACLBLASLT_MATMUL_DESC_C_SCALE_POINTER = 7,
    ACLBLASLT_MATMUL_DESC_D_SCALE_POINTER = 8,
    ACLBLASLT_MATMUL_DESC_EPILOGUE_AUX_SCALE_POINTER = 9,
    ACLBLASLT_MATMUL_DESC_EPILOGUE_AUX_POINTER = 10,
    ACLBLASLT_MATMUL_DESC_EPILOGUE_AUX_LD = 11,
    ACLBLASLT_MATMUL_DESC_EPILOGUE_AUX_BATCH_STRIDE = 12,
    ACLBLASLT_MATMUL_DESC_POINTER_MODE = 13,
    ACLBLASLT_MATMUL_DESC_AMAX_D_POINTER = 14,
    ACLBLASLT_MATMUL_DESC_EPILOGUE_AUX_DATA_TYPE = 22,
    ACLBLASLT_MATMUL_DESC_A_SCALE_MODE = 31,    /* accepted and ignored, as the reference */
    ACLBLASLT_MATMUL_DESC_B_SCALE_MODE = 32,
    ACLBLASLT_MATMUL_DESC_COMPUTE_INPUT_TYPE_A_EXT = 100,
    ACLBLASLT_MATMUL_DESC_COMPUTE_INPUT_TYPE_B_EXT = 101,
    ACLBLASLT_MATMUL_DESC_EPILOGUE_ACT_ARG0_EXT = 102,
    ACLBLASLT_MATMUL_DESC_EPILOGUE_ACT_ARG1_EXT = 103,
    ACLBLASLT_MATMUL_DESC_MAX
} aclblasLtMatmulDescAttribute_t;

typedef enum aclblasLtMatmulPreferenceAttribute {
    ACLBLASLT_MATMUL_PREF_SEARCH_MODE = 0,         /* uint32_t */
    ACLBLASLT_MATMUL_PREF_MAX_WORKSPACE_BYTES = 1, /* uint64_t */
    ACLBLASLT_MATMUL_PREF_MAX = 2
} aclblasLtMatmulPreferenceAttribute_t;

/* Version (reference :301-306): major << 24 | minor << 16 | patch, 1.0.0. */
aclblasStatus_t aclblasLtGetVersion(size_t* version);
aclblasStatus_t aclblasLtGetProperty(aclblasLtPropertyType_t type, int* value);

/* Handle (reference :329, :349): bound to the device current at creation. */
aclblasStatus_t aclblasLtCreate(aclblasLtHandle_t* lightHandle);
aclblasStatus_t aclblasLtDestroy(const aclblasLtHandle_t lightHandle);

/* Matrix layouts (reference :378-433): ld 0 means rows; ld < 0 is INVALID_VALUE. */
aclblasStatus_t aclblasLtMatrixLayoutCreate(aclblasLtMatrixLayout_t* matLayout, aclDataType type, uint64_t rows,
                                            uint64_t cols, int64_t ld);
aclblasStatus_t aclblasLtMatrixLayoutDestroy(const aclblasLtMatrixLayout_t matLayout);
aclblasStatus_t aclblasLtMatrixLayoutSetAttribute(aclblasLtMatrixLayout_t matLayout,
                                                  aclblasLtMatrixLayoutAttribute_t attr, const void* buf,
                                                  size_t sizeInBytes);
aclblasStatus_t aclblasLtMatrixLayoutGetAttribute(const aclblasLtMatrixLayout_t matLayout,
                                                  aclblasLtMatrixLayoutAttribute_t attr, void* buf,
                                                  size_t sizeInBytes, size_t* sizeWritten);

/* Matmul descriptors (reference :457-559).  Setting an attribute the matmul
 * does not use returns NOT_SUPPORTED, as the reference. */
aclblasStatus_t aclblasLtMatmulDescCreate(aclblasLtMatmulDesc_t* matmulDesc, aclblasComputeType_t computeType,
                                          aclDataType scaleType);
aclblasStatus_t aclblasLtMatmulDescDestroy(const aclblasLtMatmulDesc_t matmulDesc);
aclblasStatus_t aclblasLtMatmulDescInit(aclblasLtMatmulDesc_t matmulDesc, aclblasComputeType_t computeType,
                                        aclDataType scaleType);
aclblasStatus_t aclblasLtMatmulDescSetAttribute(aclblasLtMatmulDesc_t matmulDesc,
                                                aclblasLtMatmulDescAttribute_t attr, const void* buf,
                                                size_t sizeInBytes);
aclblasStatus_t aclblasLtMatmulDescGetAttribute(aclblasLtMatmulDesc_t desc, aclblasLtMatmulDescAttribute_t attr,
                                                void* buf, size_t sizeInBytes, size_t* sizeWritten);

/* Preferences (reference :578-658). */
aclblasStatus_t aclblasLtMatmulPreferenceCreate(aclblasLtMatmulPreference_t* pref);
aclblasStatus_t aclblasLtMatmulPreferenceDestroy(const aclblasLtMatmulPreference_t pref);
aclblasStatus_t aclblasLtMatmulPreferenceSetAttribute(aclblasLtMatmulPreference_t pref,
                                                      aclblasLtMatmulPreferenceAttribute_t attr, const void* buf,
                                                      size_t sizeInBytes);
aclblasStatus_t aclblasLtMatmulPreferenceGetAttribute(aclblasLtMatmulPreference_t pref,
                                                      aclblasLtMatmulPreferenceAttribute_t attr, void* buf,
                                                      size_t sizeInBytes, size_t* sizeWritten);

/* Algorithms (reference :689-845): one algorithm (id 0, no workspace) per
 * supported type combination; the heuristic returns it, or NOT_SUPPORTED. */
aclblasStatus_t aclblasLtMatmulAlgoInit(aclblasLtHandle_t lightHandle, aclblasComputeType_t computeType,
                                        aclDataType scaleType, aclDataType Atype, aclDataType Btype,
                                        aclDataType Ctype, aclDataType Dtype, int algoId,
                                        aclblasLtMatmulAlgo_t* algo);
aclblasStatus_t aclblasLtMatmulAlgoGetIds(aclblasLtHandle_t lightHandle, aclblasComputeType_t computeType,
                                          aclDataType scaleType, aclDataType Atype, aclDataType Btype,
                                          aclDataType Ctype, aclDataType Dtype, int* algoIdsArray,
                                          int algoIdsArrayLength, int* numAlgoIds);
aclblasStatus_t aclblasLtMatmulAlgoGetHeuristic(aclblasLtHandle_t lightHandle, aclblasLtMatmulDesc_t matmulDesc,
                                                aclblasLtMatrixLayout_t Adesc, aclblasLtMatrixLayout_t Bdesc,
                                                aclblasLtMatrixLayout_t Cdesc, aclblasLtMatrixLayout_t Ddesc,
                                                aclblasLtMatmulPreference_t pref, int requestedAlgoCount,
                                                aclblasLtMatmulHeuristicResult_t heuristicResultsArray[],
                                                int* returnAlgoCount);

/* D = alpha op(A) op(B) + beta C (reference :901-917, checks as
 * blasLt/api/aclblaslt_matmul.cpp:30-101 and
 * blasLt/matmul/engine/matmul_engine.cpp:100-161).  Row-major matrices:
 * m = rows(D), n = cols(D), k = cols(A) (transA N) or rows(A) (T); A is m x k
 * (N) or k x m (T), B is k x n (N) or n x k (T), each with its layout's ld
 * (elements; for FP4 E2M1 the bytes per row, two elements per byte with the
 * first in the low nibble, as the reference's callers create FP4 layouts).
 * alpha, beta: host floats.  C may be NULL when beta == 0.  workspace: may be
 * NULL (none is used); when given it must be 16-byte aligned.
 * MX scale layouts (reference test/blasLtMatmul/blasLtMatmul_golden.h:77-91):
 * with S = 2 ceil(K / 64) scale bytes per row, block b = k / 32,
 *   A (transA N) and B (transB T): scale[row * S + b]        (row = m or n)
 *   A (transA T) and B (transB N): scale[(b / 2) * 2 R + 2 row + b % 2]
 *                                  (R = m for A, n for B).
 * A scale byte of 0 makes its 32-element block contribute 0 (the reference's
 * golden decoding); 255 is NaN. */
aclblasStatus_t aclblasLtMatmul(aclblasLtHandle_t lightHandle, aclblasLtMatmulDesc_t computeDesc, const void* alpha,
                                const void* A, aclblasLtMatrixLayout_t Adesc, const void* B,
                                aclblasLtMatrixLayout_t Bdesc, const void* beta, const void* C,
                                aclblasLtMatrixLayout_t Cdesc, void* D, aclblasLtMatrixLayout_t Ddesc,
                                const aclblasLtMatmulAlgo_t* algo, void* workspace, size_t workspaceSizeInBytes,
                                aclrtStream stream);

#ifdef __cplusplus
}
#endif

#endif /* ACLBLAS_AMD_CANN_OPS_BLASLT_H */
