/*
 * Minimal ACL runtime types needed by the aclblas C API, defined for the HIP
 * runtime on MI355X (gfx950).
 *
 * The reference library includes the CANN SDK's <acl/acl.h> for exactly three
 * things: the stream handle type, the tensor element-type enum and the error
 * type (reference include/cann_ops_blas.h:13).  This header provides those
 * three with the same names so code written against the reference compiles
 * unchanged.  aclrtStream is the HIP stream handle (a pointer to the opaque
 * ihipStream_t), declared here without pulling in the HIP headers so that a
 * C caller needs nothing but this file.
 *
 * aclDataType values: ACL_FLOAT=0, ACL_FLOAT16=1 and ACL_BF16=27 are the values
 * the reference states (reference test/frame/csv_loader.h:330).  The FP8 values
 * are not stated anywhere in the reference's library headers; 35 (E5M2) and 36
 * (E4M3FN) follow the public CANN enum numbering, and 36 and 40 (FP4 E2M1, the
 * BLASLt MX format) are the values the reference's BLASLt tests parse
 * (reference test/blasLtMatmul/blasLtMatmul_param.h:21-31).
 */
#ifndef ACLBLAS_AMD_ACL_ACL_H
#define ACLBLAS_AMD_ACL_ACL_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* aclrtStream;

typedef int aclError;
#define ACL_SUCCESS 0

typedef enum {
    ACL_DT_UNDEFINED = -1,
    ACL_FLOAT = 0,
    ACL_FLOAT16 = 1,
    ACL_INT8 = 2,
    ACL_INT32 = 3,
    ACL_UINT8 = 4,
    ACL_INT16 = 6,
    ACL_UINT16 = 7,
    ACL_UINT32 = 8,
    ACL_INT64 = 9,
    ACL_UINT64 = 10,
    ACL_DOUBLE = 11,
    ACL_BOOL = 12,
    ACL_BF16 = 27,
    ACL_FLOAT8_E5M2 = 35,
    ACL_FLOAT8_E4M3FN = 36,
    ACL_FLOAT4_E2M1 = 40
} aclDataType;

#ifdef __cplusplus
}
#endif

#endif /* ACLBLAS_AMD_ACL_ACL_H */
