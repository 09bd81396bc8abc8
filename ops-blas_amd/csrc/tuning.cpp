// Bucket functions, the measured choice table and the tuning entry points.
#include "tuning.h"

#include "cann_ops_blas.h"
#include "internal.h"

namespace aclblas {
namespace tune {
namespace {

struct Entry {
    int op;
    int bucket;
    int a;
    int b;
};

// {op, bucket, a, b} rows written by autotune.py; the last row is a sentinel.
const Entry kTable[] = {
#include "tuning_table.inc"
    {-1, -1, 0, 0},
};

int SizeBucket(long long units, int cus, const long long (&edges)[4], int count)
{
    for (int i = 0; i < count - 1; ++i) {
        if (units < edges[i] * cus / 4) {
            return i;
        }
    }
    return count - 1;
}

}  // namespace

int GemmBucket(int cls, long long tiles128, int k, int cus)
{
    // tile-count edges in quarter compute units: < cus/4, < cus, < 4 cus, < 16 cus, more
    static const long long kEdges[4] = {1, 4, 16, 64};
    const int s = SizeBucket(tiles128, cus, kEdges, kGemmSizeBuckets);
    const int kb = k < 256 ? 0 : (k < 2048 ? 1 : 2);
    return (cls * kGemmSizeBuckets + s) * kGemmKBuckets + kb;
}

int GroupedBucket(int cls, long long tiles128, int cus)
{
    // < cus, < 4 cus, more; then the element class
    static const long long kEdges[4] = {4, 16, 0, 0};
    return cls * kGroupedSizeBuckets + SizeBucket(tiles128, cus, kEdges, kGroupedSizeBuckets);
}

int MixedBucket(int cls, long long count, int cus)
{
    // < cus, < 4 cus, < 16 cus, more; then the update class (0: BF16, 1: FP32)
    static const long long kEdges[4] = {4, 16, 64, 0};
    return cls * kMixedSizeBuckets + SizeBucket(count, cus, kEdges, kMixedSizeBuckets);
}

int LuBucket(int S, long long batch, int cus)
{
    const int si = S <= 8 ? 0 : (S <= 16 ? 1 : (S <= 32 ? 2 : 3));
    // waves of a one-row-per-lane launch: < 2, < 4, < 8 per SIMD (4 SIMDs per CU), more
    static const long long kEdges[4] = {32, 64, 128, 0};
    const int w = SizeBucket(batch * S / 64, cus, kEdges, 4);
    return si * 4 + w;
}

int BucketCount(int op)
{
    switch (op) {
        case kOpGemm:
            return kNumGemmClasses * kGemmSizeBuckets * kGemmKBuckets;
        case kOpGrouped:
            return kGroupedBuckets;
        case kOpMixed:
            return kMixedBuckets;
        case kOpLu:
            return kLuBuckets;
        default:
            return 0;
    }
}

bool Valid(int op, int bucket, Choice c)
{
    if (bucket < 0 || bucket >= BucketCount(op)) {
        return false;
    }
    switch (op) {
        case kOpGemm: {
            const int cls = bucket / (kGemmSizeBuckets * kGemmKBuckets);
            const bool tileOk = c.a == 64 || c.a == 128 || (c.a == 256 && cls != kClassF32) ||
                                ((c.a == 32 || c.a == 65) && cls == kClassF32);
            return tileOk && c.b >= 1 && c.b <= 64;
        }
        case kOpGrouped:
            return (c.a == 64 || c.a == 128) && c.b >= 1 && c.b <= 4;
        case kOpMixed:
            return c.a >= 1 && c.a <= 4 && c.b >= 1 && c.b <= 4;
        case kOpLu:
            return c.a == 1 || (c.a == 2 && bucket >= 4);
        default:
            return false;
    }
}

Choice Get(const _aclblas_handle* h, int op, int bucket, Choice fallback)
{
    if (h != nullptr && op >= 0 && op < kNumOps && h->tune_override[op][0] != 0) {
        const Choice o{h->tune_override[op][1], h->tune_override[op][2]};
        if (Valid(op, bucket, o)) {
            return o;
        }
    }
    for (const Entry& e : kTable) {
        if (e.op == op && e.bucket == bucket) {
            const Choice c{e.a, e.b};
            return Valid(op, bucket, c) ? c : fallback;
        }
    }
    return fallback;
}

}  // namespace tune
}  // namespace aclblas

extern "C" aclblasStatus_t aclblasSetTuningOverride(aclblasHandle_t handle, int op, int a, int b)
{
    using namespace aclblas;
    if (handle == nullptr) {
        return ACLBLAS_STATUS_HANDLE_IS_NULLPTR;
    }
    if (op < 0 || op >= tune::kNumOps) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (a < 0) {
        handle->tune_override[op][0] = 0;
        return ACLBLAS_STATUS_SUCCESS;
    }
    // legal for at least one bucket of the operation
    bool legal = false;
    for (int bucket = 0; bucket < tune::BucketCount(op) && !legal; ++bucket) {
        legal = tune::Valid(op, bucket, tune::Choice{a, b});
    }
    if (!legal) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    handle->tune_override[op][0] = 1;
    handle->tune_override[op][1] = a;
    handle->tune_override[op][2] = b;
    return ACLBLAS_STATUS_SUCCESS;
}

extern "C" aclblasStatus_t aclblasGetTuningChoice(aclblasHandle_t handle, int op, int bucket, int* a, int* b)
{
    using namespace aclblas;
    if (handle == nullptr) {
        return ACLBLAS_STATUS_HANDLE_IS_NULLPTR;
    }
    if (a == nullptr || b == nullptr || bucket < 0 || bucket >= tune::BucketCount(op)) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    const tune::Choice c = tune::Get(handle, op, bucket, tune::Choice{0, 0});
    *a = c.a;
    *b = c.b;
    return ACLBLAS_STATUS_SUCCESS;
}

extern "C" aclblasStatus_t aclblasTuningBucket(aclblasHandle_t handle, int op, int cls, long long units, int k,
                                               int* bucket)
{
    using namespace aclblas;
    if (handle == nullptr) {
        return ACLBLAS_STATUS_HANDLE_IS_NULLPTR;
    }
    if (bucket == nullptr || units < 0) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    const int cus = handle->cu_count > 0 ? handle->cu_count : 256;
    switch (op) {
        case tune::kOpGemm:
            if (cls < 0 || cls >= tune::kNumGemmClasses || k < 0) {
                return ACLBLAS_STATUS_INVALID_VALUE;
            }
            *bucket = tune::GemmBucket(cls, units, k, cus);
            return ACLBLAS_STATUS_SUCCESS;
        case tune::kOpGrouped:
            if (cls < 0 || cls > 1) {
                return ACLBLAS_STATUS_INVALID_VALUE;
            }
            *bucket = tune::GroupedBucket(cls, units, cus);
            return ACLBLAS_STATUS_SUCCESS;
        case tune::kOpMixed:
            if (cls < 0 || cls > 1) {
                return ACLBLAS_STATUS_INVALID_VALUE;
            }
            *bucket = tune::MixedBucket(cls, units, cus);
            return ACLBLAS_STATUS_SUCCESS;
        case tune::kOpLu:
            if (cls < 1 || cls > 64) {
                return ACLBLAS_STATUS_INVALID_VALUE;
            }
            *bucket = tune::LuBucket(cls <= 8 ? 8 : (cls <= 16 ? 16 : (cls <= 32 ? 32 : 64)), units, cus);
            return ACLBLAS_STATUS_SUCCESS;
        default:
            return ACLBLAS_STATUS_INVALID_VALUE;
    }
}
