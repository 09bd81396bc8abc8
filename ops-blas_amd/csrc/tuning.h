// Launch choices per size bucket.
//
// Every call that has more than one sensible launch shape (GEMM tile and L2
// tile-group height, grouped-GEMM tile and persistent grid, mixed-solve
// workgroups per compute unit) maps its problem to a bucket and looks the
// choice up in a table measured on the target GPU by ops-blas_amd/autotune.py
// (generated into tuning_table.inc).  Buckets the table does not hold use the
// built-in fallback of the call.  A handle-scoped override
// (aclblasSetTuningOverride) forces one choice for every bucket of an
// operation; the tuner uses it to time each candidate.
#pragma once

struct _aclblas_handle;

namespace aclblas {
namespace tune {

enum Op : int {
    kOpGemm = 0,     // a = tile edge (64, 128, 256; FP32 also 32 and 65 = 64 with deep prefetch),
                     // b = tile rows per L2 group (1..64)
    kOpGrouped = 1,  // a = tile edge (64, 128), b = workgroups per CU of the persistent grid (1..4); class 0:
                     // FP32 (SgemmGroupedBatched), 1: 16-bit and FP8 operands (GemmGroupedBatchedEx)
    kOpMixed = 2,    // a = factor workgroups per CU (1..4), b = solve workgroups per CU (1..4); class 0: one-
                     // and two-plane BF16 trailing updates, 1: FP32 and three planes
    kOpLu = 3,       // a = rows per lane of the n <= 64 LU (1, 2; 1 only for n <= 8), b ignored
    kNumOps = 4
};

struct Choice {
    int a;
    int b;
};

// Element classes of the GEMM buckets.
enum GemmClass : int { kClassF32 = 0, kClass16 = 1, kClassF8 = 2, kClassF8Scaled = 3, kNumGemmClasses = 4 };

constexpr int kGemmSizeBuckets = 5;
constexpr int kGemmKBuckets = 3;
constexpr int kGroupedSizeBuckets = 3;
constexpr int kGroupedBuckets = 2 * kGroupedSizeBuckets;  // element class x tile count
constexpr int kMixedSizeBuckets = 4;
constexpr int kMixedBuckets = 2 * kMixedSizeBuckets;  // update class x system count
constexpr int kLuBuckets = 16;

// GEMM bucket from the element class, the number of 128 x 128 output tiles
// over the whole call (batch included), k and the compute-unit count.
int GemmBucket(int cls, long long tiles128, int k, int cus);
// Grouped GEMM bucket from the element class (0 FP32, 1 16-bit / FP8) and the
// total number of 128 x 128 tiles of all problems.
int GroupedBucket(int cls, long long tiles128, int cus);
// Mixed-solve bucket from the number of systems.
int MixedBucket(int cls, long long count, int cus);
// n <= 64 LU bucket from the padded order S (8, 16, 32, 64) and the waves a
// one-row-per-lane launch would have (batch * S / 64) relative to the SIMD
// count (4 per CU): below 2, 4, 8 per SIMD, or more.
int LuBucket(int S, long long batch, int cus);
// Number of buckets of an operation (0 for an unknown one).
int BucketCount(int op);

// True when `c` is a legal choice for the operation (and, for GEMM, the
// element class of the bucket).
bool Valid(int op, int bucket, Choice c);

// Choice for (op, bucket): the handle's override, else the measured table,
// else `fallback`.
Choice Get(const _aclblas_handle* h, int op, int bucket, Choice fallback);

}  // namespace tune
}  // namespace aclblas
