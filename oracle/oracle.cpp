// CPU oracle for the aclblas GEMM / batched-LU hot path.
//
// TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by the
// tests, by the smoke check and by the benchmark's CPU-baseline leg, never by
// the product library, which has no CPU fallback.
//
// It restates, in plain C++, the algorithms of the reference's golden path and
// of its device kernels for the operations the GPU library implements:
//   - the seeded input generators of the reference's tests (std::mt19937 +
//     std::uniform_real_distribution, exactly as the reference calls them);
//   - batched LU with partial pivoting, the matching triangular solves and the
//     inverse, with the reference's pivot, zero-pivot and info rules;
//   - GEMM with FP64 accumulation and the FP16/BF16/FP8 conversions of the
//     reference's host helpers, and the BLASLt matmul golden (FP32, MXFP8,
//     MXFP4 with E8M0 block scales);
//   - an FP64 residual / backward-error evaluator (the "FP64 oracle").
// Each function cites the reference file:line it follows.  How each is pinned
// (golden vectors, LAPACK through scipy, closed forms) is stated in the tests.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// Threading helper: runs body(b) for b in [0, count) on `threads` threads.
// ---------------------------------------------------------------------------
template <typename F>
void ParallelFor(int count, int threads, F body)
{
    if (threads <= 1 || count <= 1) {
        for (int b = 0; b < count; ++b) {
            body(b);
        }
        return;
    }
    std::atomic<int> next(0);
    std::vector<std::thread> pool;
    const int nt = std::min(threads, count);
    pool.reserve(nt);
    for (int t = 0; t < nt; ++t) {
        pool.emplace_back([&]() {
            for (int b = next.fetch_add(1); b < count; b = next.fetch_add(1)) {
                body(b);
            }
        });
    }
    for (auto& th : pool) {
        th.join();
    }
}

// Pivot rule of the LU (see orc_getrf_batched).
inline float PivotMag(float v)
{
    const float a = std::fabs(v);
    return (a == a) ? a : 0.0f;
}

// ---------------------------------------------------------------------------
// One matrix: right-looking unblocked LU, reference
// blas/getrf_batched/arch35/sgetrf_batched_kernel.cpp:106-150.
// ---------------------------------------------------------------------------
int GetrfOne(float* A, int n, int lda, int* piv)
{
    for (int k = 0; k < n; ++k) {
        float* ck = A + static_cast<size_t>(k) * lda;
        int p = k;
        if (piv != nullptr) {
            // FindPivotRow (:52-79): candidates k..n-1, strict '>' against a
            // running maximum that starts at 0 with row k, so NaN and zero
            // never win; the scan keeps the lowest index on ties (LAPACK isamax).
            float best = 0.0f;
            for (int i = k; i < n; ++i) {
                const float m = PivotMag(ck[i]);
                if (m > best) {
                    best = m;
                    p = i;
                }
            }
            piv[k] = p + 1;
        }
        const float pivot = ck[p];
        if (pivot == 0.0f) {
            return k + 1;  // :128-137: record info, stop; later pivots untouched
        }
        if (p != k) {  // SwapRows (:81-91): whole row
            for (int j = 0; j < n; ++j) {
                float* cj = A + static_cast<size_t>(j) * lda;
                std::swap(cj[k], cj[p]);
            }
        }
        for (int i = k + 1; i < n; ++i) {  // ComputeMultipliers (:93-104): division
            ck[i] = ck[i] / pivot;
        }
        for (int j = k + 1; j < n; ++j) {  // Rank1Update (:38-50), one rounding
            float* cj = A + static_cast<size_t>(j) * lda;
            const float u = cj[k];
            for (int i = k + 1; i < n; ++i) {
                cj[i] = std::fmaf(-ck[i], u, cj[i]);
            }
        }
    }
    return 0;
}

// perm with (P b)[i] = b[perm[i]] from the LAPACK swap list; out-of-range and
// self pivots are skipped (reference sgetrs_batched_kernel.cpp:95-98).
void BuildPerm(const int* piv, int n, std::vector<int>& perm)
{
    perm.resize(n);
    for (int i = 0; i < n; ++i) {
        perm[i] = i;
    }
    if (piv == nullptr) {
        return;
    }
    for (int k = 0; k < n; ++k) {
        const int p = piv[k] - 1;
        if (p >= 0 && p < n && p != k) {
            std::swap(perm[k], perm[p]);
        }
    }
}

// Column-oriented substitutions with one rounding per update, in the order of
// reference sgetrs_batched_kernel.cpp:167-259 (N) and :263-362 (T).
void SolveN(const float* A, int lda, int n, float* x)
{
    for (int j = 0; j < n; ++j) {  // L, unit diagonal
        const float* cj = A + static_cast<size_t>(j) * lda;
        const float xj = x[j];
        for (int i = j + 1; i < n; ++i) {
            x[i] = std::fmaf(-cj[i], xj, x[i]);
        }
    }
    for (int j = n - 1; j >= 0; --j) {  // U
        const float* cj = A + static_cast<size_t>(j) * lda;
        x[j] = x[j] / cj[j];
        const float xj = x[j];
        for (int i = 0; i < j; ++i) {
            x[i] = std::fmaf(-cj[i], xj, x[i]);
        }
    }
}

void SolveT(const float* A, int lda, int n, float* x)
{
    for (int j = 0; j < n; ++j) {  // U^T
        const float* cj = A + static_cast<size_t>(j) * lda;
        x[j] = x[j] / cj[j];
        const float xj = x[j];
        for (int i = j + 1; i < n; ++i) {
            x[i] = std::fmaf(-A[j + static_cast<size_t>(i) * lda], xj, x[i]);
        }
    }
    for (int j = n - 1; j >= 0; --j) {  // L^T, unit diagonal
        const float xj = x[j];
        for (int i = 0; i < j; ++i) {
            x[i] = std::fmaf(-A[j + static_cast<size_t>(i) * lda], xj, x[i]);
        }
    }
}

// Round-to-nearest-even float -> bf16 with NaN quieting, reference
// blas/common/helper/dtype_cast.h:92-105.
uint16_t F32ToBf16(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7FFFFFFFu) > 0x7F800000u) {
        return static_cast<uint16_t>((u >> 16) | 0x40u);
    }
    const uint32_t lsb = (u >> 16) & 1u;
    u += 0x7FFFu + lsb;
    return static_cast<uint16_t>(u >> 16);
}

float Bf16ToF32(uint16_t h)
{
    const uint32_t u = static_cast<uint32_t>(h) << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// IEEE binary16, round to nearest even (reference dtype_cast.h:31-90).
uint16_t F32ToF16(float f)
{
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t absx = x & 0x7FFFFFFFu;
    if (absx > 0x7F800000u) {
        return static_cast<uint16_t>(sign | 0x7E00u);
    }
    if (absx >= 0x477FF000u) {  // rounds to >= 65520 -> inf
        return static_cast<uint16_t>(sign | 0x7C00u);
    }
    if (absx < 0x38800000u) {  // subnormal or zero in half
        float a;
        std::memcpy(&a, &absx, 4);
        const float scaled = a * 16777216.0f;  // 2^24: half subnormal lsb = 2^-24
        const uint32_t m = static_cast<uint32_t>(std::nearbyint(scaled));
        return static_cast<uint16_t>(sign | m);
    }
    uint32_t mant = absx & 0x7FFFFFu;
    uint32_t exp = (absx >> 23) - 127 + 15;
    uint32_t h = (exp << 10) | (mant >> 13);
    const uint32_t rem = mant & 0x1FFFu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) {
        h += 1;
    }
    return static_cast<uint16_t>(sign | h);
}

float F16ToF32(uint16_t h)
{
    const uint32_t sign = (h & 0x8000u) << 16;
    const uint32_t exp = (h >> 10) & 0x1Fu;
    const uint32_t mant = h & 0x3FFu;
    float f;
    if (exp == 0) {
        f = std::ldexp(static_cast<float>(mant), -24);
        if (sign) {
            f = -f;
        }
        return f;
    }
    uint32_t u;
    if (exp == 31) {
        u = sign | 0x7F800000u | (mant << 13);
    } else {
        u = sign | ((exp - 15 + 127) << 23) | (mant << 13);
    }
    std::memcpy(&f, &u, 4);
    return f;
}

// OCP FP8 decode.  E4M3FN: bias 7, no infinities, 0x7F/0xFF = NaN, max 448.
// E5M2: bias 15, IEEE-like with inf and NaN, max 57344.
float Fp8ToF32(uint8_t v, bool e4m3)
{
    const int sign = (v >> 7) & 1;
    float r;
    if (e4m3) {
        const int exp = (v >> 3) & 0xF;
        const int man = v & 0x7;
        if (exp == 0xF && man == 0x7) {
            return std::nanf("");
        }
        r = exp == 0 ? std::ldexp(static_cast<float>(man), -9)
                     : std::ldexp(1.0f + static_cast<float>(man) / 8.0f, exp - 7);
    } else {
        const int exp = (v >> 2) & 0x1F;
        const int man = v & 0x3;
        if (exp == 0x1F) {
            return man == 0 ? (sign ? -INFINITY : INFINITY) : std::nanf("");
        }
        r = exp == 0 ? std::ldexp(static_cast<float>(man), -16)
                     : std::ldexp(1.0f + static_cast<float>(man) / 4.0f, exp - 15);
    }
    return sign ? -r : r;
}

// Saturating round-to-nearest-even encode (reference test/frame/dtype_utils.h:
// 39-160 saturates finite values to the format's maximum; NaN stays NaN; E4M3
// maps +-inf to +-max, E5M2 keeps inf).
uint8_t F32ToFp8(float f, bool e4m3)
{
    const uint8_t sign = std::signbit(f) ? 0x80 : 0x00;
    if (std::isnan(f)) {
        return e4m3 ? 0x7F : 0x7E;
    }
    float a = std::fabs(f);
    const float maxv = e4m3 ? 448.0f : 57344.0f;
    if (std::isinf(a)) {
        return e4m3 ? static_cast<uint8_t>(sign | 0x7E) : static_cast<uint8_t>(sign | 0x7C);
    }
    if (a >= maxv) {
        return e4m3 ? static_cast<uint8_t>(sign | 0x7E) : static_cast<uint8_t>(sign | 0x7B);
    }
    // brute-force nearest (ties to even code) over the 127 non-negative codes
    int bestCode = 0;
    double bestErr = 1e300;
    for (int c = 0; c < 128; ++c) {
        const float v = Fp8ToF32(static_cast<uint8_t>(c), e4m3);
        if (!std::isfinite(v)) {
            continue;
        }
        const double err = std::fabs(static_cast<double>(v) - a);
        if (err < bestErr || (err == bestErr && (c & 1) == 0)) {
            bestErr = err;
            bestCode = c;
        }
    }
    return static_cast<uint8_t>(sign | bestCode);
}

// aclDataType values used by the GEMM oracle.
constexpr int kDtF32 = 0;
constexpr int kDtF16 = 1;
constexpr int kDtBf16 = 27;
constexpr int kDtE5M2 = 35;
constexpr int kDtE4M3 = 36;

float LoadAs(const void* p, size_t idx, int dt)
{
    switch (dt) {
        case kDtF32: return static_cast<const float*>(p)[idx];
        case kDtF16: return F16ToF32(static_cast<const uint16_t*>(p)[idx]);
        case kDtBf16: return Bf16ToF32(static_cast<const uint16_t*>(p)[idx]);
        case kDtE5M2: return Fp8ToF32(static_cast<const uint8_t*>(p)[idx], false);
        case kDtE4M3: return Fp8ToF32(static_cast<const uint8_t*>(p)[idx], true);
        default: return 0.0f;
    }
}

void StoreAs(void* p, size_t idx, int dt, float v)
{
    switch (dt) {
        case kDtF32: static_cast<float*>(p)[idx] = v; break;
        case kDtF16: static_cast<uint16_t*>(p)[idx] = F32ToF16(v); break;
        case kDtBf16: static_cast<uint16_t*>(p)[idx] = F32ToBf16(v); break;
        default: break;
    }
}

// Box-Muller normal generator over mt19937_64 (fully specified, identical on
// every host with a conforming std::mt19937_64).
struct Normal {
    std::mt19937_64 rng;
    bool has = false;
    double spare = 0.0;
    explicit Normal(uint64_t seed) : rng(seed) {}
    double U01()
    {
        return (static_cast<double>(rng() >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    }
    double Next()
    {
        if (has) {
            has = false;
            return spare;
        }
        const double u1 = U01();
        const double u2 = U01();
        const double r = std::sqrt(-2.0 * std::log(u1));
        const double th = 6.283185307179586476925286766559 * u2;
        spare = r * std::sin(th);
        has = true;
        return r * std::cos(th);
    }
};

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------
// Generators
// ---------------------------------------------------------------------------

// RANDOM_v fill of makeBlasArray / makeBlasMatrix: std::mt19937(seed ? seed :
// 42), std::uniform_real_distribution<float>(lo, hi), `count` draws in storage
// order.  reference test/frame/fill.h:280-288 and :497-506.  map_zero: apply
// the seed==0 -> 42 rule.
void orc_fill_uniform(float* out, size_t count, uint32_t seed, float lo, float hi, int map_zero)
{
    std::mt19937 rng(map_zero ? (seed ? seed : 42u) : seed);
    std::uniform_real_distribution<float> dist(lo, hi);
    for (size_t i = 0; i < count; ++i) {
        out[i] = dist(rng);
    }
}

// makeBlasLapackMatrix, reference test/frame/fill.h:586-730.  type: 0 random
// nonsingular, 1 identity, 2 diagonally dominant, 3 Hilbert, 4 diagonal,
// 5 upper triangular, 6 lower triangular, 7 singular zero column, 8 singular
// dependent row, 9 mixed.  out holds lda*max(1,n) floats, zero-filled here.
void orc_lapack_matrix(float* out, int n, int lda, int type, uint32_t seed, int batchIdx)
{
    const size_t total = static_cast<size_t>(lda) * std::max(1, n);
    std::fill(out, out + total, 0.0f);
    if (n <= 0) {
        return;
    }
    std::mt19937 rng(seed + static_cast<uint32_t>(batchIdx) * 1000u);
    std::uniform_real_distribution<float> dist01(0.0f, 1.0f);
    auto randomNonsingular = [&]() {
        for (int j = 0; j < n; ++j) {
            for (int i = 0; i < n; ++i) {
                out[j * lda + i] = dist01(rng);
            }
            out[j * lda + j] += static_cast<float>(n);
        }
    };
    switch (type) {
        case 0: randomNonsingular(); break;
        case 1:
            for (int j = 0; j < n; ++j) out[j * lda + j] = 1.0f;
            break;
        case 2:
            for (int j = 0; j < n; ++j) {
                for (int i = 0; i < n; ++i) out[j * lda + i] = dist01(rng);
                out[j * lda + j] = static_cast<float>(n);
            }
            break;
        case 3:
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) out[j * lda + i] = 1.0f / static_cast<float>(i + j + 1);
            break;
        case 4: {
            std::uniform_real_distribution<float> d(0.1f, 10.0f);
            for (int j = 0; j < n; ++j) out[j * lda + j] = d(rng);
            break;
        }
        case 5:
        case 6: {
            const bool upper = type == 5;
            std::uniform_real_distribution<float> dd(0.5f, 2.0f);
            std::uniform_real_distribution<float> od(0.0f, 1.0f);
            for (int j = 0; j < n; ++j) {
                for (int i = 0; i < n; ++i) {
                    const bool off = upper ? (i < j) : (i > j);
                    if (off) {
                        out[j * lda + i] = od(rng);
                    } else if (i == j) {
                        out[j * lda + i] = dd(rng);
                    }
                }
            }
            break;
        }
        case 7:
            randomNonsingular();
            for (int i = 0; i < n; ++i) out[i] = 0.0f;
            break;
        case 8: {
            randomNonsingular();
            const int dep = std::min(n - 1, 2);
            if (dep > 0)
                for (int j = 0; j < n; ++j) out[j * lda + dep] = 2.0f * out[j * lda + 0];
            break;
        }
        case 9:
            randomNonsingular();
            if (batchIdx % 2 == 1)
                for (int i = 0; i < n; ++i) out[i] = 0.0f;
            break;
        default: break;
    }
}

// Right-hand side of the getrs test: std::mt19937(seed + b*1000 + 500),
// U(-1, 1), columns 0..nrhs-1 rows 0..n-1 of an ldb-strided block.
// reference test/getrs_batched/arch35/getrs_batched_test.cpp:31-49.
void orc_getrs_rhs(float* out, int n, int nrhs, int ldb, uint32_t seed, int batchIdx)
{
    const size_t total = static_cast<size_t>(ldb) * std::max(1, nrhs);
    std::fill(out, out + total, 0.0f);
    if (n <= 0 || nrhs <= 0) {
        return;
    }
    std::mt19937 rng(seed + static_cast<uint32_t>(batchIdx) * 1000u + 500u);
    std::uniform_real_distribution<float> dist(-1.0f, 1.0f);
    for (int j = 0; j < nrhs; ++j)
        for (int i = 0; i < n; ++i) out[j * ldb + i] = dist(rng);
}

// Standard normal FP32 samples (mt19937_64 + Box-Muller), plus `shift` added
// on the diagonal of each of `batch` n x n column-major blocks when n > 0.
// Used for the synthetic benchmark inputs (A ~ N(0,1) + shift*I, b ~ N(0,1)).
void orc_fill_normal(float* out, size_t count, uint64_t seed, int n, int batch, float shift)
{
    Normal g(seed);
    for (size_t i = 0; i < count; ++i) {
        out[i] = static_cast<float>(g.Next());
    }
    if (n > 0) {
        for (int b = 0; b < batch; ++b)
            for (int j = 0; j < n; ++j) out[static_cast<size_t>(b) * n * n + static_cast<size_t>(j) * n + j] += shift;
    }
}

// ---------------------------------------------------------------------------
// Batched LAPACK
// ---------------------------------------------------------------------------

// Batched LU: matrices at A + b*strideA, pivots at ipiv + b*n (ipiv NULL = no
// pivoting), info[b] (may be NULL).  reference sgetrf_batched_kernel.cpp:
// 106-221 for the per-matrix algorithm; the no-pivot variant is the same
// elimination with the diagonal as pivot (:152-197).
void orc_getrf_batched(float* A, long long strideA, int n, int lda, int* ipiv, int* info, int batch, int threads)
{
    ParallelFor(batch, threads, [&](int b) {
        const int r = GetrfOne(A + b * strideA, n, lda, ipiv ? ipiv + static_cast<size_t>(b) * n : nullptr);
        if (info) {
            info[b] = r;
        }
    });
}

// Batched solve op(A) X = B with LU factors; trans 0 = N, 1 = T/C.
// reference sgetrs_batched_kernel.cpp:365-410.
void orc_getrs_batched(int trans, int n, int nrhs, const float* A, long long strideA, int lda, const int* ipiv,
                       float* B, long long strideB, int ldb, int batch, int threads)
{
    ParallelFor(batch, threads, [&](int b) {
        const float* Ab = A + b * strideA;
        float* Bb = B + b * strideB;
        std::vector<int> perm;
        BuildPerm(ipiv ? ipiv + static_cast<size_t>(b) * n : nullptr, n, perm);
        std::vector<float> x(n);
        for (int r = 0; r < nrhs; ++r) {
            float* col = Bb + static_cast<size_t>(r) * ldb;
            if (trans == 0) {
                for (int i = 0; i < n; ++i) x[i] = col[perm[i]];
                SolveN(Ab, lda, n, x.data());
                for (int i = 0; i < n; ++i) col[i] = x[i];
            } else {
                for (int i = 0; i < n; ++i) x[i] = col[i];
                SolveT(Ab, lda, n, x.data());
                for (int i = 0; i < n; ++i) col[perm[i]] = x[i];
            }
        }
    });
}

// Batched inverse C = A^-1 from LU factors.  info[b] = j+1 for the zero u_jj
// with the largest j (reference getri_batched_kernel.cpp:142-151: the scan runs
// from j = n-1 down); in that case C = L^-1 P (no backward substitution).
// Work is spread over (matrix, column) pairs.
void orc_getri_batched(int n, const float* A, long long strideA, int lda, const int* ipiv, float* C,
                       long long strideC, int ldc, int* info, int batch, int threads)
{
    std::vector<std::vector<int>> perms(batch);
    for (int b = 0; b < batch; ++b) {
        const float* Ab = A + b * strideA;
        BuildPerm(ipiv ? ipiv + static_cast<size_t>(b) * n : nullptr, n, perms[b]);
        int z = 0;
        for (int j = n - 1; j >= 0; --j) {
            if (Ab[j + static_cast<size_t>(j) * lda] == 0.0f) {
                z = j + 1;
                break;
            }
        }
        info[b] = z;
    }
    ParallelFor(batch * n, threads, [&](int bc) {
        const int b = bc / n;
        const int c = bc - b * n;
        const float* Ab = A + b * strideA;
        float* Cb = C + b * strideC;
        const std::vector<int>& perm = perms[b];
        std::vector<float> x(n);
        for (int i = 0; i < n; ++i) x[i] = (perm[i] == c) ? 1.0f : 0.0f;
        for (int j = 0; j < n; ++j) {  // L
            const float* cj = Ab + static_cast<size_t>(j) * lda;
            const float xj = x[j];
            for (int i = j + 1; i < n; ++i) x[i] = std::fmaf(-cj[i], xj, x[i]);
        }
        if (info[b] == 0) {
            for (int j = n - 1; j >= 0; --j) {  // U
                const float* cj = Ab + static_cast<size_t>(j) * lda;
                x[j] = x[j] / cj[j];
                const float xj = x[j];
                for (int i = 0; i < j; ++i) x[i] = std::fmaf(-cj[i], xj, x[i]);
            }
        }
        for (int i = 0; i < n; ++i) Cb[i + static_cast<size_t>(c) * ldc] = x[i];
    });
}

// Wall time of `reps` runs of getrf + getrs (nrhs right-hand sides, trans N)
// over copies of the given batch, on `threads` threads.  Returns seconds per
// run (median).  Inputs are left untouched.
double orc_time_getrf_getrs(const float* A, const float* B, int n, int nrhs, int batch, int threads, int reps)
{
    const size_t an = static_cast<size_t>(n) * n;
    const size_t bn = static_cast<size_t>(n) * nrhs;
    std::vector<float> a(an * batch), x(bn * batch);
    std::vector<int> piv(static_cast<size_t>(n) * batch), info(batch);
    std::vector<double> times;
    for (int r = 0; r < reps + 1; ++r) {
        std::memcpy(a.data(), A, a.size() * sizeof(float));
        std::memcpy(x.data(), B, x.size() * sizeof(float));
        const auto t0 = std::chrono::steady_clock::now();
        ParallelFor(batch, threads, [&](int b) {
            float* Ab = a.data() + b * an;
            int* pb = piv.data() + static_cast<size_t>(b) * n;
            info[b] = GetrfOne(Ab, n, n, pb);
            std::vector<int> perm;
            BuildPerm(pb, n, perm);
            std::vector<float> v(n);
            for (int c = 0; c < nrhs; ++c) {
                float* col = x.data() + b * bn + static_cast<size_t>(c) * n;
                for (int i = 0; i < n; ++i) v[i] = col[perm[i]];
                SolveN(Ab, n, n, v.data());
                for (int i = 0; i < n; ++i) col[i] = v[i];
            }
        });
        const auto t1 = std::chrono::steady_clock::now();
        if (r > 0) {
            times.push_back(std::chrono::duration<double>(t1 - t0).count());
        }
    }
    std::sort(times.begin(), times.end());
    return times.empty() ? 0.0 : times[times.size() / 2];
}

// Wall time of `reps` runs of getrf + getrs (trans N) over a batch of systems
// of mixed orders: system i is the n[i] x n[i] column-major matrix at
// A + offA[i] with nrhs[i] right-hand sides at B + offB[i] (leading dimension
// n[i]).  Systems are handed to the threads largest first.  Returns seconds
// per run (median); inputs are left untouched.
double orc_time_getrf_getrs_mixed(const float* A, const long long* offA, const float* B, const long long* offB,
                                  const int* n, const int* nrhs, int count, int threads, int reps)
{
    std::vector<int> order(static_cast<size_t>(count));
    for (int i = 0; i < count; ++i) order[static_cast<size_t>(i)] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return n[a] > n[b]; });
    size_t atotal = 0, btotal = 0;
    std::vector<size_t> wa(static_cast<size_t>(count)), wb(static_cast<size_t>(count));
    for (int i = 0; i < count; ++i) {
        wa[static_cast<size_t>(i)] = atotal;
        wb[static_cast<size_t>(i)] = btotal;
        atotal += static_cast<size_t>(n[i]) * n[i];
        btotal += static_cast<size_t>(n[i]) * nrhs[i];
    }
    std::vector<float> a(atotal), x(btotal);
    std::vector<double> times;
    for (int r = 0; r < reps + 1; ++r) {
        for (int i = 0; i < count; ++i) {
            std::memcpy(a.data() + wa[static_cast<size_t>(i)], A + offA[i],
                        sizeof(float) * static_cast<size_t>(n[i]) * n[i]);
            std::memcpy(x.data() + wb[static_cast<size_t>(i)], B + offB[i],
                        sizeof(float) * static_cast<size_t>(n[i]) * nrhs[i]);
        }
        const auto t0 = std::chrono::steady_clock::now();
        ParallelFor(count, threads, [&](int q) {
            const int i = order[static_cast<size_t>(q)];
            const int ni = n[i];
            float* Ab = a.data() + wa[static_cast<size_t>(i)];
            std::vector<int> piv(static_cast<size_t>(ni));
            GetrfOne(Ab, ni, ni, piv.data());
            std::vector<int> perm;
            BuildPerm(piv.data(), ni, perm);
            std::vector<float> v(static_cast<size_t>(ni));
            for (int c = 0; c < nrhs[i]; ++c) {
                float* col = x.data() + wb[static_cast<size_t>(i)] + static_cast<size_t>(c) * ni;
                for (int k = 0; k < ni; ++k) v[k] = col[perm[k]];
                SolveN(Ab, ni, ni, v.data());
                for (int k = 0; k < ni; ++k) col[k] = v[k];
            }
        });
        const auto t1 = std::chrono::steady_clock::now();
        if (r > 0) {
            times.push_back(std::chrono::duration<double>(t1 - t0).count());
        }
    }
    std::sort(times.begin(), times.end());
    return times.empty() ? 0.0 : times[times.size() / 2];
}

int orc_hardware_threads() { return static_cast<int>(std::thread::hardware_concurrency()); }

// ---------------------------------------------------------------------------
// Triangular solve with multiple right-hand sides (FP32), the golden of the
// reference's aclblasStrsm tests: test/trsm/strsm/strsm_golden.h:19-36 calls
// cblas_strsm (netlib reference BLAS, column-major).  This restates the loop
// order of netlib strsm.f for every side / uplo / trans / diag case: left side
// by columns of B with division by the diagonal, right side by columns of B
// with multiplication by the diagonal's reciprocal, alpha applied where
// strsm.f applies it, and every multiply-subtract rounded separately.
// side: 0 left, 1 right; upper: 1 upper; trans: 0 N, 1 T (C = T); unit: 1 unit.
// ---------------------------------------------------------------------------
void orc_trsm(int side, int upper, int trans, int unit, int m, int n, float alpha, const float* A, int lda, float* B,
              int ldb)
{
    auto a = [&](int i, int j) -> float { return A[i + static_cast<size_t>(j) * lda]; };
    auto b = [&](int i, int j) -> float& { return B[i + static_cast<size_t>(j) * ldb]; };
    if (m == 0 || n == 0) {
        return;
    }
    if (alpha == 0.0f) {
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < m; ++i) b(i, j) = 0.0f;
        return;
    }
    if (side == 0) {
        if (!trans) {
            for (int j = 0; j < n; ++j) {
                if (alpha != 1.0f)
                    for (int i = 0; i < m; ++i) b(i, j) = alpha * b(i, j);
                if (upper) {
                    for (int k = m - 1; k >= 0; --k) {
                        if (b(k, j) != 0.0f) {
                            if (!unit) b(k, j) = b(k, j) / a(k, k);
                            const float bk = b(k, j);
                            for (int i = 0; i < k; ++i) b(i, j) = b(i, j) - bk * a(i, k);
                        }
                    }
                } else {
                    for (int k = 0; k < m; ++k) {
                        if (b(k, j) != 0.0f) {
                            if (!unit) b(k, j) = b(k, j) / a(k, k);
                            const float bk = b(k, j);
                            for (int i = k + 1; i < m; ++i) b(i, j) = b(i, j) - bk * a(i, k);
                        }
                    }
                }
            }
        } else {
            for (int j = 0; j < n; ++j) {
                if (upper) {
                    for (int i = 0; i < m; ++i) {
                        float t = alpha * b(i, j);
                        for (int k = 0; k < i; ++k) t = t - a(k, i) * b(k, j);
                        if (!unit) t = t / a(i, i);
                        b(i, j) = t;
                    }
                } else {
                    for (int i = m - 1; i >= 0; --i) {
                        float t = alpha * b(i, j);
                        for (int k = i + 1; k < m; ++k) t = t - a(k, i) * b(k, j);
                        if (!unit) t = t / a(i, i);
                        b(i, j) = t;
                    }
                }
            }
        }
        return;
    }
    if (!trans) {
        if (upper) {
            for (int j = 0; j < n; ++j) {
                if (alpha != 1.0f)
                    for (int i = 0; i < m; ++i) b(i, j) = alpha * b(i, j);
                for (int k = 0; k < j; ++k) {
                    if (a(k, j) != 0.0f) {
                        const float akj = a(k, j);
                        for (int i = 0; i < m; ++i) b(i, j) = b(i, j) - akj * b(i, k);
                    }
                }
                if (!unit) {
                    const float t = 1.0f / a(j, j);
                    for (int i = 0; i < m; ++i) b(i, j) = t * b(i, j);
                }
            }
        } else {
            for (int j = n - 1; j >= 0; --j) {
                if (alpha != 1.0f)
                    for (int i = 0; i < m; ++i) b(i, j) = alpha * b(i, j);
                for (int k = j + 1; k < n; ++k) {
                    if (a(k, j) != 0.0f) {
                        const float akj = a(k, j);
                        for (int i = 0; i < m; ++i) b(i, j) = b(i, j) - akj * b(i, k);
                    }
                }
                if (!unit) {
                    const float t = 1.0f / a(j, j);
                    for (int i = 0; i < m; ++i) b(i, j) = t * b(i, j);
                }
            }
        }
    } else {
        if (upper) {
            for (int k = n - 1; k >= 0; --k) {
                if (!unit) {
                    const float t = 1.0f / a(k, k);
                    for (int i = 0; i < m; ++i) b(i, k) = t * b(i, k);
                }
                for (int j = 0; j < k; ++j) {
                    if (a(j, k) != 0.0f) {
                        const float t = a(j, k);
                        for (int i = 0; i < m; ++i) b(i, j) = b(i, j) - t * b(i, k);
                    }
                }
                if (alpha != 1.0f)
                    for (int i = 0; i < m; ++i) b(i, k) = alpha * b(i, k);
            }
        } else {
            for (int k = 0; k < n; ++k) {
                if (!unit) {
                    const float t = 1.0f / a(k, k);
                    for (int i = 0; i < m; ++i) b(i, k) = t * b(i, k);
                }
                for (int j = k + 1; j < n; ++j) {
                    if (a(j, k) != 0.0f) {
                        const float t = a(j, k);
                        for (int i = 0; i < m; ++i) b(i, j) = b(i, j) - t * b(i, k);
                    }
                }
                if (alpha != 1.0f)
                    for (int i = 0; i < m; ++i) b(i, k) = alpha * b(i, k);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// GEMM
// ---------------------------------------------------------------------------

// C = alpha*op(A)*op(B) + beta*C with FP64 products and sums, one rounding to
// the C type at the end.  Element types are aclDataType values (0 FP32, 1 FP16,
// 27 BF16, 35 E5M2, 36 E4M3FN).  transA/transB: 0 = N, 1 = T.  The reference's
// golden is cblas_sgemm on the (de-quantised) inputs followed by quantisation
// to Ctype (reference test/sgemm_ex/sgemm_ex_golden.h:121,
// test/gemm/gemm_ex/gemm_ex_golden.h); FP64 accumulation is the exact-product
// limit of that computation.  beta == 0 does not read C (BLAS convention).
// scaleAB multiplies the accumulated product (per-tensor FP8 dequantisation).
void orc_gemm(int transA, int transB, int m, int n, int k, double alpha, const void* A, int Atype, int lda,
              const void* B, int Btype, int ldb, double beta, void* C, int Ctype, int ldc, double scaleAB,
              int threads)
{
    ParallelFor(n, threads, [&](int j) {
        for (int i = 0; i < m; ++i) {
            double acc = 0.0;
            for (int l = 0; l < k; ++l) {
                const size_t ia = transA ? (static_cast<size_t>(i) * lda + l) : (i + static_cast<size_t>(l) * lda);
                const size_t ib = transB ? (j + static_cast<size_t>(l) * ldb) : (l + static_cast<size_t>(j) * ldb);
                acc += static_cast<double>(LoadAs(A, ia, Atype)) * static_cast<double>(LoadAs(B, ib, Btype));
            }
            const size_t ic = i + static_cast<size_t>(j) * ldc;
            double r = alpha * acc * scaleAB;
            if (beta != 0.0) {
                r += beta * static_cast<double>(LoadAs(C, ic, Ctype));
            }
            StoreAs(C, ic, Ctype, static_cast<float>(r));
        }
    });
}

// BLASLt matmul golden (reference test/blasLtMatmul/blasLtMatmul_golden.h):
// row-major D = alpha op(A) op(B) + beta C with every operand decoded to FP32
// first and the products summed in FP64, one rounding to FP32, then (BF16 D)
// the golden's truncation to BF16 (:318-323: the high 16 bits).  MX decoding
// as the golden's: E4M3 with exponent field 15 read as 0 (:25-39; test data
// stays below it), E2M1 per the OCP table (:43-63), E8M0 scale 0 -> 0,
// 255 -> inf, else 2^(s - 127) (:65-71), one scale per 32 K elements looked
// up in the two layouts of :77-91 (A transposed and B not transposed use the
// K-major pairs layout).  Atype: 0 FP32, 36 MXFP8, 40 MXFP4 (both operands
// of one kind).  lda / ldb: elements for FP32 and FP8, bytes per row for FP4.
// C may be null (beta ignored then).  D: m x ldd floats.
static float LtE4m3(uint8_t v)
{
    const int sign = (v >> 7) & 1, e = (v >> 3) & 0xF, mant = v & 7;
    float r;
    if ((e == 0 && mant == 0) || e == 15) {
        r = 0.0f;
    } else if (e == 0) {
        r = static_cast<float>(mant) / 8.0f * std::ldexp(1.0f, -6);
    } else {
        r = (1.0f + static_cast<float>(mant) / 8.0f) * std::ldexp(1.0f, e - 7);
    }
    return sign ? -r : r;
}
static float LtE2m1(uint8_t nib)
{
    const int sign = (nib >> 3) & 1, e = (nib >> 1) & 3, mant = nib & 1;
    const float r = e == 0 ? 0.5f * static_cast<float>(mant) : (1.0f + 0.5f * mant) * std::ldexp(1.0f, e - 1);
    return sign ? -r : r;
}
static float LtE8m0(uint8_t s)
{
    if (s == 0) return 0.0f;
    if (s == 255) return INFINITY;
    return std::ldexp(1.0f, static_cast<int>(s) - 127);
}
static float LtScale(const uint8_t* scale, bool kMajorPairs, int minorDim, int k, int minor, int kk)
{
    const int block = kk / 64, sub = (kk % 64) / 32;
    if (kMajorPairs) {
        return LtE8m0(scale[static_cast<size_t>(block) * minorDim * 2 + static_cast<size_t>(minor) * 2 + sub]);
    }
    const int stride = ((k + 63) / 64) * 2;
    return LtE8m0(scale[static_cast<size_t>(minor) * stride + static_cast<size_t>(block) * 2 + sub]);
}
// op(A)[i][kk] (idx = i, K index kk) of a stored operand: element (r, c) of the stored matrix with the
// (minor, K) roles given by kContig (stored row = minor index).
static float LtElem(int type, const void* X, long long ld, bool kContig, int minor, int kk, const uint8_t* scale,
                    bool kMajorPairs, int minorDim, int k)
{
    const long long r = kContig ? minor : kk, c = kContig ? kk : minor;
    if (type == 0) {
        return static_cast<const float*>(X)[r * ld + c];
    }
    const float s = LtScale(scale, kMajorPairs, minorDim, k, minor, kk);
    const uint8_t* b = static_cast<const uint8_t*>(X);
    if (type == 36) {
        return LtE4m3(b[r * ld + c]) * s;
    }
    const uint8_t byte = b[r * ld + c / 2];
    return LtE2m1(static_cast<uint8_t>((c & 1) ? (byte >> 4) : (byte & 0xF))) * s;
}

void orc_lt_matmul(int type, int bf16Out, int transA, int transB, int m, int n, int k, float alpha, const void* A,
                   long long lda, const void* B, long long ldb, const uint8_t* scaleA, const uint8_t* scaleB,
                   float beta, const float* C, int ldc, float* D, int ldd, int threads)
{
    ParallelFor(m, threads, [&](int i) {
        for (int j = 0; j < n; ++j) {
            double sum = 0.0;
            for (int kk = 0; kk < k; ++kk) {
                // A: stored m x k (N, K-contiguous, scales [m][S]) or k x m (T, K-major pairs)
                const float a = LtElem(type, A, lda, transA == 0, i, kk, scaleA, transA != 0, m, k);
                // B: stored k x n (N, K-major pairs) or n x k (T, K-contiguous, scales [n][S])
                const float b = LtElem(type, B, ldb, transB != 0, j, kk, scaleB, transB == 0, n, k);
                sum += static_cast<double>(a) * static_cast<double>(b);
            }
            double cv = 0.0;
            if (C != nullptr && beta != 0.0f) {
                cv = static_cast<double>(C[static_cast<size_t>(i) * ldc + j]) * static_cast<double>(beta);
            }
            float v = static_cast<float>(static_cast<double>(alpha) * sum + cv);
            if (bf16Out) {
                uint32_t u;
                std::memcpy(&u, &v, 4);
                u &= 0xFFFF0000u;
                std::memcpy(&v, &u, 4);
            }
            D[static_cast<size_t>(i) * ldd + j] = v;
        }
    });
}

// Element conversions (count elements), for building typed test inputs.
void orc_f32_to_bf16(const float* in, uint16_t* out, size_t count)
{
    for (size_t i = 0; i < count; ++i) out[i] = F32ToBf16(in[i]);
}
void orc_bf16_to_f32(const uint16_t* in, float* out, size_t count)
{
    for (size_t i = 0; i < count; ++i) out[i] = Bf16ToF32(in[i]);
}
void orc_f32_to_f16(const float* in, uint16_t* out, size_t count)
{
    for (size_t i = 0; i < count; ++i) out[i] = F32ToF16(in[i]);
}
void orc_f16_to_f32(const uint16_t* in, float* out, size_t count)
{
    for (size_t i = 0; i < count; ++i) out[i] = F16ToF32(in[i]);
}
void orc_f32_to_fp8(const float* in, uint8_t* out, size_t count, int e4m3)
{
    for (size_t i = 0; i < count; ++i) out[i] = F32ToFp8(in[i], e4m3 != 0);
}
void orc_fp8_to_f32(const uint8_t* in, float* out, size_t count, int e4m3)
{
    for (size_t i = 0; i < count; ++i) out[i] = Fp8ToF32(in[i], e4m3 != 0);
}

// ---------------------------------------------------------------------------
// FP64 residual oracle
// ---------------------------------------------------------------------------

// Normwise backward error of X for A X = B (per system, max over its columns):
//   max_r ||b_r - A x_r||_inf / (||A||_inf ||x_r||_inf + ||b_r||_inf)
// evaluated in FP64 from the FP32 data.  A, B, X column-major, contiguous per
// system with leading dimension n.
void orc_backward_error(const float* A, const float* B, const float* X, int n, int nrhs, int batch, double* out,
                        int threads)
{
    ParallelFor(batch, threads, [&](int b) {
        const float* Ab = A + static_cast<size_t>(b) * n * n;
        double anorm = 0.0;
        for (int i = 0; i < n; ++i) {
            double s = 0.0;
            for (int j = 0; j < n; ++j) s += std::fabs(static_cast<double>(Ab[i + static_cast<size_t>(j) * n]));
            anorm = std::max(anorm, s);
        }
        double worst = 0.0;
        for (int r = 0; r < nrhs; ++r) {
            const float* x = X + static_cast<size_t>(b) * n * nrhs + static_cast<size_t>(r) * n;
            const float* bb = B + static_cast<size_t>(b) * n * nrhs + static_cast<size_t>(r) * n;
            double rn = 0.0, xn = 0.0, bn = 0.0;
            for (int i = 0; i < n; ++i) {
                double s = static_cast<double>(bb[i]);
                for (int j = 0; j < n; ++j) s -= static_cast<double>(Ab[i + static_cast<size_t>(j) * n]) * x[j];
                rn = std::max(rn, std::fabs(s));
                xn = std::max(xn, std::fabs(static_cast<double>(x[i])));
                bn = std::max(bn, std::fabs(static_cast<double>(bb[i])));
            }
            const double den = anorm * xn + bn;
            worst = std::max(worst, den > 0.0 ? rn / den : rn);
        }
        out[b] = worst;
    });
}

}  // extern "C"
