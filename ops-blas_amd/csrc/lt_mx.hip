// MX block-scaled GEMM behind aclblasLtMatmul: MXFP8 (E4M3FN) x MXFP8 and
// MXFP4 (E2M1) x MXFP4, row-major, D = alpha op(A) op(B) + beta C with FP32 C
// and FP32 or BF16 D.
//
// Reference: blasLt/matmul/mxfp8/arch35/matmul_mxfp8_kernel.cpp,
// blasLt/matmul/mxfp4/arch35/matmul_mxfp4_kernel.cpp and the alpha/beta
// epilogue blasLt/matmul/epilogue/arch35/epilogue_alpha_beta_kernel.cpp; the
// scale layouts are those of the reference's golden
// (test/blasLtMatmul/blasLtMatmul_golden.h:77-91).
//
// MI355X design: the MX format is native to gfx950's
// v_mfma_scale_f32_16x16x128_f8f6f4: each lane feeds 32 K elements of one row
// of its operand (row l % 16) and one E8M0 scale byte per lane, the scale of
// one 32-element K block of that row (the slot layout is at Fragment; the
// result lane holds rows 4 (l / 16) + e, column l % 16).  So one instruction
// consumes four 32-element blocks of A and of B with their scales and no
// dequantisation pass exists.
// The whole scaled product is summed in FP32 by the instruction, alpha / beta
// and the rounding to D's type are fused in the epilogue, and D is written
// once.
//
// One workgroup of four waves per 64 x 64 tile of D, each wave 32 x 32 (two
// by two MFMA tiles); K advances 128 elements per step through LDS, where
// every operand is held K-contiguous per row (a row of A or a column of B),
// 128 bytes (FP8) or 64 bytes (FP4, two elements per byte, low nibble first)
// per step, plus four scale bytes per row.  Operands stored K-contiguous are
// copied 16 bytes at a time; the others are transposed on the way into LDS.
// Rows or columns past m / n, and K blocks past k (k is a multiple of 32)
// read as zeros with scale 1.  A scale byte of 0 makes its block contribute
// 0, the reference golden's decoding of E8M0 0 (the instruction would scale
// by 2^-127).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "drivers.h"
#include "internal.h"

namespace aclblas {
namespace {

constexpr int kTile = 64;
constexpr int kThreads = 256;
constexpr int kStepK = 128;     // K elements per step (one MFMA per fragment pair)
constexpr int kPad = 16;        // bytes of padding per LDS row

typedef int i32x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

struct MxArgs {
    const uint8_t* A;
    const uint8_t* B;
    const uint8_t* scaleA;
    const uint8_t* scaleB;
    const float* C;
    void* D;
    long long pitchA;   // bytes between rows of A as stored
    long long pitchB;
    int m, n, k;
    int ldc, ldd;
    int transA, transB; // 0 = N
    int vecA, vecB;     // 16-byte copies possible (aligned base and pitch)
    float alpha, beta;
};

// Scale byte of row `row` (of A: m index; of B: n index) and K block b.
// kMajorPairs: the [b / 2][rows][2] layout, else [rows][2 ceil(k / 64)].
__device__ __forceinline__ uint8_t ScaleAt(const uint8_t* s, bool kMajorPairs, int rows, int k, int row, int b)
{
    if (kMajorPairs) {
        return s[static_cast<size_t>(b >> 1) * 2 * rows + 2 * row + (b & 1)];
    }
    const int stride = ((k + 63) / 64) * 2;
    return s[static_cast<size_t>(row) * stride + b];
}

// Loads one operand's 64-row step into LDS rows of `rowBytes` data bytes.
// kContig: element (row, kk) of the operand is stored at row * pitch + kk
// (bytes kk / 2 for FP4), else at kk * pitch + row.
template <bool FP4>
__device__ __forceinline__ void LoadOperand(uint8_t* lds, uint8_t* sc, const uint8_t* X, long long pitch,
                                            bool kContig, bool vec, int row0, int rows, int k0, int k,
                                            const uint8_t* scale, bool kMajorPairs)
{
    constexpr int kRowBytes = FP4 ? kStepK / 2 : kStepK;
    constexpr int kLd = kRowBytes + kPad;
    const int t = threadIdx.x;
    const int kBytes0 = FP4 ? k0 / 2 : k0;       // byte offset of this step along K (K-contiguous storage)
    const int kBytesEnd = FP4 ? k / 2 : k;
    if (kContig) {
        // 16-byte chunks: kRowBytes / 16 per row
        constexpr int kChunks = kRowBytes / 16;
        for (int e = t; e < kTile * kChunks; e += kThreads) {
            const int r = e / kChunks;
            const int c = e % kChunks;
            const int row = row0 + r;
            const int kb = kBytes0 + 16 * c;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (row < rows && kb < kBytesEnd) {
                const uint8_t* src = X + static_cast<size_t>(row) * pitch + kb;
                if (vec) {
                    v = *reinterpret_cast<const uint4*>(src);
                } else {
                    uint32_t w[4];
                    for (int i = 0; i < 4; ++i) {
                        w[i] = static_cast<uint32_t>(src[4 * i]) | (static_cast<uint32_t>(src[4 * i + 1]) << 8) |
                               (static_cast<uint32_t>(src[4 * i + 2]) << 16) |
                               (static_cast<uint32_t>(src[4 * i + 3]) << 24);
                    }
                    v = make_uint4(w[0], w[1], w[2], w[3]);
                }
            }
            *reinterpret_cast<uint4*>(lds + r * kLd + 16 * c) = v;
        }
    } else if constexpr (!FP4) {
        // element (row, kk) at kk * pitch + row: one byte per (row, kk)
        for (int e = t; e < kTile * kStepK; e += kThreads) {
            const int r = e % kTile;
            const int kk = e / kTile;
            const int row = row0 + r;
            const int kg = k0 + kk;
            lds[r * kLd + kk] = (row < rows && kg < k) ? X[static_cast<size_t>(kg) * pitch + row] : 0;
        }
    } else {
        // FP4: nibble (row, kk) at byte kk * pitch + row / 2, low nibble for even rows; an LDS byte packs
        // K elements 2q (low) and 2q + 1 (high) of one row
        for (int e = t; e < kTile * (kStepK / 2); e += kThreads) {
            const int r = e % kTile;
            const int q = e / kTile;
            const int row = row0 + r;
            const int kg = k0 + 2 * q;
            uint8_t out = 0;
            if (row < rows && kg < k) {
                const uint8_t lo = X[static_cast<size_t>(kg) * pitch + row / 2];
                const uint8_t hi = X[static_cast<size_t>(kg + 1) * pitch + row / 2];
                const int sh = (row & 1) * 4;
                out = static_cast<uint8_t>(((lo >> sh) & 0xF) | (((hi >> sh) & 0xF) << 4));
            }
            lds[r * kLd + q] = out;
        }
    }
    // scales: four K blocks of 32 per step
    for (int e = t; e < kTile * 4; e += kThreads) {
        const int r = e / 4;
        const int b = (k0 / 32) + (e % 4);
        const int row = row0 + r;
        sc[e] = (row < rows && 32 * b < k) ? ScaleAt(scale, kMajorPairs, rows, k, row, b) : static_cast<uint8_t>(127);
    }
}

__device__ __forceinline__ uint16_t Bf16Rne(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7FFFFFFFu) > 0x7F800000u) {
        return static_cast<uint16_t>((u >> 16) | 0x40u);
    }
    u += 0x7FFFu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}

// One lane's operand of v_mfma_scale_f32_16x16x128_f8f6f4 for the step: `row`
// points at the LDS row (K-contiguous bytes of the step), `sc` at its four
// block scales, g = lane / 16.  The instruction's K slots, established on the
// GPU, are:
//   FP8: the lane's bytes 0-15 are slots 16 g .. 16 g + 15 and bytes 16-31 are
//        slots 64 + 16 g .. 64 + 16 g + 15; the scale the lane supplies covers
//        slots 32 g .. 32 g + 31 (held by lane groups 2 (g % 2) .. + 1);
//   FP4: the lane's 32 elements are slots 32 g .. 32 g + 31 and its scale
//        covers them.
// Slot s holds K element k0 + s, so each supplied scale is that of the step's
// K block g.  A zero scale byte zeroes the data of its block (scale 1 then).
template <bool FP4>
__device__ __forceinline__ i32x8 Fragment(const uint8_t* row, const uint8_t* sc, int g, int& scale)
{
    const int s = sc[g];
    scale = s == 0 ? 127 : s;
    if constexpr (FP4) {
        const uint4 v = s == 0 ? make_uint4(0, 0, 0, 0) : *reinterpret_cast<const uint4*>(row + 16 * g);
        return i32x8{static_cast<int>(v.x), static_cast<int>(v.y), static_cast<int>(v.z), static_cast<int>(v.w),
                     0, 0, 0, 0};
    } else {
        // bytes 0-15 belong to block g / 2, bytes 16-31 to block 2 + g / 2
        const uint4 lo = sc[g >> 1] == 0 ? make_uint4(0, 0, 0, 0) : *reinterpret_cast<const uint4*>(row + 16 * g);
        const uint4 hi =
            sc[2 + (g >> 1)] == 0 ? make_uint4(0, 0, 0, 0) : *reinterpret_cast<const uint4*>(row + 64 + 16 * g);
        return i32x8{static_cast<int>(lo.x), static_cast<int>(lo.y), static_cast<int>(lo.z), static_cast<int>(lo.w),
                     static_cast<int>(hi.x), static_cast<int>(hi.y), static_cast<int>(hi.z), static_cast<int>(hi.w)};
    }
}

template <bool FP4, bool BF16OUT>
__global__ __launch_bounds__(kThreads) void LtMxKernel(MxArgs p)
{
    constexpr int kRowBytes = FP4 ? kStepK / 2 : kStepK;
    constexpr int kLd = kRowBytes + kPad;
    constexpr int kFmt = FP4 ? 4 : 0;  // operand format code: 0 E4M3, 4 E2M1
    __shared__ __attribute__((aligned(16))) uint8_t sA[kTile * kLd];
    __shared__ __attribute__((aligned(16))) uint8_t sB[kTile * kLd];
    __shared__ uint8_t sSA[kTile * 4];
    __shared__ uint8_t sSB[kTile * 4];

    const int lane = static_cast<int>(threadIdx.x) & 63;
    const int wave = static_cast<int>(threadIdx.x) >> 6;
    const int wm = (wave >> 1) * 32;
    const int wn = (wave & 1) * 32;
    const int r16 = lane & 15;
    const int g = lane >> 4;
    const int m0 = static_cast<int>(blockIdx.y) * kTile;
    const int n0 = static_cast<int>(blockIdx.x) * kTile;

    f32x4 acc[2][2];
    for (int i = 0; i < 2; ++i) {
        for (int j = 0; j < 2; ++j) {
            acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        }
    }
    // A: row index m, K-contiguous when transA == N; B: row index n, K-contiguous when transB != N.
    // Scales: A [m][2 ceil(k/64)] when transA == N, else K-major pairs; B the other way round.
    for (int k0 = 0; k0 < p.k; k0 += kStepK) {
        LoadOperand<FP4>(sA, sSA, p.A, p.pitchA, p.transA == 0, p.vecA != 0, m0, p.m, k0, p.k, p.scaleA,
                         p.transA != 0);
        LoadOperand<FP4>(sB, sSB, p.B, p.pitchB, p.transB != 0, p.vecB != 0, n0, p.n, k0, p.k, p.scaleB,
                         p.transB == 0);
        __syncthreads();
        i32x8 a[2], b[2];
        int sa[2], sb[2];
        for (int i = 0; i < 2; ++i) {
            a[i] = Fragment<FP4>(sA + (wm + 16 * i + r16) * kLd, sSA + (wm + 16 * i + r16) * 4, g, sa[i]);
        }
        for (int j = 0; j < 2; ++j) {
            b[j] = Fragment<FP4>(sB + (wn + 16 * j + r16) * kLd, sSB + (wn + 16 * j + r16) * 4, g, sb[j]);
        }
        for (int i = 0; i < 2; ++i) {
            for (int j = 0; j < 2; ++j) {
                acc[i][j] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a[i], b[j], acc[i][j], kFmt, kFmt, 0,
                                                                             sa[i], 0, sb[j]);
            }
        }
        __syncthreads();
    }
    // epilogue: lane holds rows 4 g + e, column r16 of each 16 x 16 tile
    for (int i = 0; i < 2; ++i) {
        for (int j = 0; j < 2; ++j) {
            const int col = n0 + wn + 16 * j + r16;
            for (int e = 0; e < 4; ++e) {
                const int row = m0 + wm + 16 * i + 4 * g + e;
                if (row < p.m && col < p.n) {
                    float v = p.alpha * acc[i][j][e];
                    if (p.beta != 0.0f) {
                        v = fmaf(p.beta, p.C[static_cast<size_t>(row) * p.ldc + col], v);
                    }
                    if constexpr (BF16OUT) {
                        static_cast<uint16_t*>(p.D)[static_cast<size_t>(row) * p.ldd + col] = Bf16Rne(v);
                    } else {
                        static_cast<float*>(p.D)[static_cast<size_t>(row) * p.ldd + col] = v;
                    }
                }
            }
        }
    }
}

}  // namespace

aclblasStatus_t LtMxLaunch(hipStream_t st, bool fp4, bool transA, bool transB, int m, int n, int k, float alpha,
                           const void* A, long long lda, const void* B, long long ldb, const void* scaleA,
                           const void* scaleB, float beta, const float* C, int ldc, void* D, int ldd, bool bf16Out)
{
    MxArgs p{};
    p.A = static_cast<const uint8_t*>(A);
    p.B = static_cast<const uint8_t*>(B);
    p.scaleA = static_cast<const uint8_t*>(scaleA);
    p.scaleB = static_cast<const uint8_t*>(scaleB);
    p.C = C;
    p.D = D;
    p.pitchA = lda;  // bytes per stored row (FP4: two elements per byte)
    p.pitchB = ldb;
    p.m = m;
    p.n = n;
    p.k = k;
    p.ldc = ldc;
    p.ldd = ldd;
    p.transA = transA ? 1 : 0;
    p.transB = transB ? 1 : 0;
    p.vecA = (reinterpret_cast<uintptr_t>(A) % 16 == 0 && p.pitchA % 16 == 0) ? 1 : 0;
    p.vecB = (reinterpret_cast<uintptr_t>(B) % 16 == 0 && p.pitchB % 16 == 0) ? 1 : 0;
    p.alpha = alpha;
    p.beta = beta;
    const dim3 grid((n + kTile - 1) / kTile, (m + kTile - 1) / kTile);
    if (fp4) {
        if (bf16Out) {
            hipLaunchKernelGGL((LtMxKernel<true, true>), grid, dim3(kThreads), 0, st, p);
        } else {
            hipLaunchKernelGGL((LtMxKernel<true, false>), grid, dim3(kThreads), 0, st, p);
        }
    } else {
        if (bf16Out) {
            hipLaunchKernelGGL((LtMxKernel<false, true>), grid, dim3(kThreads), 0, st, p);
        } else {
            hipLaunchKernelGGL((LtMxKernel<false, false>), grid, dim3(kThreads), 0, st, p);
        }
    }
    return LaunchStatus("aclblasLtMatmul");
}

}  // namespace aclblas
