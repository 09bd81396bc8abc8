// MFMA GEMM kernel template for gfx950 (wave64, v_mfma_*_16x16x*).
//
// C[m x n] = alpha * op(A) * op(B) + beta * C, column-major, FP32 accumulate,
// fused epilogue (no temporary, no second pass over C).
//
// Tiling.  A workgroup of 256 threads (2 x 2 waves) computes a BM x BN tile of
// C; each wave owns a (BM/2) x (BN/2) sub-tile made of 16 x 16 MFMA tiles.  The
// K loop advances by one "k-block" of 128 bytes per row (32 FP32, 64 BF16/FP16
// or 128 FP8 values).  Both operands are staged in LDS in k-contiguous rows
// (row = one m index of op(A) or one n index of op(B)), 128 bytes per row,
// with the eight 16-byte chunks of row r XOR-swizzled by (r & 7) so that the
// MFMA operand reads (16 rows x 32 bytes per quarter-wave) are free of bank
// conflicts.  Two LDS stages: the next k-block is fetched from global memory
// into registers while the MFMAs consume the current one.
//
// MFMA operand mapping (16x16 family): lane l supplies row (l % 16) of the A
// fragment and column (l % 16) of the B fragment, and reads the 32-byte quarter
// (l / 16) of its row in the k-block: 8 FP32 values (8 x 16x16x4 MFMAs), 16
// BF16/FP16 values (2 x 16x16x32) or 32 FP8 values (1 x 16x16x128).  A and B
// lanes consume the same k indices, so every k is summed exactly once.  The
// accumulator of lane l holds rows 4*(l/16)..+3 of column l%16: four
// consecutive elements of one column of C, stored with one vector write.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace aclblas {
namespace gemm {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef int i32x8 __attribute__((ext_vector_type(8)));

// Element kinds (also the LDS byte layouts).
enum Kind { kF32 = 0, kBF16 = 1, kF16 = 2, kE4M3 = 3, kE5M2 = 4 };

template <int K>
struct KindInfo;
template <>
struct KindInfo<kF32> {
    static constexpr int bytes = 4;
};
template <>
struct KindInfo<kBF16> {
    static constexpr int bytes = 2;
};
template <>
struct KindInfo<kF16> {
    static constexpr int bytes = 2;
};
template <>
struct KindInfo<kE4M3> {
    static constexpr int bytes = 1;
};
template <>
struct KindInfo<kE5M2> {
    static constexpr int bytes = 1;
};

constexpr int kRowBytes = 128;   // one k-block per LDS row
constexpr int kThreads = 256;

struct GemmArgs {
    const void* A;
    const void* B;
    void* C;
    int m, n, k;
    int lda, ldb, ldc;
    int transA, transB;     // 0 = N, 1 = T
    float alpha, beta;
    const float* scaleA;    // optional device scalars multiplying op(A), op(B)
    const float* scaleB;
    int vecA, vecB;         // 1 when 16-byte vector loads are legal
    int tilesM, tilesN;
};

__device__ __forceinline__ uint32_t SwizzledOffset(int row, int byteInRow)
{
    const int chunk = (byteInRow >> 4) ^ (row & 7);
    return static_cast<uint32_t>(row * kRowBytes + (chunk << 4) + (byteInRow & 15));
}

// Fetches of one operand tile (ROWS rows of op(X), one k-block) from global
// memory into registers, then into LDS.  Row r of the tile is op(X) row/col
// index r0 + r; the k-block starts at k0.  kcontig: element (r, kk) sits at
// X[kk + r*ld] (k contiguous); otherwise at X[r + kk*ld] (rows contiguous).
template <int ROWS, int EB>
struct TileLoader {
    static constexpr int kElems = kRowBytes / EB;                 // k per block
    static constexpr int kChunks = ROWS * kRowBytes / 16;         // 16-byte chunks per tile
    static constexpr int kPerThread = kChunks / kThreads;
    static_assert(kChunks % kThreads == 0, "tile must split evenly");
    uint4 v[kPerThread];

    __device__ __forceinline__ static uint32_t LoadElem(const uint8_t* p)
    {
        if (EB == 4) return *reinterpret_cast<const uint32_t*>(p);
        if (EB == 2) return *reinterpret_cast<const uint16_t*>(p);
        return *p;
    }

    __device__ __forceinline__ void Fetch(const uint8_t* X, int ld, bool kcontig, bool vec, int r0, int rlim,
                                          int k0, int klim)
    {
        const int t = threadIdx.x;
#pragma unroll
        for (int q = 0; q < kPerThread; ++q) {
            const int id = t + kThreads * q;
            uint32_t w[4] = {0u, 0u, 0u, 0u};
            if (kcontig) {
                // chunk = 16 bytes along k of one row
                const int r = id / (kRowBytes / 16);
                const int kc = (id % (kRowBytes / 16)) * (16 / EB);
                const int gr = r0 + r;
                const int gk = k0 + kc;
                const uint8_t* p = X + (static_cast<size_t>(gr) * ld + gk) * EB;
                if (vec && gr < rlim && gk + 16 / EB <= klim) {
                    v[q] = *reinterpret_cast<const uint4*>(p);
                    continue;
                }
#pragma unroll
                for (int e = 0; e < 16 / EB; ++e) {
                    uint32_t x = 0;
                    if (gr < rlim && gk + e < klim) {
                        x = LoadElem(p + e * EB);
                    }
                    w[(e * EB) / 4] |= x << (((e * EB) % 4) * 8);
                }
            } else {
                // chunk = 16 bytes along rows at one k
                constexpr int kRowChunks = ROWS * EB / 16;
                const int kk = id / kRowChunks;
                const int rc = (id % kRowChunks) * (16 / EB);
                const int gr = r0 + rc;
                const int gk = k0 + kk;
                const uint8_t* p = X + (static_cast<size_t>(gk) * ld + gr) * EB;
                if (vec && gk < klim && gr + 16 / EB <= rlim) {
                    v[q] = *reinterpret_cast<const uint4*>(p);
                    continue;
                }
#pragma unroll
                for (int e = 0; e < 16 / EB; ++e) {
                    uint32_t x = 0;
                    if (gk < klim && gr + e < rlim) {
                        x = LoadElem(p + e * EB);
                    }
                    w[(e * EB) / 4] |= x << (((e * EB) % 4) * 8);
                }
            }
            v[q] = make_uint4(w[0], w[1], w[2], w[3]);
        }
    }

    __device__ __forceinline__ void Store(uint8_t* lds, bool kcontig) const
    {
        const int t = threadIdx.x;
#pragma unroll
        for (int q = 0; q < kPerThread; ++q) {
            const int id = t + kThreads * q;
            if (kcontig) {
                const int r = id / (kRowBytes / 16);
                const int b = (id % (kRowBytes / 16)) * 16;
                *reinterpret_cast<uint4*>(lds + SwizzledOffset(r, b)) = v[q];
            } else {
                constexpr int kRowChunks = ROWS * EB / 16;
                const int kk = id / kRowChunks;
                const int rc = (id % kRowChunks) * (16 / EB);
                const uint32_t w[4] = {v[q].x, v[q].y, v[q].z, v[q].w};
#pragma unroll
                for (int e = 0; e < 16 / EB; ++e) {
                    const uint32_t x = (w[(e * EB) / 4] >> (((e * EB) % 4) * 8));
                    uint8_t* dst = lds + SwizzledOffset(rc + e, kk * EB);
                    if (EB == 4) {
                        *reinterpret_cast<uint32_t*>(dst) = x;
                    } else if (EB == 2) {
                        *reinterpret_cast<uint16_t*>(dst) = static_cast<uint16_t>(x);
                    } else {
                        *dst = static_cast<uint8_t>(x);
                    }
                }
            }
        }
    }
};

// One MFMA step over a 32-byte quarter-row fragment of A and B.
template <int KA, int KB>
struct Mfma;

template <>
struct Mfma<kF32, kF32> {
    __device__ __forceinline__ static void Run(const uint4 (&a)[2], const uint4 (&b)[2], f32x4& acc)
    {
        const float* fa = reinterpret_cast<const float*>(a);
        const float* fb = reinterpret_cast<const float*>(b);
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[s], fb[s], acc, 0, 0, 0);
        }
    }
};

template <>
struct Mfma<kBF16, kBF16> {
    __device__ __forceinline__ static void Run(const uint4 (&a)[2], const uint4 (&b)[2], f32x4& acc)
    {
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a[s]),
                                                          __builtin_bit_cast(bf16x8, b[s]), acc, 0, 0, 0);
        }
    }
};

template <>
struct Mfma<kF16, kF16> {
    __device__ __forceinline__ static void Run(const uint4 (&a)[2], const uint4 (&b)[2], f32x4& acc)
    {
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a[s]),
                                                         __builtin_bit_cast(f16x8, b[s]), acc, 0, 0, 0);
        }
    }
};

// FP8/BF8 operand format codes of v_mfma_f32_16x16x128_f8f6f4 (cbsz / blgp):
// 0 = E4M3 (fp8), 1 = E5M2 (bf8).
template <int KA, int KB>
struct Fp8Mfma {
    static constexpr int fa = KA == kE5M2 ? 1 : 0;
    static constexpr int fb = KB == kE5M2 ? 1 : 0;
    __device__ __forceinline__ static void Run(const uint4 (&a)[2], const uint4 (&b)[2], f32x4& acc)
    {
        struct P {
            uint4 x, y;
        };
        const i32x8 va = __builtin_bit_cast(i32x8, P{a[0], a[1]});
        const i32x8 vb = __builtin_bit_cast(i32x8, P{b[0], b[1]});
        acc = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(va, vb, acc, fa, fb, 0, 0, 0, 0);
    }
};
template <>
struct Mfma<kE4M3, kE4M3> : Fp8Mfma<kE4M3, kE4M3> {};
template <>
struct Mfma<kE4M3, kE5M2> : Fp8Mfma<kE4M3, kE5M2> {};
template <>
struct Mfma<kE5M2, kE4M3> : Fp8Mfma<kE5M2, kE4M3> {};
template <>
struct Mfma<kE5M2, kE5M2> : Fp8Mfma<kE5M2, kE5M2> {};

// Output element conversions (round to nearest even).
__device__ __forceinline__ uint16_t ToBf16(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7FFFFFFFu) > 0x7F800000u) {
        return static_cast<uint16_t>((u >> 16) | 0x40u);
    }
    u += 0x7FFFu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}
__device__ __forceinline__ float FromBf16(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }
__device__ __forceinline__ uint16_t ToF16(float f)
{
    const _Float16 h = static_cast<_Float16>(f);
    return __builtin_bit_cast(uint16_t, h);
}
__device__ __forceinline__ float FromF16(uint16_t h) { return static_cast<float>(__builtin_bit_cast(_Float16, h)); }

enum OutKind { kOutF32 = 0, kOutBF16 = 1, kOutF16 = 2 };

template <int OK>
__device__ __forceinline__ float LoadC(const void* C, size_t idx)
{
    if (OK == kOutF32) return static_cast<const float*>(C)[idx];
    if (OK == kOutBF16) return FromBf16(static_cast<const uint16_t*>(C)[idx]);
    return FromF16(static_cast<const uint16_t*>(C)[idx]);
}

template <int OK>
__device__ __forceinline__ void StoreC(void* C, size_t idx, float v)
{
    if (OK == kOutF32) {
        static_cast<float*>(C)[idx] = v;
    } else if (OK == kOutBF16) {
        static_cast<uint16_t*>(C)[idx] = ToBf16(v);
    } else {
        static_cast<uint16_t*>(C)[idx] = ToF16(v);
    }
}

// Block-id to tile mapping.  Workgroups are dispatched round-robin over the 8
// XCDs (each with its own L2); consecutive ids of one XCD are given
// neighbouring tiles, grouped by GROUP rows of tiles so that a group shares its
// A row-panels and B column-panels in that XCD's L2.
__device__ __forceinline__ void TileOf(int bid, int nblocks, int tilesM, int tilesN, int& tm, int& tn)
{
    constexpr int kXcd = 8;
    int id = bid;
    if (nblocks % kXcd == 0) {
        id = (bid % kXcd) * (nblocks / kXcd) + bid / kXcd;
    }
    constexpr int kGroup = 8;
    const int perGroup = kGroup * tilesN;
    const int g = id / perGroup;
    const int firstM = g * kGroup;
    const int rowsInGroup = min(kGroup, tilesM - firstM);
    const int inGroup = id - g * perGroup;
    tm = firstM + inGroup % rowsInGroup;
    tn = inGroup / rowsInGroup;
}

// FP8 non-finite results become NaN (reference test/gemm/gemm_ex/gemm_ex_golden.h,
// QuantizeGoldenOutput: the reference's FP8 path yields NaN for non-finite).
template <int KA, int KB, int OK, int BM, int BN>
__global__ __launch_bounds__(kThreads) void GemmKernel(GemmArgs p)
{
    constexpr int EA = KindInfo<KA>::bytes;
    constexpr int EB = KindInfo<KB>::bytes;
    static_assert(EA == EB, "operands share the k-block size");
    constexpr int kK = kRowBytes / EA;   // k per block
    constexpr int WM = BM / 2;           // wave tile
    constexpr int WN = BN / 2;
    constexpr int TM = WM / 16;          // MFMA tiles per wave
    constexpr int TN = WN / 16;
    constexpr bool kFp8 = (KA == kE4M3 || KA == kE5M2);

    __shared__ __attribute__((aligned(16))) uint8_t sA[2][BM * kRowBytes];
    __shared__ __attribute__((aligned(16))) uint8_t sB[2][BN * kRowBytes];

    int tm, tn;
    TileOf(blockIdx.x, gridDim.x, p.tilesM, p.tilesN, tm, tn);
    const int m0 = tm * BM;
    const int n0 = tn * BN;
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wave = t >> 6;
    const int wm = (wave & 1) * WM;
    const int wn = (wave >> 1) * WN;

    const uint8_t* A = static_cast<const uint8_t*>(p.A);
    const uint8_t* B = static_cast<const uint8_t*>(p.B);
    // op(A) row r = C row; k-contiguous when A is transposed.
    const bool aK = p.transA != 0;
    // op(B) column c = C column; k-contiguous when B is not transposed.
    const bool bK = p.transB == 0;

    TileLoader<BM, EA> la;
    TileLoader<BN, EB> lb;

    f32x4 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    const int nk = (p.k + kK - 1) / kK;
    if (nk > 0) {
        la.Fetch(A, p.lda, aK, p.vecA, m0, p.m, 0, p.k);
        lb.Fetch(B, p.ldb, bK, p.vecB, n0, p.n, 0, p.k);
        la.Store(sA[0], aK);
        lb.Store(sB[0], bK);
    }
    __syncthreads();

    const int fr = lane & 15;         // fragment row / column
    const int fq = (lane >> 4) * 32;  // byte offset of this lane's quarter

    for (int kb = 0; kb < nk; ++kb) {
        const int cur = kb & 1;
        const bool more = kb + 1 < nk;
        if (more) {
            la.Fetch(A, p.lda, aK, p.vecA, m0, p.m, (kb + 1) * kK, p.k);
            lb.Fetch(B, p.ldb, bK, p.vecB, n0, p.n, (kb + 1) * kK, p.k);
        }
        uint4 fb[TN][2];
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            const int row = wn + j * 16 + fr;
            fb[j][0] = *reinterpret_cast<const uint4*>(sB[cur] + SwizzledOffset(row, fq));
            fb[j][1] = *reinterpret_cast<const uint4*>(sB[cur] + SwizzledOffset(row, fq + 16));
        }
#pragma unroll
        for (int i = 0; i < TM; ++i) {
            const int row = wm + i * 16 + fr;
            uint4 fa[2];
            fa[0] = *reinterpret_cast<const uint4*>(sA[cur] + SwizzledOffset(row, fq));
            fa[1] = *reinterpret_cast<const uint4*>(sA[cur] + SwizzledOffset(row, fq + 16));
#pragma unroll
            for (int j = 0; j < TN; ++j) {
                Mfma<KA, KB>::Run(fa, fb[j], acc[i][j]);
            }
        }
        if (more) {
            la.Store(sA[cur ^ 1], aK);
            lb.Store(sB[cur ^ 1], bK);
        }
        __syncthreads();
    }

    // ---- fused epilogue: C = alpha * scale * acc + beta * C ----
    float alpha = p.alpha;
    if (p.scaleA != nullptr) alpha *= *p.scaleA;
    if (p.scaleB != nullptr) alpha *= *p.scaleB;
    const float beta = p.beta;
    const int rq = (lane >> 4) * 4;
#pragma unroll
    for (int i = 0; i < TM; ++i) {
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            const int col = n0 + wn + j * 16 + fr;
            const int row = m0 + wm + i * 16 + rq;
            if (col >= p.n) {
                continue;
            }
            const size_t base = static_cast<size_t>(col) * p.ldc + row;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                if (row + e < p.m) {
                    float v = alpha * acc[i][j][e];
                    if (beta != 0.0f) {
                        v = v + beta * LoadC<OK>(p.C, base + e);
                    }
                    if (kFp8 && !__builtin_isfinite(v)) {
                        v = __builtin_nanf("");
                    }
                    StoreC<OK>(p.C, base + e, v);
                }
            }
        }
    }
}

}  // namespace gemm
}  // namespace aclblas
