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
// BF16/FP16 values (2 x 16x16x32) or 32 FP8 values (widened to BF16 in
// registers, 4 x 16x16x32).  A and B lanes consume the same k indices, so every
// k is summed exactly once.  The accumulator of lane l holds rows
// 4*(l/16)..+3 of column l%16: four consecutive elements of one column of C,
// stored with one vector write.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_common.h"

namespace aclblas {
namespace gemm {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef int i32x8 __attribute__((ext_vector_type(8)));

// Element kinds (also the LDS byte layouts).  kE4M3/kE5M2: FP8 operands with
// FP32-grade accumulation (widened to BF16); kE4M3S/kE5M2S: FP8 operands on the
// native FP8 matrix instruction (the per-tensor-scaled entry).
enum Kind { kF32 = 0, kBF16 = 1, kF16 = 2, kE4M3 = 3, kE5M2 = 4, kE4M3S = 5, kE5M2S = 6 };

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
template <>
struct KindInfo<kE4M3S> {
    static constexpr int bytes = 1;
};
template <>
struct KindInfo<kE5M2S> {
    static constexpr int bytes = 1;
};

constexpr int kRowBytes = 128;   // one k-block per LDS row
constexpr int kThreads = 256;     // default workgroup: 2 x 2 waves
constexpr int kWaveSize = 64;

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
    int vecC;               // 1 when four consecutive C elements can move as one vector
    int tilesM, tilesN;
    int group;              // tile rows per L2 group (TileOf)
    // batched form: device arrays of per-problem pointers (A/B/C unused), or null
    const void* const* Aarr;
    const void* const* Barr;
    void* const* Carr;
};

// The XOR key mixes row bits 0-2 and 3-5 so that both access patterns are
// conflict-free: eight consecutive rows at one chunk (MFMA fragment reads) and
// one chunk of rows eight apart (transposed tile stores).
__device__ __forceinline__ uint32_t SwizzledOffset(int row, int byteInRow)
{
    const int chunk = (byteInRow >> 4) ^ ((row ^ (row >> 3)) & 7);
    return static_cast<uint32_t>(row * kRowBytes + (chunk << 4) + (byteInRow & 15));
}

// Fetches of one operand tile (ROWS rows of op(X), one k-block) from global
// memory into registers, then into LDS.  Row r of the tile is op(X) row/col
// index r0 + r; the k-block starts at k0.  kcontig: element (r, kk) sits at
// X[kk + r*ld] (k contiguous): each thread moves 16-byte chunks along k.
// Otherwise element (r, kk) sits at X[r + kk*ld] (rows contiguous): each thread
// loads KB 16-byte chunks of E rows at KB consecutive k, transposes the E x KB
// block in registers and stores E row segments of KB elements (2- and 4-byte
// types; FP8 falls back to element stores).
template <int ROWS, int EB, int NT = kThreads>
struct TileLoader {
    static constexpr int kElems = kRowBytes / EB;                 // k per block
    static constexpr int kChunks = ROWS * kRowBytes / 16;         // 16-byte chunks per tile
    static constexpr int kPerThread = kChunks / NT;               // = KB
    static constexpr int E = 16 / EB;                             // elements per chunk
    static constexpr int kMChunks = ROWS / E;                     // row chunks (transposed path)
    static_assert(kChunks % NT == 0, "tile must split evenly");
    static_assert(kMChunks * (kElems / kPerThread) == NT, "transposed blocks cover the tile");
    static constexpr int kWords = kPerThread * EB / 4;  // 32-bit words per transposed row segment
    uint4 v[kPerThread];
    // addressing of the unchecked path (interior tiles), fixed for one tile
    const uint8_t* base;                 // this thread's chunk 0 at k-block 0
    size_t qstride;                      // bytes between this thread's consecutive chunks
    size_t kstep;                        // bytes per k-block

    __device__ __forceinline__ void Setup(const uint8_t* X, int ld, bool kcontig, int r0)
    {
        const int t = threadIdx.x;
        if (kcontig) {
            const int r = t / (kRowBytes / 16);
            const int kc = t % (kRowBytes / 16);
            base = X + (static_cast<size_t>(r0 + r) * ld + kc * E) * EB;
            qstride = static_cast<size_t>(NT / (kRowBytes / 16)) * ld * EB;
            kstep = static_cast<size_t>(kElems) * EB;
        } else {
            const int mc = t % kMChunks;
            const int kg = t / kMChunks;
            base = X + (static_cast<size_t>(kg * kPerThread) * ld + r0 + mc * E) * EB;
            qstride = static_cast<size_t>(ld) * EB;
            kstep = static_cast<size_t>(kElems) * ld * EB;
        }
    }

    // All chunks of k-block kb in range and 16-byte aligned.
    __device__ __forceinline__ void FetchFull(int kb) { FetchFullInto(kb, v); }
    __device__ __forceinline__ void FetchFullInto(int kb, uint4 (&dst)[kPerThread]) const
    {
        const uint8_t* p = base + static_cast<size_t>(kb) * kstep;
#pragma unroll
        for (int q = 0; q < kPerThread; ++q) {
            dst[q] = dev::LoadG<uint4>(p + q * qstride);
        }
    }

    // Stores the fetched chunks (EB 2 and 4 use the register transpose).  The
    // LDS offsets are recomputed at each store: offsets kept in arrays per
    // layout end up in private memory once the compiler merges the two
    // layouts' stores, and a private-memory load waits for every global load
    // in flight.
    __device__ __forceinline__ void StoreFast(uint8_t* lds, bool kcontig) const
    {
        if constexpr (EB == 4 || EB == 2) {
            StoreFastFrom(lds, kcontig, v);
        } else if (kcontig) {
            StoreFastFrom(lds, kcontig, v);
        } else {
            Store(lds, kcontig);
        }
    }
    // The same from registers src (EB 2 and 4, or a k-contiguous operand).
    __device__ __forceinline__ void StoreFastFrom(uint8_t* lds, bool kcontig, const uint4 (&src)[kPerThread]) const
    {
        const int t = threadIdx.x;
        if (kcontig) {
#pragma unroll
            for (int q = 0; q < kPerThread; ++q) {
                const int id = t + NT * q;
                *reinterpret_cast<uint4*>(lds + SwizzledOffset(id / (kRowBytes / 16), (id % (kRowBytes / 16)) * 16)) =
                    src[q];
            }
            return;
        }
        const int mc = t % kMChunks;
        const int kg = t / kMChunks;
        const int byte0 = kg * kPerThread * EB;
        if constexpr (EB == 4 || EB == 2) {
            const uint32_t* w = reinterpret_cast<const uint32_t*>(src);
#pragma unroll
            for (int i = 0; i < E; ++i) {
                uint32_t seg[kWords];
#pragma unroll
                for (int o = 0; o < kWords; ++o) {
                    if constexpr (EB == 4) {
                        seg[o] = w[o * 4 + i];
                    } else {
                        const uint32_t lo = w[(2 * o) * 4 + i / 2];
                        const uint32_t hi = w[(2 * o + 1) * 4 + i / 2];
                        seg[o] = __builtin_amdgcn_perm(hi, lo, (i & 1) ? 0x07060302u : 0x05040100u);
                    }
                }
                const int row = mc * E + i;
                if constexpr (kWords >= 4) {
#pragma unroll
                    for (int o = 0; o < kWords; o += 4) {
                        *reinterpret_cast<uint4*>(lds + SwizzledOffset(row, byte0 + o * 4)) =
                            make_uint4(seg[o], seg[o + 1], seg[o + 2], seg[o + 3]);
                    }
                } else if constexpr (kWords == 2) {
                    *reinterpret_cast<uint2*>(lds + SwizzledOffset(row, byte0)) = make_uint2(seg[0], seg[1]);
                } else {
                    *reinterpret_cast<uint32_t*>(lds + SwizzledOffset(row, byte0)) = seg[0];
                }
            }
        }
    }

    __device__ __forceinline__ static uint32_t LoadElem(const uint8_t* p)
    {
        if (EB == 4) return dev::LoadG<uint32_t>(p);
        if (EB == 2) return dev::LoadG<uint16_t>(p);
        return dev::LoadG<uint8_t>(p);
    }

    __device__ __forceinline__ static uint4 LoadChunk(const uint8_t* p, bool full, int valid)
    {
        if (full) {
            return dev::LoadG<uint4>(p);
        }
        uint32_t w[4] = {0u, 0u, 0u, 0u};
#pragma unroll
        for (int e = 0; e < E; ++e) {
            if (e < valid) {
                w[(e * EB) / 4] |= LoadElem(p + e * EB) << (((e * EB) % 4) * 8);
            }
        }
        return make_uint4(w[0], w[1], w[2], w[3]);
    }

    __device__ __forceinline__ void Fetch(const uint8_t* X, int ld, bool kcontig, bool vec, int r0, int rlim,
                                          int k0, int klim)
    {
        FetchInto(X, ld, kcontig, vec, r0, rlim, k0, klim, v);
    }
    __device__ __forceinline__ void FetchInto(const uint8_t* X, int ld, bool kcontig, bool vec, int r0, int rlim,
                                              int k0, int klim, uint4 (&v)[kPerThread]) const
    {
        const int t = threadIdx.x;
        if (kcontig) {
#pragma unroll
            for (int q = 0; q < kPerThread; ++q) {
                const int id = t + NT * q;
                const int r = id / (kRowBytes / 16);
                const int gr = r0 + r;
                const int gk = k0 + (id % (kRowBytes / 16)) * E;
                const uint8_t* p = X + (static_cast<size_t>(gr) * ld + gk) * EB;
                const int valid = gr < rlim ? klim - gk : 0;
                v[q] = LoadChunk(p, vec && valid >= E, valid);
            }
        } else {
            const int mc = t % kMChunks;
            const int kg = t / kMChunks;
            const int gr = r0 + mc * E;
#pragma unroll
            for (int q = 0; q < kPerThread; ++q) {
                const int gk = k0 + kg * kPerThread + q;
                const uint8_t* p = X + (static_cast<size_t>(gk) * ld + gr) * EB;
                const int valid = gk < klim ? rlim - gr : 0;
                v[q] = LoadChunk(p, vec && valid >= E, valid);
            }
        }
    }

    __device__ __forceinline__ void Store(uint8_t* lds, bool kcontig) const
    {
        const int t = threadIdx.x;
        if (kcontig) {
#pragma unroll
            for (int q = 0; q < kPerThread; ++q) {
                const int id = t + NT * q;
                const int r = id / (kRowBytes / 16);
                const int b = (id % (kRowBytes / 16)) * 16;
                *reinterpret_cast<uint4*>(lds + SwizzledOffset(r, b)) = v[q];
            }
            return;
        }
        const int mc = t % kMChunks;
        const int kg = t / kMChunks;
        const int byte0 = kg * kPerThread * EB;
        if constexpr (EB == 4 || EB == 2) {
            constexpr int kWords = kPerThread * EB / 4;  // 32-bit words per row segment
            const uint32_t* w = reinterpret_cast<const uint32_t*>(v);
#pragma unroll
            for (int i = 0; i < E; ++i) {
                uint32_t seg[kWords];
#pragma unroll
                for (int o = 0; o < kWords; ++o) {
                    if constexpr (EB == 4) {
                        seg[o] = w[o * 4 + i];
                    } else {
                        // element i of chunks 2o and 2o+1 (16-bit halves)
                        const uint32_t lo = w[(2 * o) * 4 + i / 2];
                        const uint32_t hi = w[(2 * o + 1) * 4 + i / 2];
                        seg[o] = __builtin_amdgcn_perm(hi, lo, (i & 1) ? 0x07060302u : 0x05040100u);
                    }
                }
                const int row = mc * E + i;
                if constexpr (kWords >= 4) {
#pragma unroll
                    for (int o = 0; o < kWords; o += 4) {
                        *reinterpret_cast<uint4*>(lds + SwizzledOffset(row, byte0 + o * 4)) =
                            make_uint4(seg[o], seg[o + 1], seg[o + 2], seg[o + 3]);
                    }
                } else if constexpr (kWords == 2) {
                    *reinterpret_cast<uint2*>(lds + SwizzledOffset(row, byte0)) = make_uint2(seg[0], seg[1]);
                } else {
                    *reinterpret_cast<uint32_t*>(lds + SwizzledOffset(row, byte0)) = seg[0];
                }
            }
        } else {
            // FP8: element stores
#pragma unroll
            for (int q = 0; q < kPerThread; ++q) {
                const uint32_t w[4] = {v[q].x, v[q].y, v[q].z, v[q].w};
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    lds[SwizzledOffset(mc * E + e, byte0 + q)] = static_cast<uint8_t>(w[e / 4] >> ((e % 4) * 8));
                }
            }
        }
    }
};

// One MFMA step over a 32-byte quarter-row fragment of A and B.
template <int KA, int KB>
struct Mfma;

// Run: one k-block quarter (32 bytes per lane); RunHalf: its first or second
// 16 bytes (kSplit: the instruction sequence splits at 16 bytes, so a k-block
// can be consumed half by half with the same per-accumulator order).
template <>
struct Mfma<kF32, kF32> {
    static constexpr bool kSplit = true;
    __device__ __forceinline__ static void RunHalf(const uint4& a, const uint4& b, f32x4& acc)
    {
        const float fa[4] = {__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z), __uint_as_float(a.w)};
        const float fb[4] = {__uint_as_float(b.x), __uint_as_float(b.y), __uint_as_float(b.z), __uint_as_float(b.w)};
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[s], fb[s], acc, 0, 0, 0);
        }
    }
    __device__ __forceinline__ static void Run(const uint4 (&a)[2], const uint4 (&b)[2], f32x4& acc)
    {
        RunHalf(a[0], b[0], acc);
        RunHalf(a[1], b[1], acc);
    }
};

template <>
struct Mfma<kBF16, kBF16> {
    static constexpr bool kSplit = true;
    __device__ __forceinline__ static void RunHalf(const uint4& a, const uint4& b, f32x4& acc)
    {
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b),
                                                      acc, 0, 0, 0);
    }
    __device__ __forceinline__ static void Run(const uint4 (&a)[2], const uint4 (&b)[2], f32x4& acc)
    {
        RunHalf(a[0], b[0], acc);
        RunHalf(a[1], b[1], acc);
    }
};

template <>
struct Mfma<kF16, kF16> {
    static constexpr bool kSplit = true;
    __device__ __forceinline__ static void RunHalf(const uint4& a, const uint4& b, f32x4& acc)
    {
        acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b), acc,
                                                     0, 0, 0);
    }
    __device__ __forceinline__ static void Run(const uint4 (&a)[2], const uint4 (&b)[2], f32x4& acc)
    {
        RunHalf(a[0], b[0], acc);
        RunHalf(a[1], b[1], acc);
    }
};

// FP8 operands with FP32-grade accumulation.  v_mfma_f32_16x16x128_f8f6f4
// sums its 128 products in a narrow internal window (about 2^-14 of the
// largest product is kept), which misses the FP32-accumulate bar of GemmEx.
// Instead every FP8 value is widened to BF16 in registers (exact: E4M3 and
// E5M2 are subsets of BF16) and fed to v_mfma_f32_16x16x32_bf16, whose products
// are exact and whose sums are FP32.  The LDS tiles stay 1 byte per element.
template <int K>
__device__ __forceinline__ uint32_t Fp8PairToBf16(uint32_t w, bool hi)
{
    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
    bf16x2 r;
    if (K == kE5M2) {
        r = hi ? __builtin_amdgcn_cvt_scalef32_pk_bf16_bf8(w, 1.0f, true)
               : __builtin_amdgcn_cvt_scalef32_pk_bf16_bf8(w, 1.0f, false);
    } else {
        r = hi ? __builtin_amdgcn_cvt_scalef32_pk_bf16_fp8(w, 1.0f, true)
               : __builtin_amdgcn_cvt_scalef32_pk_bf16_fp8(w, 1.0f, false);
    }
    return __builtin_bit_cast(uint32_t, r);
}

// 8 FP8 values (two 32-bit words) -> 8 BF16 values.
template <int K>
__device__ __forceinline__ bf16x8 Fp8x8ToBf16(uint32_t w0, uint32_t w1)
{
    const uint4 r = make_uint4(Fp8PairToBf16<K>(w0, false), Fp8PairToBf16<K>(w0, true),
                               Fp8PairToBf16<K>(w1, false), Fp8PairToBf16<K>(w1, true));
    return __builtin_bit_cast(bf16x8, r);
}

template <int KA, int KB>
struct Fp8ViaBf16Mfma {
    static constexpr bool kSplit = true;
    __device__ __forceinline__ static void RunHalf(const uint4& a, const uint4& b, f32x4& acc)
    {
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Fp8x8ToBf16<KA>(a.x, a.y), Fp8x8ToBf16<KB>(b.x, b.y), acc, 0,
                                                      0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Fp8x8ToBf16<KA>(a.z, a.w), Fp8x8ToBf16<KB>(b.z, b.w), acc, 0,
                                                      0, 0);
    }
    __device__ __forceinline__ static void Run(const uint4 (&a)[2], const uint4 (&b)[2], f32x4& acc)
    {
        RunHalf(a[0], b[0], acc);
        RunHalf(a[1], b[1], acc);
    }
};
template <>
struct Mfma<kE4M3, kE4M3> : Fp8ViaBf16Mfma<kE4M3, kE4M3> {};
template <>
struct Mfma<kE4M3, kE5M2> : Fp8ViaBf16Mfma<kE4M3, kE5M2> {};
template <>
struct Mfma<kE5M2, kE4M3> : Fp8ViaBf16Mfma<kE5M2, kE4M3> {};
template <>
struct Mfma<kE5M2, kE5M2> : Fp8ViaBf16Mfma<kE5M2, kE5M2> {};

// Native FP8: one v_mfma_f32_16x16x128_f8f6f4 per 32-byte fragment.  Operand
// format codes (cbsz / blgp): 0 = E4M3, 1 = E5M2.  Its 128 products are summed
// in a narrower window than FP32 (relative error about 2^-14 of the largest
// product, measured), which the scaled entry states as its accuracy.
template <int KA, int KB>
struct Fp8NativeMfma {
    static constexpr bool kSplit = false;
    __device__ __forceinline__ static void RunHalf(const uint4&, const uint4&, f32x4&) {}
    static constexpr int fa = KA == kE5M2S ? 1 : 0;
    static constexpr int fb = KB == kE5M2S ? 1 : 0;
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
struct Mfma<kE4M3S, kE4M3S> : Fp8NativeMfma<kE4M3S, kE4M3S> {};
template <>
struct Mfma<kE4M3S, kE5M2S> : Fp8NativeMfma<kE4M3S, kE5M2S> {};
template <>
struct Mfma<kE5M2S, kE4M3S> : Fp8NativeMfma<kE5M2S, kE4M3S> {};
template <>
struct Mfma<kE5M2S, kE5M2S> : Fp8NativeMfma<kE5M2S, kE5M2S> {};

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
// FP16 rounding to nearest with ties away from zero: the nearest-even result,
// moved one step away from zero when f lies exactly halfway and nearest-even
// went toward zero (both differences below are exact in FP32).
__device__ __forceinline__ uint16_t ToF16Away(float f)
{
    const uint16_t h = ToF16(f);
    const float back = FromF16(h);
    const float mag = __builtin_fabsf(f);
    const float hm = __builtin_fabsf(back);
    if (mag > hm && (h & 0x7FFFu) < 0x7C00u) {
        const uint16_t up = static_cast<uint16_t>(h + 1u);  // same sign, next magnitude
        if (mag - hm == __builtin_fabsf(FromF16(up)) - mag) {
            return up;
        }
    }
    return h;
}

// kOutF16A: FP16 with ties rounded away from zero (the grouped 16-bit/FP8 GEMM's
// epilogue rounding), no FP8 non-finite rule.
enum OutKind { kOutF32 = 0, kOutBF16 = 1, kOutF16 = 2, kOutF16A = 3 };

template <int OK>
__device__ __forceinline__ float LoadC(const void* C, size_t idx)
{
    if (OK == kOutF32) return dev::G(static_cast<const float*>(C))[idx];
    if (OK == kOutBF16) return FromBf16(dev::G(static_cast<const uint16_t*>(C))[idx]);
    return FromF16(dev::G(static_cast<const uint16_t*>(C))[idx]);
}

template <int OK>
__device__ __forceinline__ void StoreC(void* C, size_t idx, float v)
{
    if (OK == kOutF32) {
        dev::G(static_cast<float*>(C))[idx] = v;
    } else if (OK == kOutBF16) {
        dev::G(static_cast<uint16_t*>(C))[idx] = ToBf16(v);
    } else if (OK == kOutF16A) {
        dev::G(static_cast<uint16_t*>(C))[idx] = ToF16Away(v);
    } else {
        dev::G(static_cast<uint16_t*>(C))[idx] = ToF16(v);
    }
}

// Four consecutive elements of one column of C (rows row..row+3); `full`
// selects one vector access, otherwise the first `valid` elements move singly.
template <int OK>
__device__ __forceinline__ void LoadC4(const void* C, size_t idx, bool full, int valid, float (&c)[4])
{
    if (full) {
        if (OK == kOutF32) {
            const float4 x = dev::LoadG<float4>(static_cast<const float*>(C) + idx);
            c[0] = x.x; c[1] = x.y; c[2] = x.z; c[3] = x.w;
        } else {
            const uint2 x = dev::LoadG<uint2>(static_cast<const uint16_t*>(C) + idx);
            const uint16_t h[4] = {static_cast<uint16_t>(x.x), static_cast<uint16_t>(x.x >> 16),
                                   static_cast<uint16_t>(x.y), static_cast<uint16_t>(x.y >> 16)};
#pragma unroll
            for (int e = 0; e < 4; ++e) c[e] = OK == kOutBF16 ? FromBf16(h[e]) : FromF16(h[e]);
        }
        return;
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) c[e] = e < valid ? LoadC<OK>(C, idx + e) : 0.0f;
}

template <int OK>
__device__ __forceinline__ void StoreC4(void* C, size_t idx, bool full, int valid, const float (&v)[4])
{
    if (full) {
        if (OK == kOutF32) {
            dev::StoreG(static_cast<float*>(C) + idx, make_float4(v[0], v[1], v[2], v[3]));
        } else {
            uint16_t h[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                h[e] = OK == kOutBF16 ? ToBf16(v[e]) : (OK == kOutF16A ? ToF16Away(v[e]) : ToF16(v[e]));
            }
            dev::StoreG(static_cast<uint16_t*>(C) + idx,
                        make_uint2(h[0] | (static_cast<uint32_t>(h[1]) << 16),
                                   h[2] | (static_cast<uint32_t>(h[3]) << 16)));
        }
        return;
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        if (e < valid) StoreC<OK>(C, idx + e, v[e]);
    }
}

// Tile of one problem from its (XCD-ordered) block index.  Workgroups are
// dispatched round-robin over the 8 XCDs (each with its own L2); the kernel
// first renumbers blocks so that consecutive numbers run on one XCD, then
// walks tiles in groups of `group` tile-rows so that the blocks of one XCD
// share their A row-panels and B column-panels in that XCD's L2.
__device__ __forceinline__ void TileOf(int id, int tilesM, int tilesN, int group, int& tm, int& tn)
{
    const int perGroup = group * tilesN;
    const int g = id / perGroup;
    const int firstM = g * group;
    const int rowsInGroup = min(group, tilesM - firstM);
    const int inGroup = id - g * perGroup;
    tm = firstM + inGroup % rowsInGroup;
    tn = inGroup / rowsInGroup;
}

// One BM x BN tile of one problem: everything the tile computation needs.
struct TileJob {
    const uint8_t* A;
    const uint8_t* B;
    void* C;
    int m, n, k;
    int lda, ldb, ldc;
    bool aK, bK;            // operand tiles are k-contiguous in memory
    bool vecA, vecB, vecC;  // 16-byte operand loads / 4-element C moves are legal
    float alpha, beta;
};

// Fused epilogue of one wave's TM x TN MFMA tiles whose top-left output
// element is (row0, col0): C = alpha * acc + beta * C (C is not read when beta
// is 0); FP8_NAN turns non-finite results into NaN.
template <int OK, int TM, int TN, bool FP8_NAN>
__device__ __forceinline__ void StoreTileC(const TileJob& q, const f32x4 (&acc)[TM][TN], int row0, int col0, int lane)
{
    const float alpha = q.alpha;
    const float beta = q.beta;
    const int fr = lane & 15;
    const int rq = (lane >> 4) * 4;
#pragma unroll
    for (int i = 0; i < TM; ++i) {
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            const int col = col0 + j * 16 + fr;
            const int row = row0 + i * 16 + rq;
            if (col >= q.n || row >= q.m) {
                continue;
            }
            const size_t base = static_cast<size_t>(col) * q.ldc + row;
            float v[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                v[e] = alpha * acc[i][j][e];
            }
            const bool full = q.vecC && row + 4 <= q.m;
            if (beta != 0.0f) {
                float c[4];
                LoadC4<OK>(q.C, base, full, q.m - row, c);
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    v[e] = v[e] + beta * c[e];
                }
            }
            if (FP8_NAN) {
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    if (!__builtin_isfinite(v[e])) {
                        v[e] = __builtin_nanf("");
                    }
                }
            }
            StoreC4<OK>(q.C, base, full, q.m - row, v);
        }
    }
}

// C of a wave's TM x TN MFMA tiles read ahead of the k loop (beta != 0), in
// the accumulator layout: one round trip less after the last MFMA.
template <int OK, int TM, int TN>
__device__ __forceinline__ void PreloadTileC(const TileJob& q, int row0, int col0, int lane, float (&c)[TM][TN][4])
{
    const int fr = lane & 15;
    const int rq = (lane >> 4) * 4;
#pragma unroll
    for (int i = 0; i < TM; ++i) {
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            const int col = col0 + j * 16 + fr;
            const int row = row0 + i * 16 + rq;
            const bool in = col < q.n && row < q.m;
            // rows past m or columns past n read element (0, 0) and are never stored
            const size_t base = in ? static_cast<size_t>(col) * q.ldc + row : 0;
            LoadC4<OK>(q.C, base, in && q.vecC && row + 4 <= q.m, in ? q.m - row : 1, c[i][j]);
        }
    }
}

// StoreTileC with C already in registers (PreloadTileC).
template <int OK, int TM, int TN>
__device__ __forceinline__ void StoreTileCPre(const TileJob& q, const f32x4 (&acc)[TM][TN], const float (&c)[TM][TN][4],
                                              int row0, int col0, int lane)
{
    const float alpha = q.alpha;
    const float beta = q.beta;
    const int fr = lane & 15;
    const int rq = (lane >> 4) * 4;
#pragma unroll
    for (int i = 0; i < TM; ++i) {
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            const int col = col0 + j * 16 + fr;
            const int row = row0 + i * 16 + rq;
            if (col >= q.n || row >= q.m) {
                continue;
            }
            const size_t base = static_cast<size_t>(col) * q.ldc + row;
            float v[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                v[e] = alpha * acc[i][j][e] + beta * c[i][j][e];
            }
            StoreC4<OK>(q.C, base, q.vecC && row + 4 <= q.m, q.m - row, v);
        }
    }
}

// The MFMAs of one k-block held in LDS (a, b: the stage's A and B rows; offA,
// offB: a wave's swizzled fragment offsets).  Every tile shape runs the same
// instruction sequence per 16 x 16 accumulator, so results do not depend on
// the tile.
template <int KA, int KB, int TM, int TN>
__device__ __forceinline__ void KBlockMfma(const uint8_t* a, const uint8_t* b, const uint32_t (&offA)[4],
                                           const uint32_t (&offB)[4], f32x4 (&acc)[TM][TN])
{
    if constexpr (Mfma<KA, KB>::kSplit) {
        // half-quarter by half-quarter: half the fragment registers live
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            uint4 fb[TN];
#pragma unroll
            for (int j = 0; j < TN; ++j) {
                fb[j] = *reinterpret_cast<const uint4*>(b + ((offB[j % 4] ^ (16u * h)) + j * 16 * kRowBytes));
            }
#pragma unroll
            for (int i = 0; i < TM; ++i) {
                const uint4 fa = *reinterpret_cast<const uint4*>(a + ((offA[i % 4] ^ (16u * h)) +
                                                                      i * 16 * kRowBytes));
#pragma unroll
                for (int j = 0; j < TN; ++j) {
                    Mfma<KA, KB>::RunHalf(fa, fb[j], acc[i][j]);
                }
            }
        }
    } else {
        uint4 fb[TN][2];
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            fb[j][0] = *reinterpret_cast<const uint4*>(b + (offB[j % 4] + j * 16 * kRowBytes));
            fb[j][1] = *reinterpret_cast<const uint4*>(b + ((offB[j % 4] ^ 16u) + j * 16 * kRowBytes));
        }
#pragma unroll
        for (int i = 0; i < TM; ++i) {
            uint4 fa[2];
            fa[0] = *reinterpret_cast<const uint4*>(a + (offA[i % 4] + i * 16 * kRowBytes));
            fa[1] = *reinterpret_cast<const uint4*>(a + ((offA[i % 4] ^ 16u) + i * 16 * kRowBytes));
#pragma unroll
            for (int j = 0; j < TN; ++j) {
                Mfma<KA, KB>::Run(fa, fb[j], acc[i][j]);
            }
        }
    }
}

// Computes C[m0:m0+BM, n0:n0+BN] = alpha * op(A) op(B) + beta * C for one job
// with the LDS stages sA/sB.  FP8 non-finite results become NaN (reference
// test/gemm/gemm_ex/gemm_ex_golden.h, QuantizeGoldenOutput: the reference's FP8
// path yields NaN for non-finite).  Ends with the LDS free for the next tile.
template <int KA, int KB, int OK, int BM, int BN, int WGM = 2, int WGN = 2>
__device__ __forceinline__ void GemmTile(const TileJob& q, int m0, int n0, uint8_t (*sA)[BM * kRowBytes],
                                         uint8_t (*sB)[BN * kRowBytes])
{
    constexpr int NT = kWaveSize * WGM * WGN;
    constexpr int EA = KindInfo<KA>::bytes;
    constexpr int EB = KindInfo<KB>::bytes;
    static_assert(EA == EB, "operands share the k-block size");
    constexpr int kK = kRowBytes / EA;   // k per block
    constexpr int WM = BM / WGM;         // wave tile
    constexpr int WN = BN / WGN;
    constexpr int TM = WM / 16;          // MFMA tiles per wave
    constexpr int TN = WN / 16;
    constexpr bool kFp8 = (KA == kE4M3 || KA == kE5M2);

    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wave = t >> 6;
    const int wm = (wave % WGM) * WM;
    const int wn = (wave / WGM) * WN;

    TileLoader<BM, EA, NT> la;
    TileLoader<BN, EB, NT> lb;
    la.Setup(q.A, q.lda, q.aK, m0);
    lb.Setup(q.B, q.ldb, q.bK, n0);
    // interior tiles take the unchecked path for every full k-block
    const int fullK = q.k / kK;
    const bool interiorA = q.vecA && m0 + BM <= q.m;
    const bool interiorB = q.vecB && n0 + BN <= q.n;

    f32x4 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    const int nk = (q.k + kK - 1) / kK;
    auto fetch = [&](int kb) {
        if (interiorA && kb < fullK) {
            la.FetchFull(kb);
        } else {
            la.Fetch(q.A, q.lda, q.aK, q.vecA, m0, q.m, kb * kK, q.k);
        }
        if (interiorB && kb < fullK) {
            lb.FetchFull(kb);
        } else {
            lb.Fetch(q.B, q.ldb, q.bK, q.vecB, n0, q.n, kb * kK, q.k);
        }
    };
    if (nk > 0) {
        fetch(0);
        la.StoreFast(sA[0], q.aK);
        lb.StoreFast(sB[0], q.bK);
    }
    __syncthreads();

    const int fr = lane & 15;         // fragment row / column
    const int fq = (lane >> 4) * 32;  // byte offset of this lane's quarter
    // Fragment rows r = w + 16 i + fr: the swizzle key (r ^ r >> 3) & 7 depends
    // on i only through i mod 4, so four offsets per operand and half-row cover
    // every fragment (plus 16 rows * 128 bytes per i).
    // The second 16 bytes of a quarter sit at offset ^ 16 (fq / 16 is even).
    uint32_t offA[4], offB[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        offA[i] = SwizzledOffset(wm + i * 16 + fr, fq) - i * 16 * kRowBytes;
        offB[i] = SwizzledOffset(wn + i * 16 + fr, fq) - i * 16 * kRowBytes;
    }

    for (int kb = 0; kb < nk; ++kb) {
        const int cur = kb & 1;
        const bool more = kb + 1 < nk;
        if (more) {
            fetch(kb + 1);
        }
        KBlockMfma<KA, KB, TM, TN>(sA[cur], sB[cur], offA, offB, acc);
        if (more) {
            la.StoreFast(sA[cur ^ 1], q.aK);
            lb.StoreFast(sB[cur ^ 1], q.bK);
        }
        __syncthreads();
    }

    StoreTileC<OK, TM, TN, kFp8 && OK != kOutF16A>(q, acc, m0 + wm, n0 + wn, lane);
}

// GemmTile with DEPTH k-blocks in flight: register slots hold the global loads
// of k-blocks kb+1 .. kb+DEPTH while the MFMAs consume k-block kb from LDS, so
// a workgroup whose k loop is short on work per k-block (small tiles, one
// workgroup per compute unit) pays the memory latency about once instead of
// once per k-block.  Same LDS layout and MFMA sequence as GemmTile: the results
// are bit-identical.  2- and 4-byte elements.
template <int KA, int KB, int OK, int BM, int BN, int DEPTH>
__device__ __forceinline__ void GemmTileDeep(const TileJob& q, int m0, int n0, uint8_t (*sA)[BM * kRowBytes],
                                             uint8_t (*sB)[BN * kRowBytes])
{
    constexpr int NT = kThreads;
    constexpr int EA = KindInfo<KA>::bytes;
    static_assert(EA == KindInfo<KB>::bytes && (EA == 4 || EA == 2), "2- and 4-byte operands");
    constexpr int kK = kRowBytes / EA;
    constexpr int WM = BM / 2;
    constexpr int WN = BN / 2;
    constexpr int TM = WM / 16;
    constexpr int TN = WN / 16;
    using LA = TileLoader<BM, EA, NT>;
    using LB = TileLoader<BN, EA, NT>;

    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wave = t >> 6;
    const int wm = (wave % 2) * WM;
    const int wn = (wave / 2) * WN;

    LA la;
    LB lb;
    la.Setup(q.A, q.lda, q.aK, m0);
    lb.Setup(q.B, q.ldb, q.bK, n0);
    const int fullK = q.k / kK;
    const bool interiorA = q.vecA && m0 + BM <= q.m;
    const bool interiorB = q.vecB && n0 + BN <= q.n;

    f32x4 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    // C first (beta != 0): its loads are the oldest in flight and complete
    // during the k loop
    float cpre[TM][TN][4];
    if (q.beta != 0.0f) {
        PreloadTileC<OK, TM, TN>(q, m0 + wm, n0 + wn, lane, cpre);
    }

    uint4 ra[DEPTH][LA::kPerThread];
    uint4 rb[DEPTH][LB::kPerThread];
    const int nk = (q.k + kK - 1) / kK;
    auto fetch = [&](int kb, uint4 (&da)[LA::kPerThread], uint4 (&db)[LB::kPerThread]) {
        if (interiorA && kb < fullK) {
            la.FetchFullInto(kb, da);
        } else {
            la.FetchInto(q.A, q.lda, q.aK, q.vecA, m0, q.m, kb * kK, q.k, da);
        }
        if (interiorB && kb < fullK) {
            lb.FetchFullInto(kb, db);
        } else {
            lb.FetchInto(q.B, q.ldb, q.bK, q.vecB, n0, q.n, kb * kK, q.k, db);
        }
    };
    const int fr = lane & 15;
    const int fq = (lane >> 4) * 32;
    uint32_t offA[4], offB[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        offA[i] = SwizzledOffset(wm + i * 16 + fr, fq) - i * 16 * kRowBytes;
        offB[i] = SwizzledOffset(wn + i * 16 + fr, fq) - i * 16 * kRowBytes;
    }

    // Interior tile, whole k-blocks, at least 2 DEPTH of them: the refill
    // groups run without branches, so the compiler can wait for each slot's
    // own loads (vmcnt counts) instead of for every load in flight.
    int kbStart = 0;
    if (interiorA && interiorB && fullK * kK == q.k && nk >= 2 * DEPTH) {
        dev::StaticFor<0, DEPTH>([&](auto sc) {
            constexpr int sl = decltype(sc)::value;
            la.FetchFullInto(sl, ra[sl]);
            lb.FetchFullInto(sl, rb[sl]);
        });
        la.StoreFastFrom(sA[0], q.aK, ra[0]);
        lb.StoreFastFrom(sB[0], q.bK, rb[0]);
        __syncthreads();
        for (; kbStart + 2 * DEPTH <= nk; kbStart += DEPTH) {
            dev::StaticFor<0, DEPTH>([&](auto sc) {
                constexpr int sl = decltype(sc)::value;
                constexpr int nx = (sl + 1) % DEPTH;
                const int kb = kbStart + sl;
                la.FetchFullInto(kb + DEPTH, ra[sl]);
                lb.FetchFullInto(kb + DEPTH, rb[sl]);
                KBlockMfma<KA, KB, TM, TN>(sA[kb & 1], sB[kb & 1], offA, offB, acc);
                la.StoreFastFrom(sA[(kb + 1) & 1], q.aK, ra[nx]);
                lb.StoreFastFrom(sB[(kb + 1) & 1], q.bK, rb[nx]);
                __syncthreads();
            });
        }
    } else {
        dev::StaticFor<0, DEPTH>([&](auto sc) {
            constexpr int sl = decltype(sc)::value;
            if (sl < nk) {
                fetch(sl, ra[sl], rb[sl]);
            }
        });
        if (nk > 0) {
            la.StoreFastFrom(sA[0], q.aK, ra[0]);
            lb.StoreFastFrom(sB[0], q.bK, rb[0]);
        }
        __syncthreads();
    }

    // slot kb % DEPTH holds k-block kb until it is stored to LDS (one
    // iteration before kb is consumed), then takes k-block kb + DEPTH
    for (int kb0 = kbStart; kb0 < nk; kb0 += DEPTH) {
        dev::StaticFor<0, DEPTH>([&](auto sc) {
            constexpr int sl = decltype(sc)::value;
            constexpr int nx = (sl + 1) % DEPTH;
            const int kb = kb0 + sl;
            if (kb < nk) {
                if (kb + DEPTH < nk) {
                    fetch(kb + DEPTH, ra[sl], rb[sl]);
                }
                KBlockMfma<KA, KB, TM, TN>(sA[kb & 1], sB[kb & 1], offA, offB, acc);
                if (kb + 1 < nk) {
                    la.StoreFastFrom(sA[(kb + 1) & 1], q.aK, ra[nx]);
                    lb.StoreFastFrom(sB[(kb + 1) & 1], q.bK, rb[nx]);
                }
                __syncthreads();
            }
        });
    }

    if (q.beta != 0.0f) {
        StoreTileCPre<OK, TM, TN>(q, acc, cpre, m0 + wm, n0 + wn, lane);
    } else {
        StoreTileC<OK, TM, TN, false>(q, acc, m0 + wm, n0 + wn, lane);
    }
}

// One problem (or a uniform batch through pointer arrays): one workgroup per
// tile, WGM x WGN waves.
// DEPTH > 0: GemmTileDeep with that many k-blocks in flight (2 x 2 waves).
template <int KA, int KB, int OK, int BM, int BN, int WGM = 2, int WGN = 2, int DEPTH = 0>
__global__ __launch_bounds__(kWaveSize * WGM * WGN) void GemmKernel(GemmArgs p)
{
    __shared__ __attribute__((aligned(16))) uint8_t sA[2][BM * kRowBytes];
    __shared__ __attribute__((aligned(16))) uint8_t sB[2][BN * kRowBytes];

    // XCD-aware order over the whole grid, then (problem, tile) for batches
    int bid = blockIdx.x;
    if (gridDim.x % 8 == 0) {
        bid = (bid % 8) * (gridDim.x / 8) + bid / 8;
    }
    const int tilesPer = p.tilesM * p.tilesN;
    const int prob = bid / tilesPer;
    int tm, tn;
    TileOf(bid - prob * tilesPer, p.tilesM, p.tilesN, p.group, tm, tn);

    TileJob q;
    q.A = static_cast<const uint8_t*>(p.Aarr != nullptr ? p.Aarr[prob] : p.A);
    q.B = static_cast<const uint8_t*>(p.Barr != nullptr ? p.Barr[prob] : p.B);
    q.C = p.Carr != nullptr ? p.Carr[prob] : p.C;
    q.m = p.m;
    q.n = p.n;
    q.k = p.k;
    q.lda = p.lda;
    q.ldb = p.ldb;
    q.ldc = p.ldc;
    // op(A) row r = C row; k-contiguous when A is transposed.  op(B) column c =
    // C column; k-contiguous when B is not transposed.
    q.aK = p.transA != 0;
    q.bK = p.transB == 0;
    // per-problem pointers of a batch may be less aligned than the host assumed
    q.vecA = p.vecA && (reinterpret_cast<uintptr_t>(q.A) % 16 == 0);
    q.vecB = p.vecB && (reinterpret_cast<uintptr_t>(q.B) % 16 == 0);
    q.vecC = p.vecC && (reinterpret_cast<uintptr_t>(q.C) % 16 == 0);
    float alpha = p.alpha;
    if (p.scaleA != nullptr) alpha *= *p.scaleA;
    if (p.scaleB != nullptr) alpha *= *p.scaleB;
    q.alpha = alpha;
    q.beta = p.beta;
    if constexpr (DEPTH > 0) {
        GemmTileDeep<KA, KB, OK, BM, BN, DEPTH>(q, tm * BM, tn * BN, sA, sB);
    } else {
        GemmTile<KA, KB, OK, BM, BN, WGM, WGN>(q, tm * BM, tn * BN, sA, sB);
    }
}

// Problem record of the persistent grouped kernel (built on the host).
struct GroupedProblem {
    const void* A;
    const void* B;
    void* C;
    int m, n, k;
    int lda, ldb, ldc;
    int transA, transB;
    float alpha, beta;
    int tilesM, tilesN;
    int tileStart;          // first global tile index of this problem
    int vecA, vecB, vecC;
    int slot;               // index into the device pointer arrays (when the launch passes them)
};

// L2 group height of the grouped kernel's tile walk (problems are small: a
// few tile rows each).
constexpr int kGroupedL2Rows = 8;

// Persistent grouped GEMM: the grid strides over the concatenated tile lists
// of all problems (problems ordered by decreasing k on the host, so the
// longest tiles start first); each workgroup finds its problem by binary
// search over tileStart.
// With device pointer arrays (Aarr != nullptr) a record's A, B and C come from
// Aarr/Barr/Carr[slot], and its vector flags hold only the leading-dimension
// part of the alignment test: the pointers are checked here.
template <int KA, int KB, int OK, int BM, int BN>
__global__ __launch_bounds__(kThreads) void GroupedGemmKernel(const GroupedProblem* __restrict__ probs,
                                                             int numProbs, int totalTiles,
                                                             const void* const* __restrict__ Aarr,
                                                             const void* const* __restrict__ Barr,
                                                             void* const* __restrict__ Carr)
{
    __shared__ __attribute__((aligned(16))) uint8_t sA[2][BM * kRowBytes];
    __shared__ __attribute__((aligned(16))) uint8_t sB[2][BN * kRowBytes];
    for (int tile = blockIdx.x; tile < totalTiles; tile += gridDim.x) {
        int lo = 0;
        int hi = numProbs - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (probs[mid].tileStart <= tile) {
                lo = mid;
            } else {
                hi = mid - 1;
            }
        }
        const GroupedProblem& g = probs[lo];
        int tm, tn;
        TileOf(tile - g.tileStart, g.tilesM, g.tilesN, kGroupedL2Rows, tm, tn);
        TileJob q;
        q.A = static_cast<const uint8_t*>(g.A);
        q.B = static_cast<const uint8_t*>(g.B);
        q.C = g.C;
        q.m = g.m;
        q.n = g.n;
        q.k = g.k;
        q.lda = g.lda;
        q.ldb = g.ldb;
        q.ldc = g.ldc;
        q.aK = g.transA != 0;
        q.bK = g.transB == 0;
        q.vecA = g.vecA != 0;
        q.vecB = g.vecB != 0;
        q.vecC = g.vecC != 0;
        if (Aarr != nullptr) {
            q.A = static_cast<const uint8_t*>(Aarr[g.slot]);
            q.B = static_cast<const uint8_t*>(Barr[g.slot]);
            q.C = Carr[g.slot];
            q.vecA = q.vecA && (reinterpret_cast<uintptr_t>(q.A) % 16 == 0);
            q.vecB = q.vecB && (reinterpret_cast<uintptr_t>(q.B) % 16 == 0);
            q.vecC = q.vecC && (reinterpret_cast<uintptr_t>(q.C) % 16 == 0);
        }
        q.alpha = g.alpha;
        q.beta = g.beta;
        GemmTile<KA, KB, OK, BM, BN>(q, tm * BM, tn * BN, sA, sB);
    }
}

}  // namespace gemm
}  // namespace aclblas

#include "gemm_dma.h"

namespace aclblas {
namespace gemm {

// The direct-to-LDS kernels copy an index-contiguous operand (A with transA ==
// N: m rows; B with transB != N: n rows) in chunks of `chunk` elements: they
// need its extent to be a multiple of the chunk.
inline bool DmaExtentsWhole(const GemmArgs& p, int chunk)
{
    return (p.transA != 0 || p.m % chunk == 0) && (p.transB == 0 || p.n % chunk == 0);
}

// Launch with a tuned choice: tile edge `tile` (256: 256 x 256 over 8 waves of
// 128 x 64, 64 KiB of LDS per stage, so that two waves share each SIMD; 128 and
// 64: 2 x 2 waves) and `group` tile rows per L2 group.  A 256 tile is never
// instantiated for 4-byte operands (it would not fit the register budget).
template <int KA, int KB, int OK>
inline void LaunchGemmTyped(hipStream_t st, GemmArgs p, int batch, int tile, int group)
{
    constexpr int EB = KindInfo<KA>::bytes;
    p.group = group;
    if constexpr (KA == kBF16 || KA == kF16) {
        // direct-to-LDS kernel: one problem, 16-byte aligned operands, and an
        // index-contiguous operand's extent a whole number of 16-byte chunks
        // (its DMA copies whole chunks: a partial last chunk would read past
        // the end of the operand's last column)
        if (tile == 256 && p.Aarr == nullptr && p.vecA != 0 && p.vecB != 0 && DmaExtentsWhole(p, 8)) {
            p.tilesM = (p.m + 255) / 256;
            p.tilesN = (p.n + 255) / 256;
            const dim3 grid(p.tilesM * p.tilesN);
            const dim3 block(dma::kThreads);
            // m-contiguous A: 32-k stages (transposing reads, conflict-free);
            // k-contiguous A: 64-k stages (128-byte rows, conflict-free)
            if (p.transA == 0 && p.transB == 0) {
                hipLaunchKernelGGL((GemmDmaKernel<KA, OK, true, false>), grid, block, 0, st, p);
            } else if (p.transA == 0) {
                hipLaunchKernelGGL((GemmDmaKernel<KA, OK, true, true>), grid, block, 0, st, p);
            } else if (p.transB == 0) {
                hipLaunchKernelGGL((GemmDma64Kernel<KA, OK, false, false>), grid, block, 0, st, p);
            } else {
                hipLaunchKernelGGL((GemmDma64Kernel<KA, OK, false, true>), grid, block, 0, st, p);
            }
            return;
        }
    }
    if constexpr (KA == kE4M3S || KA == kE5M2S) {
        // direct-to-LDS FP8 kernel: one problem, 16-byte aligned operands, whole
        // 16-byte chunks along an index-contiguous operand
        if (tile == 256 && p.Aarr == nullptr && p.vecA != 0 && p.vecB != 0 && DmaExtentsWhole(p, 16)) {
            p.tilesM = (p.m + 255) / 256;
            p.tilesN = (p.n + 255) / 256;
            const dim3 grid(p.tilesM * p.tilesN);
            const dim3 block(dma::kThreads);
            if (p.transA == 0 && p.transB == 0) {
                hipLaunchKernelGGL((GemmDmaFp8Kernel<KA, KB, OK, true, false>), grid, block, 0, st, p);
            } else if (p.transA == 0) {
                hipLaunchKernelGGL((GemmDmaFp8Kernel<KA, KB, OK, true, true>), grid, block, 0, st, p);
            } else if (p.transB == 0) {
                hipLaunchKernelGGL((GemmDmaFp8Kernel<KA, KB, OK, false, false>), grid, block, 0, st, p);
            } else {
                hipLaunchKernelGGL((GemmDmaFp8Kernel<KA, KB, OK, false, true>), grid, block, 0, st, p);
            }
            return;
        }
    }
    if constexpr (EB <= 2) {
        if (tile == 256) {
            p.tilesM = (p.m + 255) / 256;
            p.tilesN = (p.n + 255) / 256;
            hipLaunchKernelGGL((GemmKernel<KA, KB, OK, 256, 256, 2, 4>), dim3(p.tilesM * p.tilesN * batch),
                               dim3(kWaveSize * 8), 0, st, p);
            return;
        }
    }
    if (tile == 128) {
        p.tilesM = (p.m + 127) / 128;
        p.tilesN = (p.n + 127) / 128;
        hipLaunchKernelGGL((GemmKernel<KA, KB, OK, 128, 128>), dim3(p.tilesM * p.tilesN * batch), dim3(kThreads),
                           0, st, p);
        return;
    }
    if constexpr (EB == 4) {
        // 32 x 32 tiles with eight k-blocks in flight, 64 x 64 with four
        if (tile == 32) {
            p.tilesM = (p.m + 31) / 32;
            p.tilesN = (p.n + 31) / 32;
            hipLaunchKernelGGL((GemmKernel<KA, KB, OK, 32, 32, 2, 2, 8>), dim3(p.tilesM * p.tilesN * batch),
                               dim3(kThreads), 0, st, p);
            return;
        }
        if (tile == 65) {
            p.tilesM = (p.m + 63) / 64;
            p.tilesN = (p.n + 63) / 64;
            hipLaunchKernelGGL((GemmKernel<KA, KB, OK, 64, 64, 2, 2, 4>), dim3(p.tilesM * p.tilesN * batch),
                               dim3(kThreads), 0, st, p);
            return;
        }
    }
    p.tilesM = (p.m + 63) / 64;
    p.tilesN = (p.n + 63) / 64;
    hipLaunchKernelGGL((GemmKernel<KA, KB, OK, 64, 64>), dim3(p.tilesM * p.tilesN * batch), dim3(kThreads), 0, st,
                       p);
}

// Built-in tile when the tuning table has no entry: 256 for large 1- and
// 2-byte GEMMs, 128 when its grid fills the compute units, else 64.
inline int DefaultGemmTile(int elemBytes, int m, int n, int batch, int cus)
{
    const long long t256 = static_cast<long long>((m + 255) / 256) * ((n + 255) / 256) * batch;
    const long long t128 = static_cast<long long>((m + 127) / 128) * ((n + 127) / 128) * batch;
    if (elemBytes <= 2 && t256 >= cus) {
        return 256;
    }
    return t128 >= cus ? 128 : 64;
}

}  // namespace gemm
}  // namespace aclblas
