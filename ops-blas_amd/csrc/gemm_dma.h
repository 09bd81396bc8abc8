// 16-bit (BF16 / FP16) GEMM tile kernel with direct-to-LDS operand loads.
//
// One workgroup of 8 waves (2 x 4, each 128 x 64) computes a 256 x 256 tile
// of C.  The k dimension is consumed in stages of 32; every operand stage
// (256 x 32 values, 16 KiB) is copied from global memory straight into LDS by
// 16 one-KiB wave instructions (global_load_lds_dwordx4), with no register
// staging, so a ring of four LDS stages (128 KiB) keeps three stages of loads
// in flight while the MFMAs consume the fourth:
//
//   issue stages 0, 1, 2
//   for s: wait for this wave's loads of stage s; barrier; compute stage s,
//          issuing stage s + 3 into the buffer of stage s - 1 after its first
//          eight MFMAs
//
// LDS layouts (both conflict-free for their reads):
//   K-major (operand stored k-contiguous): 256 rows x 64 bytes (32 k), 16-byte
//     chunk q of row r at slot q ^ KmKey(r); a fragment is one ds_read_b128 of
//     8 consecutive k (served 8 lanes = 128 bytes at a time).
//   M-major (operand stored m- or n-contiguous): 32 rows (one k each) x 512
//     bytes, chunk c of row kk at slot c ^ MmKey(kk); a fragment is two
//     transposing reads (ds_read_b64_tr_b16) of 4 k each.
// The DMA writes 16-byte slots in address order, so each lane picks the global
// chunk that belongs in its slot.
//
// Stage k order: stage 2p + h holds k = 64p + 16 (kk / 8) + 8h + kk % 8 for
// stage-local kk, so each stage feeds one v_mfma_f32_16x16x32 whose products
// and order are those of the 128-byte-row kernel (GemmTile): both kernels give
// bit-identical results.  Whole 64-k pairs are loaded by DMA; a final partial
// pair is loaded through registers with bounds checks and zero fill.  Rows past
// the matrix edge read clamped (valid) addresses; their results are not stored.
#pragma once

#include "lds_dma.h"

namespace aclblas {
namespace gemm {
namespace dma {

constexpr int kBM = 256;
constexpr int kBN = 256;
constexpr int kKS = 32;                           // k per stage
constexpr int kStages = 4;
constexpr int kWGM = 2;
constexpr int kWGN = 4;
constexpr int kThreads = kWaveSize * kWGM * kWGN;  // 512
constexpr int kOpBytes = 256 * kKS * 2;           // one operand stage: 16 KiB
constexpr int kWaves = kWGM * kWGN;

__device__ __forceinline__ int StageK(int s, int kk)
{
    return 64 * (s >> 1) + 16 * (kk >> 3) + 8 * (s & 1) + (kk & 7);
}
// K-major swizzle key: the 16-byte slots of rows r..r+7 (and r..r+15) read at
// one chunk land on distinct positions of a 128-byte (and 256-byte) bank
// window.  It depends on r mod 16 only, so fragments 16 rows apart share it.
__device__ __forceinline__ int KmKey(int r) { return ((r >> 1) & 3) ^ ((r >> 3) & 1); }
__device__ __forceinline__ uint32_t KmOffset(int r, int q)
{
    return static_cast<uint32_t>(r * 64 + ((q ^ KmKey(r)) << 4));
}
__device__ __forceinline__ int MmKey(int kk) { return (2 * (kk & 3)) ^ (8 * ((kk >> 3) & 1)); }
__device__ __forceinline__ uint32_t MmOffset(int kk, int c)
{
    return static_cast<uint32_t>(kk * 512 + ((c ^ MmKey(kk)) << 4));
}

// Per-lane DMA sources of one operand: element pointers of this lane's two
// 1-KiB instructions (wave w issues instructions w and w + 8) at stage 0, and
// the element steps from one stage pair to the next and from the first to the
// second stage of a pair.
struct DmaSource {
    const uint16_t* src[2];
    long long pairStep;
    long long halfStep;

    // MM: the operand is stored index-contiguous (element (r, k) at X[r + k ld]),
    // else k-contiguous (X[k + r ld]).  rows: extent of the tile index.
    template <bool MM>
    __device__ __forceinline__ void Setup(const uint16_t* X, int ld, int row0, int rows, int wave, int lane)
    {
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int j = wave + kWaves * i;
            if (MM) {
                const int kk = 2 * j + (lane >> 5);
                const int c = (lane & 31) ^ MmKey(kk);
                const int last = ((rows + 7) & ~7) - 8;  // last 8-aligned chunk inside the column
                const int r = min(row0 + 8 * c, last);
                const int k = 16 * (kk >> 3) + (kk & 7);
                src[i] = X + static_cast<size_t>(k) * ld + r;
            } else {
                const int r = min(row0 + 16 * j + (lane >> 2), rows - 1);
                const int q = (lane & 3) ^ KmKey(16 * j + (lane >> 2));
                src[i] = X + static_cast<size_t>(r) * ld + 16 * q;
            }
        }
        pairStep = MM ? 64ll * ld : 64ll;
        halfStep = MM ? 8ll * ld : 8ll;
    }

    __device__ __forceinline__ void Issue(int s, uint8_t* buf, int wave) const
    {
        const long long off = static_cast<long long>(s >> 1) * pairStep + (s & 1) * halfStep;
        const uint32_t base = dev::LdsAddr(buf);
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            dev::LdsDma16(src[i] + off, __builtin_amdgcn_readfirstlane(base + 1024u * (wave + kWaves * i)));
        }
    }
};

// Register-path load of stage s of one operand (bounds-checked, zero fill):
// used for the stages of a final partial 64-k pair.
template <bool MM>
__device__ __forceinline__ void FillStage(uint8_t* buf, const uint16_t* X, int ld, int row0, int rows, int s, int K,
                                          int t)
{
    for (int e = t; e < 256 * kKS; e += kThreads) {
        const int r = MM ? (e & 255) : (e >> 5);
        const int kk = MM ? (e >> 8) : (e & 31);
        const int k = StageK(s, kk);
        uint16_t v = 0;
        if (row0 + r < rows && k < K) {
            v = MM ? X[static_cast<size_t>(k) * ld + row0 + r] : X[static_cast<size_t>(row0 + r) * ld + k];
        }
        const uint32_t off = MM ? MmOffset(kk, r >> 3) + 2u * (r & 7) : KmOffset(r, kk >> 3) + 2u * (kk & 7);
        *reinterpret_cast<uint16_t*>(buf + off) = v;
    }
}

// Byte offset base of a lane's transposing fragment reads in an M-major stage
// for the wave's fragments starting at tile index wbase: fragment f (16
// indices, f below the alignment of wbase / 16) reads at TrBase ^ (f << 5),
// and its second 4 k at that + 2048.
__device__ __forceinline__ uint32_t TrBase(int wbase, int lane)
{
    const int g = lane >> 4;
    const int j = lane & 15;
    const int kk = 8 * g + (j >> 2);
    const int f = MmKey(kk) >> 1;
    return static_cast<uint32_t>(kk * 512 + (((wbase >> 4) ^ f) << 5) + (j & 3) * 8);
}

}  // namespace dma

// 256 x 256 tile of one problem (no batch), both operands 16-bit (KA),
// direct-to-LDS loads.  AMM: op(A) stored m-contiguous (A not transposed);
// BMM: op(B) stored n-contiguous (B transposed).  The host launches it only
// when A, B and their leading dimensions are 16-byte aligned.
template <int KA, int OK, bool AMM, bool BMM>
__global__ __launch_bounds__(dma::kThreads) void GemmDmaKernel(GemmArgs p)
{
    using namespace dma;
    __shared__ __attribute__((aligned(1024))) uint8_t lds[kStages][2][kOpBytes];

    int bid = blockIdx.x;
    if (gridDim.x % 8 == 0) {
        bid = (bid % 8) * (gridDim.x / 8) + bid / 8;
    }
    int tm, tn;
    TileOf(bid, p.tilesM, p.tilesN, p.group, tm, tn);
    const int m0 = tm * kBM;
    const int n0 = tn * kBN;
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
    const int wm = (wave % kWGM) * (kBM / kWGM);
    const int wn = (wave / kWGM) * (kBN / kWGN);
    const uint16_t* A = static_cast<const uint16_t*>(p.A);
    const uint16_t* B = static_cast<const uint16_t*>(p.B);
    const int K = p.k;
    const int nFull = 2 * (K / 64);
    const int nst = 2 * ((K + 63) / 64);

    DmaSource da;
    DmaSource db;
    da.Setup<AMM>(A, p.lda, m0, p.m, wave, lane);
    db.Setup<BMM>(B, p.ldb, n0, p.n, wave, lane);

    auto issue = [&](int s) {
        uint8_t* bufA = lds[s % kStages][0];
        uint8_t* bufB = lds[s % kStages][1];
        if (s < nFull) {
            da.Issue(s, bufA, wave);
            db.Issue(s, bufB, wave);
        } else {
            FillStage<AMM>(bufA, A, p.lda, m0, p.m, s, K, t);
            FillStage<BMM>(bufB, B, p.ldb, n0, p.n, s, K, t);
        }
    };

    // fragment read offsets (bytes within an operand stage)
    const uint32_t offA = AMM ? TrBase(wm, lane) : KmOffset(wm + (lane & 15), lane >> 4);
    const uint32_t offB = BMM ? TrBase(wn, lane) : KmOffset(wn + (lane & 15), lane >> 4);

    constexpr int TM = kBM / kWGM / 16;  // 8
    constexpr int TN = kBN / kWGN / 16;  // 4
    f32x4 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    for (int s = 0; s < 3 && s < nst; ++s) {
        issue(s);
    }
    for (int s = 0; s < nst; ++s) {
        // this wave's DMA stages issued after s may stay in flight (4 loads each)
        const int pending = (s + 1 < nFull ? 1 : 0) + (s + 2 < nFull ? 1 : 0);
        if (pending == 2) {
            dev::WaitVmcnt<8>();
        } else if (pending == 1) {
            dev::WaitVmcnt<4>();
        } else {
            dev::WaitVmcnt<0>();
        }
        dev::BarrierLds();
        // raised wave priority over the MFMA block: the SIMD's other wave
        // (address arithmetic, waits) yields issue slots to it
        dev::SetPriority<1>();
        const uint8_t* a = lds[s % kStages][0];
        const uint8_t* b = lds[s % kStages][1];
        uint4 fb[TN];
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            if (BMM) {
                const uint32_t o = offB ^ (static_cast<uint32_t>(j) << 5);
                const uint2 lo = dev::ReadTr16(b + o);
                const uint2 hi = dev::ReadTr16(b + o + 2048);
                fb[j] = make_uint4(lo.x, lo.y, hi.x, hi.y);
            } else {
                fb[j] = *reinterpret_cast<const uint4*>(b + offB + j * 1024);
            }
        }
#pragma unroll
        for (int i = 0; i < TM; ++i) {
            uint4 fa;
            if (AMM) {
                const uint32_t o = offA ^ (static_cast<uint32_t>(i) << 5);
                const uint2 lo = dev::ReadTr16(a + o);
                const uint2 hi = dev::ReadTr16(a + o + 2048);
                fa = make_uint4(lo.x, lo.y, hi.x, hi.y);
            } else {
                fa = *reinterpret_cast<const uint4*>(a + offA + i * 1024);
            }
#pragma unroll
            for (int j = 0; j < TN; ++j) {
                Mfma<KA, KA>::RunHalf(fa, fb[j], acc[i][j]);
            }
            // the next stage's loads go out once the MFMA pipe is fed
            if (i == 1 && s + 3 < nst) {
                issue(s + 3);
            }
        }
        dev::SetPriority<0>();
    }

    TileJob q;
    q.C = p.C;
    q.m = p.m;
    q.n = p.n;
    q.k = p.k;
    q.ldc = p.ldc;
    q.vecC = p.vecC != 0;
    q.alpha = p.alpha;
    q.beta = p.beta;
    StoreTileC<OK, TM, TN, false>(q, acc, m0 + wm, n0 + wn, lane);
}

// ---------------------------------------------------------------------------
// FP8 operands on the native v_mfma_scale_f32_16x16x128_f8f6f4 (the scaled
// entry).  Same tile and waves; stages of 128 k (one MFMA per fragment pair),
// two LDS stages of 64 KiB, one stage in flight while the other is consumed.
//   K-major: 256 rows x 128 bytes, the register-staged kernel's swizzle
//     (SwizzledOffset); a fragment is two ds_read_b128 of 16 consecutive k.
//   M-major: 128 rows (one k each) x 256 bytes, 16-byte chunk c of row kk at
//     slot c ^ Mm8Key(kk); a fragment is four ds_read_b64_tr_b8 of 8 k each.
// The MFMA of stage s sums k = 128 s + 32 (lane / 16) + 0..31, as the
// register-staged kernel's does for its k-block s: bit-identical results.
// ---------------------------------------------------------------------------
namespace dma8 {

constexpr int kKS = 128;
constexpr int kStages = 2;
constexpr int kOpBytes = 256 * kKS;  // 32 KiB per operand and stage

__device__ __forceinline__ int Mm8Key(int kk) { return (kk & 7) ^ (8 * ((kk >> 5) & 1)); }
__device__ __forceinline__ uint32_t Mm8Offset(int kk, int c)
{
    return static_cast<uint32_t>(kk * 256 + ((c ^ Mm8Key(kk)) << 4));
}

// Per-lane DMA sources of one operand for its four 1-KiB instructions per
// wave and stage (wave w issues w, w + 8, w + 16, w + 24), as byte pointers at
// stage 0, and the byte step from one stage to the next.
struct DmaSource8 {
    const uint8_t* src[4];
    long long step;

    template <bool MM>
    __device__ __forceinline__ void Setup(const uint8_t* X, int ld, int row0, int rows, int wave, int lane)
    {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int j = wave + dma::kWaves * i;
            if (MM) {
                const int kk = 4 * j + (lane >> 4);
                const int c = (lane & 15) ^ Mm8Key(kk);
                const int last = ((rows + 15) & ~15) - 16;  // last 16-aligned chunk inside the column
                const int r = min(row0 + 16 * c, last);
                src[i] = X + static_cast<size_t>(kk) * ld + r;
            } else {
                const int rl = 8 * j + (lane >> 3);
                const int r = min(row0 + rl, rows - 1);
                const int c = (lane & 7) ^ ((rl ^ (rl >> 3)) & 7);
                src[i] = X + static_cast<size_t>(r) * ld + 16 * c;
            }
        }
        step = MM ? static_cast<long long>(kKS) * ld : kKS;
    }

    __device__ __forceinline__ void Issue(int s, uint8_t* buf, int wave) const
    {
        const long long off = static_cast<long long>(s) * step;
        const uint32_t base = dev::LdsAddr(buf);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            dev::LdsDma16(src[i] + off, __builtin_amdgcn_readfirstlane(base + 1024u * (wave + dma::kWaves * i)));
        }
    }
};

// Register-path load of a partial stage (k tail), bounds-checked, zero fill.
template <bool MM>
__device__ __forceinline__ void FillStage8(uint8_t* buf, const uint8_t* X, int ld, int row0, int rows, int s, int K,
                                           int t)
{
    for (int e = t; e < 256 * kKS; e += dma::kThreads) {
        const int r = MM ? (e & 255) : (e >> 7);
        const int kk = MM ? (e >> 8) : (e & 127);
        const int k = kKS * s + kk;
        uint8_t v = 0;
        if (row0 + r < rows && k < K) {
            v = MM ? X[static_cast<size_t>(k) * ld + row0 + r] : X[static_cast<size_t>(row0 + r) * ld + k];
        }
        const uint32_t off = MM ? Mm8Offset(kk, r >> 4) + (r & 15) : SwizzledOffset(r, kk);
        buf[off] = v;
    }
}

// Byte offset base of a lane's transposing reads in an M-major stage for the
// wave's fragments starting at tile index wbase: fragment f (16 indices, f
// below the alignment of wbase / 16) reads 8 k at (TrBase8 ^ (f << 4)) +
// 2048 t, t = 0..3.
__device__ __forceinline__ uint32_t TrBase8(int wbase, int lane)
{
    const int g = lane >> 4;
    const int j = lane & 15;
    const int key = (j >> 1) ^ (8 * (g & 1));
    return static_cast<uint32_t>((32 * g + (j >> 1)) * 256 + 8 * (j & 1) + (((wbase >> 4) ^ key) << 4));
}

}  // namespace dma8

// 256 x 256 tile, FP8 operands (KA, KB in kE4M3S / kE5M2S), direct-to-LDS
// loads; AMM / BMM as for GemmDmaKernel.  scaleA * scaleB * alpha in the
// epilogue.
template <int KA, int KB, int OK, bool AMM, bool BMM>
__global__ __launch_bounds__(dma::kThreads) void GemmDmaFp8Kernel(GemmArgs p)
{
    using namespace dma8;
    __shared__ __attribute__((aligned(1024))) uint8_t lds[kStages][2][kOpBytes];

    int bid = blockIdx.x;
    if (gridDim.x % 8 == 0) {
        bid = (bid % 8) * (gridDim.x / 8) + bid / 8;
    }
    int tm, tn;
    TileOf(bid, p.tilesM, p.tilesN, p.group, tm, tn);
    const int m0 = tm * dma::kBM;
    const int n0 = tn * dma::kBN;
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
    const int wm = (wave % dma::kWGM) * (dma::kBM / dma::kWGM);
    const int wn = (wave / dma::kWGM) * (dma::kBN / dma::kWGN);
    const uint8_t* A = static_cast<const uint8_t*>(p.A);
    const uint8_t* B = static_cast<const uint8_t*>(p.B);
    const int K = p.k;
    const int nFull = K / kKS;
    const int nst = (K + kKS - 1) / kKS;

    DmaSource8 da;
    DmaSource8 db;
    da.Setup<AMM>(A, p.lda, m0, p.m, wave, lane);
    db.Setup<BMM>(B, p.ldb, n0, p.n, wave, lane);

    auto issue = [&](int s) {
        uint8_t* bufA = lds[s % kStages][0];
        uint8_t* bufB = lds[s % kStages][1];
        if (s < nFull) {
            da.Issue(s, bufA, wave);
            db.Issue(s, bufB, wave);
        } else {
            FillStage8<AMM>(bufA, A, p.lda, m0, p.m, s, K, t);
            FillStage8<BMM>(bufB, B, p.ldb, n0, p.n, s, K, t);
        }
    };

    const int fr = lane & 15;
    const int fq = (lane >> 4) * 32;
    // K-major fragment offsets for row groups i % 4 (the swizzle key repeats every 64 rows)
    uint32_t kmA[4], kmB[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        kmA[i] = SwizzledOffset(wm + i * 16 + fr, fq);
        kmB[i] = SwizzledOffset(wn + i * 16 + fr, fq);
    }
    const uint32_t trA = TrBase8(wm, lane);
    const uint32_t trB = TrBase8(wn, lane);

    constexpr int TM = dma::kBM / dma::kWGM / 16;  // 8
    constexpr int TN = dma::kBN / dma::kWGN / 16;  // 4
    f32x4 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    auto fragment = [&](const uint8_t* buf, bool mm, uint32_t km, uint32_t tr, int f, uint4 (&out)[2]) {
        if (mm) {
            const uint32_t o = tr ^ (static_cast<uint32_t>(f) << 4);
            const uint2 r0 = dev::ReadTr8(buf + o);
            const uint2 r1 = dev::ReadTr8(buf + o + 2048);
            const uint2 r2 = dev::ReadTr8(buf + o + 4096);
            const uint2 r3 = dev::ReadTr8(buf + o + 6144);
            out[0] = make_uint4(r0.x, r0.y, r1.x, r1.y);
            out[1] = make_uint4(r2.x, r2.y, r3.x, r3.y);
        } else {
            out[0] = *reinterpret_cast<const uint4*>(buf + km);
            out[1] = *reinterpret_cast<const uint4*>(buf + (km ^ 16u));
        }
    };

    if (nst > 0) {
        issue(0);
    }
    for (int s = 0; s < nst; ++s) {
        // this wave's DMA for stage s is the only one in flight
        dev::WaitVmcnt<0>();
        dev::BarrierLds();
        dev::SetPriority<1>();
        const uint8_t* a = lds[s % kStages][0];
        const uint8_t* b = lds[s % kStages][1];
        uint4 fb[TN][2];
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            fragment(b, BMM, kmB[j % 4] + (j / 4) * 8192u, trB, j, fb[j]);
        }
#pragma unroll
        for (int i = 0; i < TM; ++i) {
            uint4 fa[2];
            fragment(a, AMM, kmA[i % 4] + (i / 4) * 8192u, trA, i, fa);
#pragma unroll
            for (int j = 0; j < TN; ++j) {
                Mfma<KA, KB>::Run(fa, fb[j], acc[i][j]);
            }
            // the next stage's loads go out once the MFMA pipe is fed
            if (i == 1 && s + 1 < nst) {
                issue(s + 1);
            }
        }
        dev::SetPriority<0>();
    }

    TileJob q;
    q.C = p.C;
    q.m = p.m;
    q.n = p.n;
    q.k = p.k;
    q.ldc = p.ldc;
    q.vecC = p.vecC != 0;
    float alpha = p.alpha;
    if (p.scaleA != nullptr) alpha *= *p.scaleA;
    if (p.scaleB != nullptr) alpha *= *p.scaleB;
    q.alpha = alpha;
    q.beta = p.beta;
    StoreTileC<OK, TM, TN, false>(q, acc, m0 + wm, n0 + wn, lane);
}

// ---------------------------------------------------------------------------
// 16-bit operands with 64-k stages: two LDS stages of 64 KiB, one barrier per
// 64 k, one stage in flight (the FP8 kernel's pipeline).
//   K-major: 256 rows x 128 bytes (64 k), the register-staged kernel's
//     swizzle (SwizzledOffset); a fragment half is one ds_read_b128.
//   M-major: 64 rows (one k each) x 512 bytes, chunk c of row kk at slot
//     c ^ Mm2Key(kk); a fragment half is two ds_read_b64_tr_b16.
// Per 64 k each accumulator takes the MFMA of k = 16 (lane / 16) + 0..7 and
// then that of + 8..15, as in the register-staged kernel: bit-identical.
// ---------------------------------------------------------------------------
namespace dma2 {

constexpr int kKS = 64;
constexpr int kStages = 2;
constexpr int kOpBytes = 256 * 128;  // 32 KiB per operand and stage

__device__ __forceinline__ int Mm2Key(int kk) { return (2 * (kk & 3)) ^ (8 * ((kk >> 4) & 1)); }
__device__ __forceinline__ uint32_t Mm2Offset(int kk, int c)
{
    return static_cast<uint32_t>(kk * 512 + ((c ^ Mm2Key(kk)) << 4));
}

// DMA sources (byte pointers at stage 0) of this lane's four 1-KiB
// instructions per operand and stage, and the byte step per stage.
struct DmaSource2 {
    const uint8_t* src[4];
    long long step;

    template <bool MM>
    __device__ __forceinline__ void Setup(const uint16_t* X, int ld, int row0, int rows, int wave, int lane)
    {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int j = wave + dma::kWaves * i;
            if (MM) {
                const int kk = 2 * j + (lane >> 5);
                const int c = (lane & 31) ^ Mm2Key(kk);
                const int last = ((rows + 7) & ~7) - 8;
                const int r = min(row0 + 8 * c, last);
                src[i] = reinterpret_cast<const uint8_t*>(X + static_cast<size_t>(kk) * ld + r);
            } else {
                const int rl = 8 * j + (lane >> 3);
                const int r = min(row0 + rl, rows - 1);
                const int c = (lane & 7) ^ ((rl ^ (rl >> 3)) & 7);
                src[i] = reinterpret_cast<const uint8_t*>(X + static_cast<size_t>(r) * ld) + 16 * c;
            }
        }
        step = MM ? 2ll * kKS * ld : 2ll * kKS;
    }

    __device__ __forceinline__ void Issue(int s, uint8_t* buf, int wave) const
    {
        const long long off = static_cast<long long>(s) * step;
        const uint32_t base = dev::LdsAddr(buf);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            dev::LdsDma16(src[i] + off, __builtin_amdgcn_readfirstlane(base + 1024u * (wave + dma::kWaves * i)));
        }
    }
};

template <bool MM>
__device__ __forceinline__ void FillStage2(uint8_t* buf, const uint16_t* X, int ld, int row0, int rows, int s, int K,
                                           int t)
{
    for (int e = t; e < 256 * kKS; e += dma::kThreads) {
        const int r = MM ? (e & 255) : (e >> 6);
        const int kk = MM ? (e >> 8) : (e & 63);
        const int k = kKS * s + kk;
        uint16_t v = 0;
        if (row0 + r < rows && k < K) {
            v = MM ? X[static_cast<size_t>(k) * ld + row0 + r] : X[static_cast<size_t>(row0 + r) * ld + k];
        }
        const uint32_t off = MM ? Mm2Offset(kk, r >> 3) + 2u * (r & 7) : SwizzledOffset(r, 2 * kk);
        *reinterpret_cast<uint16_t*>(buf + off) = v;
    }
}

// Transposing-read base: fragment f reads its half h at
// ((Tr2Base ^ (f << 5)) + 4096 h) and the second 4 k at + 2048.
__device__ __forceinline__ uint32_t Tr2Base(int wbase, int lane)
{
    const int g = lane >> 4;
    const int j = lane & 15;
    const int key = ((j >> 2) & 3) ^ (4 * (g & 1));
    return static_cast<uint32_t>((16 * g + (j >> 2)) * 512 + (((wbase >> 4) ^ key) << 5) + (j & 3) * 8);
}

}  // namespace dma2

template <int KA, int OK, bool AMM, bool BMM>
__global__ __launch_bounds__(dma::kThreads) void GemmDma64Kernel(GemmArgs p)
{
    using namespace dma2;
    __shared__ __attribute__((aligned(1024))) uint8_t lds[kStages][2][kOpBytes];

    int bid = blockIdx.x;
    if (gridDim.x % 8 == 0) {
        bid = (bid % 8) * (gridDim.x / 8) + bid / 8;
    }
    int tm, tn;
    TileOf(bid, p.tilesM, p.tilesN, p.group, tm, tn);
    const int m0 = tm * dma::kBM;
    const int n0 = tn * dma::kBN;
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
    const int wm = (wave % dma::kWGM) * (dma::kBM / dma::kWGM);
    const int wn = (wave / dma::kWGM) * (dma::kBN / dma::kWGN);
    const uint16_t* A = static_cast<const uint16_t*>(p.A);
    const uint16_t* B = static_cast<const uint16_t*>(p.B);
    const int K = p.k;
    const int nFull = K / kKS;
    const int nst = (K + kKS - 1) / kKS;

    DmaSource2 da;
    DmaSource2 db;
    da.Setup<AMM>(A, p.lda, m0, p.m, wave, lane);
    db.Setup<BMM>(B, p.ldb, n0, p.n, wave, lane);

    auto issue = [&](int s) {
        uint8_t* bufA = lds[s % kStages][0];
        uint8_t* bufB = lds[s % kStages][1];
        if (s < nFull) {
            da.Issue(s, bufA, wave);
            db.Issue(s, bufB, wave);
        } else {
            FillStage2<AMM>(bufA, A, p.lda, m0, p.m, s, K, t);
            FillStage2<BMM>(bufB, B, p.ldb, n0, p.n, s, K, t);
        }
    };

    const int fr = lane & 15;
    const int fq = (lane >> 4) * 32;
    uint32_t kmA[4], kmB[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        kmA[i] = SwizzledOffset(wm + i * 16 + fr, fq);
        kmB[i] = SwizzledOffset(wn + i * 16 + fr, fq);
    }
    const uint32_t trA = Tr2Base(wm, lane);
    const uint32_t trB = Tr2Base(wn, lane);

    constexpr int TM = dma::kBM / dma::kWGM / 16;  // 8
    constexpr int TN = dma::kBN / dma::kWGN / 16;  // 4
    f32x4 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    auto half = [&](const uint8_t* buf, bool mm, uint32_t km, uint32_t tr, int f, int h) -> uint4 {
        if (mm) {
            const uint32_t o = (tr ^ (static_cast<uint32_t>(f) << 5)) + 4096u * h;
            const uint2 lo = dev::ReadTr16(buf + o);
            const uint2 hi = dev::ReadTr16(buf + o + 2048);
            return make_uint4(lo.x, lo.y, hi.x, hi.y);
        }
        return *reinterpret_cast<const uint4*>(buf + (km ^ (16u * h)));
    };

    if (nst > 0) {
        issue(0);
    }
    for (int s = 0; s < nst; ++s) {
        // this wave's DMA for stage s is the only one in flight
        dev::WaitVmcnt<0>();
        dev::BarrierLds();
        dev::SetPriority<1>();
        const uint8_t* a = lds[s % kStages][0];
        const uint8_t* b = lds[s % kStages][1];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            uint4 fb[TN];
#pragma unroll
            for (int j = 0; j < TN; ++j) {
                fb[j] = half(b, BMM, kmB[j % 4] + (j / 4) * 8192u, trB, j, h);
            }
#pragma unroll
            for (int i = 0; i < TM; ++i) {
                const uint4 fa = half(a, AMM, kmA[i % 4] + (i / 4) * 8192u, trA, i, h);
#pragma unroll
                for (int j = 0; j < TN; ++j) {
                    Mfma<KA, KA>::RunHalf(fa, fb[j], acc[i][j]);
                }
                // the next stage's loads go out once the MFMA pipe is fed
                if (h == 0 && i == 1 && s + 1 < nst) {
                    issue(s + 1);
                }
            }
        }
        dev::SetPriority<0>();
    }

    TileJob q;
    q.C = p.C;
    q.m = p.m;
    q.n = p.n;
    q.k = p.k;
    q.ldc = p.ldc;
    q.vecC = p.vecC != 0;
    q.alpha = p.alpha;
    q.beta = p.beta;
    StoreTileC<OK, TM, TN, false>(q, acc, m0 + wm, n0 + wn, lane);
}

}  // namespace gemm
}  // namespace aclblas
