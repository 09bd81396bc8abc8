// aclblasSgesvBatchedMixed / aclblasSgesvBatchedMixedEx: batched dense solves
// of mixed orders (n <= 256, nrhs <= 16) with a blocked LU whose trailing
// updates run on matrix cores in a chosen precision, and FP64-residual
// iterative refinement.
//
// The reference library has no such entry.  Its building blocks follow the
// reference's getrf/getrs semantics (partial pivoting, LAPACK-style pivot
// sequence; blas/getrf_batched/arch35/sgetrf_batched_kernel.cpp:52-150,
// blas/getrs_batched/arch35/sgetrs_batched_kernel.cpp:83-259); the algorithm is
// blocked and mixed-precision, so its results are compared with an FP64
// solver through the backward error, not bit by bit.
//
// Trailing-update precisions (aclblasMixedUpdate_t):
//   FP32    A22 += (-L21) U12 on the FP32 matrix instruction
//           (v_mfma_f32_16x16x4_f32, eight per 16 x 16 x 32 block);
//   BF16    both operands rounded to BF16, one v_mfma_f32_16x16x32_bf16;
//   BF16X2  both operands split into two BF16 planes (hi + lo, about 16
//           significant bits), three products (lo*hi, hi*lo, hi*hi);
//   BF16X3  three planes (the FP32 value exactly), six products (the terms
//           below 2^-24 relative dropped).
// Panels, pivoting, the 32 x 32 diagonal-block inverses, U12 and every
// accumulation are FP32 in all modes.
//
// Factorisation (GesvFactorKernel<P>, persistent, one 256-thread workgroup
// per system, systems in order of decreasing n).  The system is copied to an
// FP32 working copy and factored in panels of 32 columns:
//   1. the four waves hold the panel's rows (one row per lane) in registers
//      and factor it with partial pivoting: rows keep their place and carry
//      their position; per column each wave reduces its |a| keys with DPP
//      (lowest position on exact ties), publishes its best row in LDS, and
//      after one barrier every lane picks the same winner; multipliers are
//      a * (1/p);
//   2. the panel is written back in pivoted row order with its row map (new
//      position -> old row), L11 kept in LDS and -L21 in the update's operand
//      format;
//   3. inv(L11) (32 x 32, identity-padded for a short last panel) by column
//      substitution, kept for U12 and for the solves;
//   4. the four waves sweep the trailing 16-column tiles: a tile's rows are
//      gathered through the row map into MFMA accumulators (one row per lane
//      of a 16-lane group, four columns per group, so a group reads 64
//      contiguous bytes), U12 = inv(L11) A12 runs on the FP32 matrix
//      instruction, U12 is stored in the update's operand format, and
//      A22 += (-L21) U12 in the chosen precision with the accumulators in the
//      transposed orientation (U12^T times (-L21)^T); the tile goes back to
//      the working copy in position order.
// Columns left of a panel are not permuted by later panels; the solves apply
// each panel's row map just before that panel's columns, which is equivalent.
//
// Solve and refinement (GesvRefineKernel, persistent, one workgroup per
// system): x0 from the factors, then until every right-hand side meets
// |b - A x|_inf / (|A|_inf |x|_inf + |b|_inf) <= tol, or the refinement stops
// contracting, or maxIter corrections have been applied: r = b - A x in FP64
// (A read in FP32, x kept in FP64 in LDS), d = LU \ r in FP32, x += d; a
// correction that increases the backward error is undone.  The triangular solves run blockwise on the FP32
// matrix instruction: each 32-row block is multiplied by its stored diagonal
// inverse, the rows below (or above) are updated by the block's L21 (U12).
//
// Fallback: with a BF16 mode, a system whose refinement does not reach tol
// (the backward error fails to halve in a correction, maxIter is exhausted,
// a zero pivot or a NaN appears) is marked in the workspace; two more
// launches re-factor the marked systems with FP32 updates and solve them
// again (every other system is skipped by a flag read).
#include <algorithm>
#include <vector>

#include "cann_ops_blas.h"
#include "device_common.h"
#include "internal.h"
#include "lu_blocked.h"
#include "tuning.h"

namespace aclblas {
namespace {

using dev::kWave;
using lub::f32x4;
using lub::FactorShared;
using lub::kMaxN;
using lub::kNb;
using lub::kThreads;
using lub::kWaves;
using lub::PanelView;

constexpr int kMaxRhs = 16;
constexpr int kInvFloats = kNb * kNb;    // per panel: inv(L11), 32 x 32 column-major

struct MixedSystem {
    const float* A;
    const float* B;
    float* X;
    float* LU;       // workspace: n x n working copy, leading dimension n
    float* dinv;     // workspace: per panel inv(L11)
    int* rowmap;     // workspace: per-panel row maps, panel p at p*n - 16*p*(p-1)
    int* state;      // workspace: [0] zero-pivot column + 1 or 0; [1] 1 = solve again with FP32 updates
    int n, nrhs, lda, ldb, ldx, index;
};

__device__ __forceinline__ int RowMapOffset(int n, int p) { return p * n - 16 * p * (p - 1); }

template <int P>
__global__ __launch_bounds__(kThreads, 3) void GesvFactorKernel(const MixedSystem* __restrict__ systems, int count,
                                                            int onlyMarked)
{
    __shared__ __attribute__((aligned(16))) FactorShared<P> sh;
    __shared__ int failed;
    const int t = threadIdx.x;
    for (int idx = blockIdx.x; idx < count; idx += gridDim.x) {
        const MixedSystem s = systems[idx];
        if (onlyMarked != 0 && dev::G(s.state)[1] == 0) {
            continue;
        }
        const int n = s.n;
        if (t == 0) {
            failed = 0;
        }
        // the first panel and its sweep read A and write the working copy
        for (int k0 = 0, p = 0; k0 < n; k0 += kNb, ++p) {
            const PanelView view = k0 == 0 ? PanelView{s.LU, n, n, s.A, s.lda} : PanelView{s.LU, n, n, s.LU, n};
            int wl = 0;
            const int z = lub::FactorPanel<P, true, false>(view, k0, s.rowmap + RowMapOffset(n, p), sh, &wl);
            if (z != 0) {
                if (t == 0) {
                    failed = z;
                }
                break;
            }
            lub::PanelInverses<P>(view, k0, wl, s.dinv + static_cast<size_t>(p) * kInvFloats, sh);
            __syncthreads();
            lub::SweepTrailing<P>(view, k0, sh);  // panels before the last are full
            __syncthreads();
        }
        if (t == 0) {
            dev::G(s.state)[0] = failed;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Solve and refinement
// ---------------------------------------------------------------------------
struct RefineShared {
    alignas(16) float bs[kMaxN][kMaxRhs];     // right-hand sides / solutions (FP32); rows >= n stay 0
    alignas(16) double x[kMaxN][kMaxRhs];     // current solution (FP64); rows >= n and columns >= nrhs stay 0
    double red[3][kWaves][kMaxRhs];           // per-wave max |r|, |x|, |b|
    double anorm[kWaves];                     // per-wave max row sum of |A|
    double berr[kMaxRhs];
};

__device__ __forceinline__ double MaxD(double a, double b) { return b > a ? b : a; }

// ok ? p[idx] : 0 with an unconditional load (of p[0] when !ok, which is in
// bounds wherever this is used): the compiler can then issue the loads of a
// sequence together, where a load under a lane-dependent condition is issued
// inside a branch and waited for there.
// The pointer, made opaque to the optimiser at this point: addresses derived
// from it are computed here instead of being hoisted out of the enclosing
// refinement loop, where they would stay live across the solves and spill.
template <typename T>
__device__ __forceinline__ T* Opaque(T* p)
{
    asm volatile("" : "+s"(p));
    return p;
}

// Keeps the compiler's scheduler from moving instructions across this point:
// placed after a group of loads, it keeps them issued together ahead of their
// uses (under register pressure the scheduler otherwise pairs each load with
// its first use and waits for it there).
__device__ __forceinline__ void LoadsBeforeUses() { __builtin_amdgcn_sched_barrier(0); }

template <typename T, typename P>
__device__ __forceinline__ T LoadOr0(P p, size_t idx, bool ok)
{
    const T v = p[ok ? idx : 0];
    return ok ? v : T(0);
}

// p[idx] when ok, else p[0] (in bounds wherever this is used; the caller
// replaces it by 0 after LoadsBeforeUses, so the loads do not wait one by one
// for a select).
template <typename P>
__device__ __forceinline__ float LoadAt(P p, size_t idx, bool ok)
{
    return p[ok ? idx : 0];
}

// bs <- LU^-1 bs: per panel the row map, inv(L11) on the diagonal block and
// the block's L21 below it; then from the bottom each diagonal block of U by
// substitution and its U column block above it.  FP32 matrix instruction: a
// lane holds rows 4 (lane >> 4) + j of a 16-row block and right-hand side
// lane & 15.
__device__ __forceinline__ void FactoredSolve(const MixedSystem& s, RefineShared& sh)
{
    const int n = s.n;
    const int t = threadIdx.x;
    const int lane = dev::LaneId();
    const int wave = t / kWave;
    const int r16 = lane & 15;
    const int grp = lane >> 4;
    const auto LU = dev::G(Opaque(s.LU));
    const int blocks = (n + kNb - 1) / kNb;
    for (int p = 0; p < blocks; ++p) {
        const int k0 = p * kNb;
        const int rows = n - k0;
        {
            // row interchanges of this panel: gather through the map in registers, write back
            const auto map = dev::G(s.rowmap + RowMapOffset(n, p));
            const int src = LoadOr0<int>(map, t, t < rows);
            float g[kMaxRhs];
            if (t < rows) {
                dev::StaticFor<0, kMaxRhs / 4>([&](auto cc) {
                    constexpr int c = decltype(cc)::value;
                    const float4 v = *reinterpret_cast<const float4*>(&sh.bs[k0 + src][4 * c]);
                    g[4 * c] = v.x;
                    g[4 * c + 1] = v.y;
                    g[4 * c + 2] = v.z;
                    g[4 * c + 3] = v.w;
                });
            }
            __syncthreads();
            if (t < rows) {
                dev::StaticFor<0, kMaxRhs / 4>([&](auto cc) {
                    constexpr int c = decltype(cc)::value;
                    *reinterpret_cast<float4*>(&sh.bs[k0 + t][4 * c]) =
                        make_float4(g[4 * c], g[4 * c + 1], g[4 * c + 2], g[4 * c + 3]);
                });
            }
        }
        __syncthreads();
        if (wave == 0) {
            // diagonal block: y = inv(L11) b (rows past n are zero in b, identity in the inverse)
            const auto Li = dev::G(s.dinv + static_cast<size_t>(p) * kInvFloats);
            // operands first, then the MFMA chains (the loads issue together)
            float li[2][kNb / 4];
            dev::StaticFor<0, 2>([&](auto bc) {
                constexpr int b = decltype(bc)::value;
                dev::StaticFor<0, kNb / 4>([&](auto sc) {
                    constexpr int st = decltype(sc)::value;
                    li[b][st] = Li[(16 * b + r16) + kNb * (4 * st + grp)];
                });
            });
            LoadsBeforeUses();
            f32x4 y[2];
            dev::StaticFor<0, 2>([&](auto bc) {
                constexpr int b = decltype(bc)::value;
                f32x4 v = {0.0f, 0.0f, 0.0f, 0.0f};
                dev::StaticFor<0, kNb / 4>([&](auto sc) {
                    constexpr int st = decltype(sc)::value;
                    v = __builtin_amdgcn_mfma_f32_16x16x4f32(li[b][st], sh.bs[k0 + 4 * st + grp][r16], v, 0, 0, 0);
                });
                y[b] = v;
            });
            dev::StaticFor<0, 2>([&](auto bc) {
                constexpr int b = decltype(bc)::value;
                dev::StaticFor<0, 4>([&](auto jc) {
                    constexpr int j = decltype(jc)::value;
                    sh.bs[k0 + 16 * b + 4 * grp + j][r16] = y[b][j];
                });
            });
        }
        __syncthreads();
        // rows below the block: b -= L21 y
        const int below = rows - kNb;
        if (below > 0) {
            const int nb16 = (below + 15) / 16;
            for (int blk = wave; blk < nb16; blk += kWaves) {
                const int r0 = k0 + kNb + 16 * blk;
                const int row = r0 + r16;
                float a[kNb / 4];
                dev::StaticFor<0, kNb / 4>([&](auto sc) {
                    constexpr int st = decltype(sc)::value;
                    a[st] = LoadAt(LU, row + static_cast<size_t>(k0 + 4 * st + grp) * n, row < n);
                });
                LoadsBeforeUses();
                f32x4 v = {0.0f, 0.0f, 0.0f, 0.0f};
                dev::StaticFor<0, kNb / 4>([&](auto sc) {
                    constexpr int st = decltype(sc)::value;
                    v = __builtin_amdgcn_mfma_f32_16x16x4f32(row < n ? a[st] : 0.0f, sh.bs[k0 + 4 * st + grp][r16], v,
                                                             0, 0, 0);
                });
                dev::StaticFor<0, 4>([&](auto jc) {
                    constexpr int j = decltype(jc)::value;
                    const int rr = r0 + 4 * grp + j;
                    if (rr < n) {
                        sh.bs[rr][r16] -= v[j];
                    }
                });
            }
        }
        __syncthreads();
    }
    for (int bi = blocks - 1; bi >= 0; --bi) {
        const int r0 = bi * kNb;
        const int w = min(kNb, n - r0);
        {
            // diagonal block by substitution (U_bb can be ill conditioned: no
            // inverse here): lane q holds row q of U_bb, wave v the right-hand
            // sides v, v + 4, ...
            const bool valid = lane < w;
            float u[kNb];
            float x[kMaxRhs / kWaves];
            dev::StaticFor<0, kNb>([&](auto jc) {
                constexpr int j = decltype(jc)::value;
                u[j] = LoadAt(LU, (r0 + lane) + static_cast<size_t>(r0 + j) * n, valid && j >= lane && j < w);
            });
            dev::StaticFor<0, kMaxRhs / kWaves>([&](auto cc) {
                constexpr int ci = decltype(cc)::value;
                x[ci] = valid ? sh.bs[r0 + lane][ci * kWaves + wave] : 0.0f;
            });
            LoadsBeforeUses();
            dev::StaticFor<0, kNb>([&](auto jc) {
                constexpr int j = decltype(jc)::value;
                u[j] = (valid && j >= lane && j < w) ? u[j] : 0.0f;
            });
            dev::StaticFor<0, kNb>([&](auto jc) {
                constexpr int j = kNb - 1 - decltype(jc)::value;
                if (j < w) {
                    dev::StaticFor<0, kMaxRhs / kWaves>([&](auto cc) {
                        constexpr int ci = decltype(cc)::value;
                        if (ci * kWaves + wave < s.nrhs) {
                            if (lane == j) {
                                x[ci] = x[ci] / u[j];
                            }
                            const float xj = dev::ReadLane(x[ci], j);
                            if (lane < j) {
                                x[ci] = fmaf(-u[j], xj, x[ci]);
                            }
                        }
                    });
                }
            });
            if (valid) {
                dev::StaticFor<0, kMaxRhs / kWaves>([&](auto cc) {
                    constexpr int ci = decltype(cc)::value;
                    sh.bs[r0 + lane][ci * kWaves + wave] = x[ci];
                });
            }
        }
        __syncthreads();
        // rows above the block: b -= U[0:r0, block] y
        const int nb16 = r0 / 16;
        for (int blk = wave; blk < nb16; blk += kWaves) {
            const int row = 16 * blk + r16;
            float a[kNb / 4];
            dev::StaticFor<0, kNb / 4>([&](auto sc) {
                constexpr int st = decltype(sc)::value;
                const int k = 4 * st + grp;
                a[st] = LoadAt(LU, row + static_cast<size_t>(r0 + k) * n, r0 + k < n);
            });
            LoadsBeforeUses();
            f32x4 v = {0.0f, 0.0f, 0.0f, 0.0f};
            dev::StaticFor<0, kNb / 4>([&](auto sc) {
                constexpr int st = decltype(sc)::value;
                v = __builtin_amdgcn_mfma_f32_16x16x4f32(r0 + 4 * st + grp < n ? a[st] : 0.0f,
                                                         sh.bs[r0 + 4 * st + grp][r16], v, 0, 0, 0);
            });
            dev::StaticFor<0, 4>([&](auto jc) {
                constexpr int j = decltype(jc)::value;
                sh.bs[16 * blk + 4 * grp + j][r16] -= v[j];
            });
        }
        __syncthreads();
    }
}

__device__ __forceinline__ double WaveMaxD(double v)
{
    for (int off = 32; off >= 1; off >>= 1) {
        v = MaxD(v, __shfl_xor(v, off, kWave));
    }
    return v;
}

// Thread t holds rows 4 (t / 4) .. +3 and right-hand sides 4 (t % 4) .. +3 of
// the residual.
constexpr int kResRows = 4;
constexpr int kResRhs = 4;
typedef double ResidualTile[kResRows][kResRhs];

// Lane I of each aligned group of four lanes, to all four (DPP quad_perm
// [I, I, I, I]; every lane is active where this is used).
template <int I>
__device__ __forceinline__ float QuadBroadcast(float v)
{
    return __uint_as_float(dev::Dpp<I * 0x55>(__float_as_uint(v)));
}

__device__ __forceinline__ double MaxOverQuadLanes(double v)
{
    // lanes with the same t % 4 (xor over lane bits 2..5)
    for (int off = 4; off < kWave; off <<= 1) {
        v = MaxD(v, __shfl_xor(v, off, kWave));
    }
    return v;
}

// r = b - A x in FP64 (A read in FP32, x from LDS) into the caller's registers
// (ResidualTile layout); the worst normwise backward error over the
// right-hand sides is returned (uniform).  With withNorm, |A|_inf is computed
// into sh.anorm first.
// Per column of A a lane loads one value (row 4 (t / 4) + t % 4) and its group
// of four lanes shares the four rows by DPP; each lane then reads 32 bytes of
// x for 16 FMAs: a quarter of the LDS bytes per FMA of one row per thread with
// all sixteen right-hand sides, whose broadcast reads bound the LDS.
template <bool NORM>
__device__ __forceinline__ void ResidualColumns(const MixedSystem& s, const RefineShared& sh, int row0, int q,
                                                ResidualTile& r, double (&rs)[kResRows])
{
    const int n = s.n;
    const int t = threadIdx.x;
    const int myRow = row0 + (t & 3);
    const bool rowOk = myRow < n;
    const auto A = dev::G(Opaque(s.A));
    const size_t lda = static_cast<size_t>(s.lda);
    // eight columns per step, the next eight loaded while these are used
    constexpr int kCols = 8;
    float cur[kCols], nxt[kCols];
    // a column's base address is uniform (scalar registers), the row a 32-bit
    // lane offset; columns past n and rows past n read element (0, 0) and are
    // replaced by 0 at their use
    const unsigned rowOff = rowOk ? static_cast<unsigned>(myRow) : 0u;
    dev::StaticFor<0, kCols>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        cur[j] = (A + static_cast<size_t>(j < n ? j : 0) * lda)[rowOff];
    });
    for (int j0 = 0; j0 < n; j0 += kCols) {
        dev::StaticFor<0, kCols>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            const int col = j0 + kCols + j;
            nxt[j] = (A + static_cast<size_t>(col < n ? col : 0) * lda)[rowOff];
        });
        LoadsBeforeUses();
        dev::StaticFor<0, kCols>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            // x rows past n are zero, so columns past n add nothing (j0 + j < kMaxN: n <= kMaxN, kCols | kMaxN)
            const double2 x01 = *reinterpret_cast<const double2*>(&sh.x[j0 + j][4 * q]);
            const double2 x23 = *reinterpret_cast<const double2*>(&sh.x[j0 + j][4 * q + 2]);
            const double xv[4] = {x01.x, x01.y, x23.x, x23.y};
            const float mine = (rowOk && j0 + j < n) ? cur[j] : 0.0f;
            double a[kResRows];
            a[0] = static_cast<double>(QuadBroadcast<0>(mine));
            a[1] = static_cast<double>(QuadBroadcast<1>(mine));
            a[2] = static_cast<double>(QuadBroadcast<2>(mine));
            a[3] = static_cast<double>(QuadBroadcast<3>(mine));
            dev::StaticFor<0, kResRows>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                if constexpr (NORM) {
                    rs[i] += fabs(a[i]);
                }
                dev::StaticFor<0, kResRhs>([&](auto ec) {
                    constexpr int e = decltype(ec)::value;
                    r[i][e] = fma(-a[i], xv[e], r[i][e]);
                });
            });
        });
        dev::StaticFor<0, kCols>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            cur[j] = nxt[j];
        });
    }
}

__device__ __forceinline__ double Residual(const MixedSystem& s, RefineShared& sh, bool withNorm, ResidualTile& r)
{
    const int n = s.n;
    const int nrhs = s.nrhs;
    const int t = threadIdx.x;
    const int lane = dev::LaneId();
    const int wave = t / kWave;
    const int q = t & 3;
    const int row0 = (t >> 2) * kResRows;
    const int uwave = dev::Uniform(wave);
    const auto Bm = dev::G(Opaque(s.B));
    const size_t ldb = static_cast<size_t>(s.ldb);
    float bf[kResRows][kResRhs];
    dev::StaticFor<0, kResRhs>([&](auto ec) {
        constexpr int e = decltype(ec)::value;
        const int c = kResRhs * q + e;
        dev::StaticFor<0, kResRows>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            const int row = row0 + i;
            bf[i][e] = LoadAt(Bm, row + static_cast<size_t>(c) * ldb, row < n && c < nrhs);
        });
    });
    LoadsBeforeUses();
    double bmax[kResRhs];
    dev::StaticFor<0, kResRhs>([&](auto ec) {
        constexpr int e = decltype(ec)::value;
        bmax[e] = 0.0;
        dev::StaticFor<0, kResRows>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            r[i][e] = (row0 + i < n && kResRhs * q + e < nrhs) ? static_cast<double>(bf[i][e]) : 0.0;
            bmax[e] = MaxD(bmax[e], fabs(r[i][e]));
        });
    });
    double rs[kResRows] = {0.0, 0.0, 0.0, 0.0};
    if (uwave * kWave < n) {  // the wave's rows 64 wave .. +63 reach into the system
        if (withNorm) {
            ResidualColumns<true>(s, sh, row0, q, r, rs);
        } else {
            ResidualColumns<false>(s, sh, row0, q, r, rs);
        }
    }
    if (withNorm) {
        double m = MaxD(MaxD(rs[0], rs[1]), MaxD(rs[2], rs[3]));
        m = WaveMaxD(m);
        if (lane == 0) {
            sh.anorm[wave] = m;
        }
    }
    dev::StaticFor<0, kResRhs>([&](auto ec) {
        constexpr int e = decltype(ec)::value;
        const int c = kResRhs * q + e;
        double rv = 0.0, xv = 0.0;
        dev::StaticFor<0, kResRows>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            rv = MaxD(rv, fabs(r[i][e]));
            xv = MaxD(xv, row0 + i < n ? fabs(sh.x[row0 + i][c]) : 0.0);
        });
        rv = MaxOverQuadLanes(rv);
        xv = MaxOverQuadLanes(xv);
        const double bv = MaxOverQuadLanes(bmax[e]);
        if (lane < 4) {
            sh.red[0][wave][c] = rv;
            sh.red[1][wave][c] = xv;
            sh.red[2][wave][c] = bv;
        }
    });
    __syncthreads();
    if (t < nrhs) {
        double an = 0.0, rn = 0.0, xn = 0.0, bn = 0.0;
        for (int w = 0; w < kWaves; ++w) {
            an = MaxD(an, sh.anorm[w]);
            rn = MaxD(rn, sh.red[0][w][t]);
            xn = MaxD(xn, sh.red[1][w][t]);
            bn = MaxD(bn, sh.red[2][w][t]);
        }
        const double den = an * xn + bn;
        sh.berr[t] = den > 0.0 ? rn / den : (rn > 0.0 ? __builtin_inf() : 0.0);
    }
    __syncthreads();
    double worst = 0.0;
    for (int c = 0; c < nrhs; ++c) {
        const double be = sh.berr[c];
        worst = (be > worst || !(be == be)) ? be : worst;
    }
    __syncthreads();
    return worst;
}

// pass 0: every system; with mayMark, a system that does not reach tol is
// marked (state[1] = 1) and left for the FP32 pass.  pass 1 (onlyMarked):
// the marked systems, final.
__global__ __launch_bounds__(kThreads, 3) void GesvRefineKernel(const MixedSystem* __restrict__ systems, int count,
                                                            double tol, int maxIter, int mayMark, int onlyMarked,
                                                            int* __restrict__ iterOut, float* __restrict__ berrOut,
                                                            int* __restrict__ infoOut, int* __restrict__ fallbackOut)
{
    __shared__ __attribute__((aligned(16))) RefineShared sh;
    const int t = threadIdx.x;
    for (int idx = blockIdx.x; idx < count; idx += gridDim.x) {
        const MixedSystem s = systems[idx];
        if (onlyMarked != 0 && dev::G(s.state)[1] == 0) {
            continue;
        }
        const int n = s.n;
        const int nrhs = s.nrhs;
        const int status = dev::G(s.state)[0];
        if (status != 0 && mayMark != 0) {
            if (t == 0) {
                dev::G(s.state)[1] = 1;
            }
            continue;
        }
        if (status != 0) {
            for (int c = 0; c < nrhs; ++c) {
                for (int i = t; i < n; i += kThreads) {
                    dev::G(s.X)[i + static_cast<size_t>(c) * s.ldx] = __builtin_nanf("");
                }
            }
            if (t == 0) {
                if (iterOut != nullptr) iterOut[s.index] = 0;
                if (berrOut != nullptr) berrOut[s.index] = __builtin_inff();
                if (infoOut != nullptr) infoOut[s.index] = status;
                if (fallbackOut != nullptr) fallbackOut[s.index] = onlyMarked;
                dev::G(s.state)[1] = 0;
            }
            continue;
        }
        // x0 = LU \ b (rows >= n and columns >= nrhs of bs and x are zero)
        {
            float b0[kMaxRhs];
            dev::StaticFor<0, kMaxRhs>([&](auto cc) {
                constexpr int c = decltype(cc)::value;
                b0[c] = LoadAt(dev::G(s.B), t + static_cast<size_t>(c) * s.ldb, t < n && c < nrhs);
            });
            LoadsBeforeUses();
            dev::StaticFor<0, kMaxRhs>([&](auto cc) {
                constexpr int c = decltype(cc)::value;
                sh.bs[t][c] = (t < n && c < nrhs) ? b0[c] : 0.0f;
            });
        }
        __syncthreads();
        FactoredSolve(s, sh);
        dev::StaticFor<0, kMaxRhs>([&](auto cc) {
            constexpr int c = decltype(cc)::value;
            sh.x[t][c] = (t < n && c < nrhs) ? static_cast<double>(sh.bs[t][c]) : 0.0;
        });
        __syncthreads();
        int iter = 0;
        double worst = 0.0, prev = 0.0;
        bool converged = false;
        ResidualTile r;
        for (;;) {
            worst = Residual(s, sh, iter == 0, r);
            if (iter > 0 && !(worst <= prev)) {
                // the last correction made things worse: take it back (bs still holds it)
                if (t < n) {
                    dev::StaticFor<0, kMaxRhs>([&](auto cc) {
                        constexpr int c = decltype(cc)::value;
                        if (c < nrhs) {
                            sh.x[t][c] -= static_cast<double>(sh.bs[t][c]);
                        }
                    });
                }
                worst = prev;
                __syncthreads();
                break;
            }
            converged = worst <= tol;
            const bool stalled = !(worst == worst) || iter >= maxIter || (iter > 0 && worst > 0.5 * prev);
            if (converged || stalled) {
                break;
            }
            // d = LU \ r (FP32), x += d (the tiles of r cover bs: rows >= n and columns >= nrhs get 0)
            dev::StaticFor<0, kResRows>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                *reinterpret_cast<float4*>(&sh.bs[(t >> 2) * kResRows + i][kResRhs * (t & 3)]) =
                    make_float4(static_cast<float>(r[i][0]), static_cast<float>(r[i][1]),
                                static_cast<float>(r[i][2]), static_cast<float>(r[i][3]));
            });
            __syncthreads();
            FactoredSolve(s, sh);
            dev::StaticFor<0, kMaxRhs>([&](auto cc) {
                constexpr int c = decltype(cc)::value;
                if (t < n && c < nrhs) {
                    sh.x[t][c] += static_cast<double>(sh.bs[t][c]);
                }
            });
            __syncthreads();
            prev = worst;
            ++iter;
        }
        if (!converged && mayMark != 0) {
            if (t == 0) {
                dev::G(s.state)[1] = 1;
            }
            __syncthreads();
            continue;
        }
        if (t < n) {
            for (int c = 0; c < nrhs; ++c) {
                dev::G(s.X)[t + static_cast<size_t>(c) * s.ldx] = static_cast<float>(sh.x[t][c]);
            }
        }
        if (t == 0) {
            if (iterOut != nullptr) iterOut[s.index] = iter;
            if (berrOut != nullptr) berrOut[s.index] = static_cast<float>(worst);
            if (infoOut != nullptr) infoOut[s.index] = 0;
            if (fallbackOut != nullptr) fallbackOut[s.index] = onlyMarked;
            dev::G(s.state)[1] = 0;
        }
        __syncthreads();
    }
}

}  // namespace
}  // namespace aclblas

namespace aclblas {
namespace {

constexpr size_t kAlign = 256;

size_t AlignUp(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

size_t RowMapInts(int n)
{
    size_t total = 0;
    for (int k0 = 0; k0 < n; k0 += kNb) {
        total += static_cast<size_t>(n - k0);
    }
    return total;
}

aclblasStatus_t ValidateMixed(int batch, const int* nA, const int* rA, const float* const* A, const int* ldaA,
                              const float* const* B, const int* ldbA, float* const* X, const int* ldxA, double tol,
                              int maxIter, int update)
{
    if (batch < 0 || !(tol >= 0.0) || maxIter < 0) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (update < ACLBLAS_MIXED_UPDATE_FP32 || update > ACLBLAS_MIXED_UPDATE_BF16X3) {
        return ACLBLAS_STATUS_INVALID_ENUM;
    }
    if (batch == 0) {
        return ACLBLAS_STATUS_SUCCESS;
    }
    if (nA == nullptr || rA == nullptr || A == nullptr || ldaA == nullptr || B == nullptr || ldbA == nullptr ||
        X == nullptr || ldxA == nullptr) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    for (int i = 0; i < batch; ++i) {
        const int n = nA[i];
        const int r = rA[i];
        if (n < 0 || r < 0) {
            return ACLBLAS_STATUS_INVALID_VALUE;
        }
        if (ldaA[i] < MaxInt(1, n) || ldbA[i] < MaxInt(1, n) || ldxA[i] < MaxInt(1, n)) {
            return ACLBLAS_STATUS_INVALID_VALUE;
        }
        if (n > kMaxN || r > kMaxRhs) {
            return ACLBLAS_STATUS_NOT_SUPPORTED;
        }
        if (n > 0 && (A[i] == nullptr || (r > 0 && (B[i] == nullptr || X[i] == nullptr)))) {
            return ACLBLAS_STATUS_INVALID_VALUE;
        }
    }
    return ACLBLAS_STATUS_SUCCESS;
}

void LaunchFactor(int update, unsigned grid, hipStream_t stream, const MixedSystem* sys, int count, int onlyMarked)
{
    switch (update) {
    case ACLBLAS_MIXED_UPDATE_BF16:
        hipLaunchKernelGGL(GesvFactorKernel<1>, dim3(grid), dim3(kThreads), 0, stream, sys, count, onlyMarked);
        break;
    case ACLBLAS_MIXED_UPDATE_BF16X2:
        hipLaunchKernelGGL(GesvFactorKernel<2>, dim3(grid), dim3(kThreads), 0, stream, sys, count, onlyMarked);
        break;
    case ACLBLAS_MIXED_UPDATE_BF16X3:
        hipLaunchKernelGGL(GesvFactorKernel<3>, dim3(grid), dim3(kThreads), 0, stream, sys, count, onlyMarked);
        break;
    default:
        hipLaunchKernelGGL(GesvFactorKernel<0>, dim3(grid), dim3(kThreads), 0, stream, sys, count, onlyMarked);
        break;
    }
}

}  // namespace
}  // namespace aclblas

extern "C" aclblasStatus_t aclblasSgesvBatchedMixedEx(
    aclblasHandle_t handle, int batchCount, const int* nArray, const int* nrhsArray, const float* const Aarray[],
    const int* ldaArray, const float* const Barray[], const int* ldbArray, float* const Xarray[],
    const int* ldxArray, double tol, int maxIter, aclblasMixedUpdate_t update, int* iterArray, float* berrArray,
    int* infoArray, int* fallbackArray)
{
    using namespace aclblas;
    if (handle == nullptr) {
        return ACLBLAS_STATUS_HANDLE_IS_NULLPTR;
    }
    const int mode = EnumBits(update);
    aclblasStatus_t st = ValidateMixed(batchCount, nArray, nrhsArray, Aarray, ldaArray, Barray, ldbArray, Xarray,
                                       ldxArray, tol, maxIter, mode);
    if (st != ACLBLAS_STATUS_SUCCESS) {
        LogError("aclblasSgesvBatchedMixedEx", "invalid argument (status %d)", static_cast<int>(st));
        return st;
    }
    if (batchCount == 0) {
        return ACLBLAS_STATUS_SUCCESS;
    }
    if (!HandleHasDevice(handle)) {
        return ACLBLAS_STATUS_EXECUTION_FAILED;
    }
    // largest systems first; workspace: table, then per system LU, inverses, row maps, state
    std::vector<int> order(static_cast<size_t>(batchCount));
    for (int i = 0; i < batchCount; ++i) {
        order[static_cast<size_t>(i)] = i;
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return nArray[a] > nArray[b]; });
    const size_t tableBytes = AlignUp(sizeof(MixedSystem) * static_cast<size_t>(batchCount));
    std::vector<size_t> offs(static_cast<size_t>(batchCount));
    size_t total = tableBytes;
    for (int i = 0; i < batchCount; ++i) {
        const int n = nArray[i];
        const size_t panels = static_cast<size_t>((n + kNb - 1) / kNb);
        offs[static_cast<size_t>(i)] = total;
        total += AlignUp(sizeof(float) * static_cast<size_t>(n) * n) + AlignUp(sizeof(float) * kInvFloats * panels) +
                 AlignUp(sizeof(int) * RowMapInts(n)) + kAlign;
    }
    st = EnsureWorkspace(handle, total);
    if (st != ACLBLAS_STATUS_SUCCESS) {
        return st;
    }
    char* base = static_cast<char*>(handle->workspace);
    std::vector<MixedSystem> table(static_cast<size_t>(batchCount));
    for (int q = 0; q < batchCount; ++q) {
        const int i = order[static_cast<size_t>(q)];
        const int n = nArray[i];
        const size_t panels = static_cast<size_t>((n + kNb - 1) / kNb);
        MixedSystem& s = table[static_cast<size_t>(q)];
        char* p = base + offs[static_cast<size_t>(i)];
        s.A = Aarray[i];
        s.B = Barray[i];
        s.X = Xarray[i];
        s.LU = reinterpret_cast<float*>(p);
        p += AlignUp(sizeof(float) * static_cast<size_t>(n) * n);
        s.dinv = reinterpret_cast<float*>(p);
        p += AlignUp(sizeof(float) * kInvFloats * panels);
        s.rowmap = reinterpret_cast<int*>(p);
        p += AlignUp(sizeof(int) * RowMapInts(n));
        s.state = reinterpret_cast<int*>(p);
        s.n = n;
        s.nrhs = nrhsArray[i];
        s.lda = ldaArray[i];
        s.ldb = ldbArray[i];
        s.ldx = ldxArray[i];
        s.index = i;
    }
    void* dev = nullptr;
    st = StageToWorkspace(handle, table.data(), sizeof(MixedSystem) * table.size(), &dev);
    if (st != ACLBLAS_STATUS_SUCCESS) {
        return st;
    }
    const int cus = handle->cu_count > 0 ? handle->cu_count : 256;
    // persistent grids, workgroups per CU from the tuning table, by the update's
    // class: 0 for the one- and two-plane BF16 updates (their factor kernel's
    // LDS lets three workgroups share a CU), 1 for FP32 and three planes (two);
    // built-in: three or two factor workgroups, three solve workgroups
    const int cls = (mode == ACLBLAS_MIXED_UPDATE_FP32 || mode == ACLBLAS_MIXED_UPDATE_BF16X3) ? 1 : 0;
    const tune::Choice c = tune::Get(handle, tune::kOpMixed, tune::MixedBucket(cls, batchCount, cus),
                                     tune::Choice{cls == 0 ? 3 : 2, 3});
    const unsigned gridF = static_cast<unsigned>(std::min(batchCount, c.a * cus));
    const unsigned gridS = static_cast<unsigned>(std::min(batchCount, c.b * cus));
    const MixedSystem* sys = static_cast<const MixedSystem*>(dev);
    hipStream_t stream = handle->stream;
    const int mayMark = mode != ACLBLAS_MIXED_UPDATE_FP32 ? 1 : 0;
    MarkLaunch(handle, 0);
    LaunchFactor(mode, gridF, stream, sys, batchCount, 0);
    MarkLaunch(handle, 1);
    MarkLaunch(handle, 2);
    hipLaunchKernelGGL(GesvRefineKernel, dim3(gridS), dim3(kThreads), 0, stream, sys, batchCount, tol, maxIter,
                       mayMark, 0, iterArray, berrArray, infoArray, fallbackArray);
    MarkLaunch(handle, 3);
    if (mayMark != 0) {
        MarkLaunch(handle, 4);
        LaunchFactor(ACLBLAS_MIXED_UPDATE_FP32, gridF, stream, sys, batchCount, 1);
        MarkLaunch(handle, 5);
        MarkLaunch(handle, 6);
        hipLaunchKernelGGL(GesvRefineKernel, dim3(gridS), dim3(kThreads), 0, stream, sys, batchCount, tol, maxIter,
                           0, 1, iterArray, berrArray, infoArray, fallbackArray);
        MarkLaunch(handle, 7);
    }
    return LaunchStatus("aclblasSgesvBatchedMixedEx");
}

extern "C" aclblasStatus_t aclblasSgesvBatchedMixed(
    aclblasHandle_t handle, int batchCount, const int* nArray, const int* nrhsArray, const float* const Aarray[],
    const int* ldaArray, const float* const Barray[], const int* ldbArray, float* const Xarray[],
    const int* ldxArray, double tol, int maxIter, int* iterArray, float* berrArray, int* infoArray)
{
    return aclblasSgesvBatchedMixedEx(handle, batchCount, nArray, nrhsArray, Aarray, ldaArray, Barray, ldbArray,
                                      Xarray, ldxArray, tol, maxIter, ACLBLAS_MIXED_UPDATE_BF16X2, iterArray,
                                      berrArray, infoArray, nullptr);
}
