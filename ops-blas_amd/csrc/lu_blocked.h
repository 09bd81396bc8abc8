// Blocked right-looking LU building blocks for systems of order <= 256 held
// in global memory, one 256-thread workgroup per system: the panel
// factorisation with partial pivoting, the inverse of the panel's unit lower
// block, and the trailing-matrix sweep whose update runs on matrix cores in
// FP32 or in split BF16 planes.  Used by the mixed-precision solve
// (gesv_mixed.hip) and by aclblasSgetrfBatched for 64 < n <= 256
// (getrf_blocked.hip).
//
// Pivoting follows the reference's getrf semantics (largest |a| in the column
// among the rows not yet used, exact ties to the lowest current position,
// NaN never wins, stop at an exact zero pivot; reference
// blas/getrf_batched/arch35/sgetrf_batched_kernel.cpp:52-150); rows never move
// during a panel: every row carries its position and the panel is written back
// in pivoted order, LAPACK's row interchanges being the position swaps.
#pragma once

#include <hip/hip_runtime.h>

#include "device_common.h"

namespace aclblas {
namespace lub {

using dev::kWave;

constexpr int kMaxN = 256;
constexpr int kNb = 32;                  // panel width
constexpr int kThreads = 256;            // four waves
constexpr int kWaves = kThreads / kWave;
constexpr int kTileN = 16;               // sweep tile width
constexpr int kMaxTiles = kMaxN / 16;    // 16-row accumulator tiles per column tile
constexpr int kLd = kNb + 4;             // row stride (floats) of 32-wide FP32 LDS blocks: 16-byte rows

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// The matrix being factored: order n, column-major with leading dimension ld.
// Loads read the current values from src (leading dimension srcLd), stores go
// to LU: src is LU itself except on a first panel that reads the caller's
// matrix and writes the working copy (the panel and its sweep together write
// every element, so no separate copy pass is needed).
struct PanelView {
    float* LU;
    int ld;
    int n;
    const float* src;
    int srcLd;
};

__device__ __forceinline__ uint16_t Bf16Rne(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7FFFFFFFu) > 0x7F800000u) {
        return static_cast<uint16_t>((u >> 16) | 0x40u);
    }
    u += 0x7FFFu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}
__device__ __forceinline__ float Bf16ToF(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

// x = sum of the P planes: each plane is the BF16 rounding of what the planes
// before it leave (three planes hold every finite FP32 value of normal range
// exactly).
template <int P>
__device__ __forceinline__ void SplitPlanes(float x, uint16_t (&h)[P])
{
    float r = x;
    dev::StaticFor<0, P>([&](auto pc) {
        constexpr int p = decltype(pc)::value;
        h[p] = Bf16Rne(r);
        if constexpr (p + 1 < P) {
            r = r - Bf16ToF(h[p]);
        }
    });
}

// -L21 and U12 of the current panel in the operand format of the update.
template <int P>
struct UpdateOperands {  // P BF16 planes each
    alignas(16) uint16_t l[P][kMaxN - kNb][kNb];          // -L21 [plane][row][k]
    alignas(16) uint16_t u[kWaves][P][kTileN][kNb];       // U12^T per wave [plane][col][k]
};
template <>
struct UpdateOperands<0> {  // FP32
    alignas(16) float l[kMaxN - kNb][kLd];                // -L21 [row][k]
    alignas(16) float u[kWaves][kTileN][kLd];             // U12^T per wave [col][k]
};

template <int P>
struct FactorShared {
    alignas(16) float cand[2][kWaves][kNb];               // each wave's best row, by column parity
    alignas(16) uint64_t cand64[2][kWaves];               // its packed key (PackCand), 0: no candidate
    int rowmap[kMaxN];                                    // panel row map (position -> row, relative)
    int pw[kNb];                                          // pivot position chosen at each panel column
    int zero;                                             // zero-pivot column of a one-wave panel
    alignas(16) float linv[kNb][kLd];                     // inv(L11), row-major
    // L11 lives from the panel's write-back to its inverse, the sweep's
    // per-wave tiles only during the sweep: they share the space
    union {
        alignas(16) float l11[kNb][kLd];                  // panel rows 0..31: L11 below, U11 on/above the diagonal
        alignas(16) float ut[kWaves][kNb][kTileN + 1];    // per-wave top rows of the sweep tile
    };
    UpdateOperands<P> op;
};

// A wave's best candidate as one integer whose maximum over the waves is the
// pivot: |a| key, then the lowest position (0xFFFF - position), then the wave.
// A candidate always packs to more than 0 (position < 0xFFFF).
__device__ __forceinline__ uint64_t PackCand(uint32_t key, int pos, int wave)
{
    return (static_cast<uint64_t>(key) << 32) | (static_cast<uint64_t>(0xFFFF - pos) << 8) |
           static_cast<uint64_t>(wave);
}

__device__ __forceinline__ void WaveSync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Panel factorisation by the whole workgroup: wave v holds rows 64v..64v+63 of
// the panel (one row per lane) in registers.  Per column, each wave finds its
// best candidate (largest |a| key by a DPP reduction; the lanes holding it by
// ballot, and only when several do, the lowest position among them by a
// second reduction), that lane publishes its row, key and position in LDS
// (double-buffered by column parity), one barrier, and every lane reads the
// four waves' keys and positions as two vectors and picks the same winner
// (without PIVOT the row at position j is the only candidate).  Waves with no
// rows left in the panel only take part in the barriers.  Multipliers are
// a / p (DIVIDE) or a * (1 / p).  The panel goes back
// in pivoted row order with its row map (sh.rowmap, and gmap when given),
// L11 / U11 in sh.l11 and -L21 in the update's operand format; sh.pw[j] is the
// position the pivot of column j came from.  At an exact zero pivot in column
// k0 + j the panel stops there: *wl = j (columns factored), the columns right
// of it hold the state after the first j steps, and the -L21 operand columns
// from j on are zero, so a sweep with this panel applies exactly those j steps
// to the trailing columns.  Returns the zero-pivot column + 1, or 0 (*wl =
// width then).  Uniform over the workgroup; ends with a barrier.
template <int P, bool PIVOT, bool DIVIDE>
__device__ __forceinline__ int FactorPanel(const PanelView& pv, int k0, int* gmap, FactorShared<P>& sh, int* wl)
{
    const int n = pv.n;
    const size_t ld = static_cast<size_t>(pv.ld);
    const size_t sld = static_cast<size_t>(pv.srcLd);
    const int lane = dev::LaneId();
    const int wave = dev::Uniform(static_cast<int>(threadIdx.x) / kWave);
    const int rows = n - k0;
    const int w = min(kNb, rows);
    const int r = wave * kWave + lane;
    const bool valid = r < rows;
    const bool active = wave * kWave < rows;  // the wave holds panel rows
    const auto LU = dev::G(pv.LU);
    const auto src = dev::G(pv.src);

    float a[kNb];
    dev::StaticFor<0, kNb>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        a[j] = (valid && j < w) ? src[(k0 + r) + static_cast<size_t>(k0 + j) * sld] : 0.0f;
    });
    if (!active && lane == 0) {
        sh.cand64[0][wave] = 0;
        sh.cand64[1][wave] = 0;
    }
    int pos = r;
    int zeroCol = 0;
    // A panel of at most 64 rows lives in wave 0 alone: the other waves skip
    // the column loop and wave 0 exchanges its pivot rows with wave-level
    // synchronisation instead of workgroup barriers (same pivots, same
    // arithmetic); the zero-pivot column is shared through LDS afterwards.
    const bool solo = rows <= kWave;  // uniform over the workgroup

    // Straight-line per column (no early exit, so the row stays in the same
    // registers): a column past the panel's width or after a zero pivot is
    // skipped as a whole by every wave alike.
    if (!solo || wave == 0) {
    dev::StaticFor<0, kNb>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        constexpr int buf = j & 1;
        const bool live = j < w && zeroCol == 0;  // uniform over the workgroup
        if (live && active) {
            const bool cand = valid && (PIVOT ? pos >= j : pos == j);
            const uint32_t key = cand ? dev::MagKey(a[j]) : 0u;
            const uint32_t best = dev::SegReduce<kWave, true>(key);
            const bool tied = cand && key == best;
            uint64_t tmask = __builtin_amdgcn_ballot_w64(tied);
            if (tmask & (tmask - 1)) {
                // several lanes hold the best key: the lowest position wins
                const uint32_t lowest = dev::SegReduce<kWave, false>(tied ? static_cast<uint32_t>(pos) : 0xFFFFu);
                tmask = __builtin_amdgcn_ballot_w64(tied && static_cast<uint32_t>(pos) == lowest);
            }
            if (tmask == 0) {
                if (lane == 0) {
                    sh.cand64[buf][wave] = 0;  // no candidate in this wave
                }
            } else if (lane == static_cast<int>(__builtin_ctzll(tmask))) {
                dev::StaticFor<0, kNb / 4>([&](auto cc) {
                    constexpr int c = decltype(cc)::value;
                    *reinterpret_cast<float4*>(&sh.cand[buf][wave][4 * c]) =
                        make_float4(a[4 * c], a[4 * c + 1], a[4 * c + 2], a[4 * c + 3]);
                });
                sh.cand64[buf][wave] = PackCand(key, pos, wave);
            }
        }
        if (live) {
            if (solo) {
                WaveSync();
            } else {
                __syncthreads();
            }
        }
        float pivot = 0.0f;
        int pw = 0;
        float u[kNb];
        if (live) {
            // the winner over the waves: the largest packed candidate (wave 0's
            // own for a one-wave panel)
            uint64_t best4 = sh.cand64[buf][0];
            if (!solo) {
                dev::StaticFor<1, kWaves>([&](auto vc) {
                    constexpr int v = decltype(vc)::value;
                    const uint64_t c = sh.cand64[buf][v];
                    best4 = c > best4 ? c : best4;
                });
            }
            const uint32_t lo = static_cast<uint32_t>(best4);
            const int wb = dev::Uniform(static_cast<int>(lo & 0xFFu));
            pw = dev::Uniform(0xFFFF - static_cast<int>((lo >> 8) & 0xFFFFu));
            if (threadIdx.x == 0) {
                sh.pw[j] = pw;
            }
            dev::StaticFor<j / 4, kNb / 4>([&](auto cc) {
                constexpr int c = decltype(cc)::value;
                const float4 v = *reinterpret_cast<const float4*>(&sh.cand[buf][wb][4 * c]);
                u[4 * c] = v.x;
                u[4 * c + 1] = v.y;
                u[4 * c + 2] = v.z;
                u[4 * c + 3] = v.w;
            });
            pivot = u[j];
            if (pivot == 0.0f) {
                zeroCol = k0 + j + 1;
            }
        }
        const bool step = live && pivot != 0.0f;  // uniform
        const bool winner = step && valid && pos == pw && (PIVOT ? pos >= j : pos == j);
        pos = (step && pos == j) ? pw : pos;
        pos = winner ? j : pos;
        if (step && active) {
            const bool upd = valid && pos > j;
            if constexpr (DIVIDE) {
                // the reference's multipliers; rows left alone keep their exact values
                const float l = a[j] / pivot;
                a[j] = upd ? l : a[j];
                dev::StaticFor<j + 1, kNb>([&](auto cc) {
                    constexpr int c = decltype(cc)::value;
                    a[c] = upd ? fmaf(-l, u[c], a[c]) : a[c];
                });
            } else {
                // rows left alone get l = 0: a - 0 * u keeps their value (up to the sign of a zero, and
                // for finite u), which is all the backward-error-checked solve needs
                const float l = upd ? a[j] * (1.0f / pivot) : 0.0f;
                a[j] = upd ? l : a[j];
                dev::StaticFor<j + 1, kNb>([&](auto cc) {
                    constexpr int c = decltype(cc)::value;
                    a[c] = fmaf(-l, u[c], a[c]);
                });
            }
        }
        if (solo) {
            // the next column's candidate slot is rewritten only after every
            // lane of the wave has read this one
            WaveSync();
        }
    });
    }
    if (solo) {
        if (threadIdx.x == 0) {
            sh.zero = zeroCol;
        }
        __syncthreads();
        zeroCol = sh.zero;
    }
    const int done = zeroCol != 0 ? zeroCol - 1 - k0 : w;
    *wl = done;

    // write back in pivoted order; row map; L11/U11 (FP32) and -L21 in the update's format
    if (valid) {
        const int p = pos;
        sh.rowmap[p] = r;
        if (gmap != nullptr) {
            dev::G(gmap)[p] = r;
        }
        dev::StaticFor<0, kNb>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            if (j < w) {
                LU[(k0 + p) + static_cast<size_t>(k0 + j) * ld] = a[j];
            }
        });
        if (p < kNb) {
            // multiplier columns past a zero pivot are not multipliers: zero (the inverse treats them as
            // the identity's)
            dev::StaticFor<0, kNb / 4>([&](auto cc) {
                constexpr int c = decltype(cc)::value;
                float v[4];
                dev::StaticFor<0, 4>([&](auto ec) {
                    constexpr int e = decltype(ec)::value;
                    v[e] = (4 * c + e < p && 4 * c + e >= done) ? 0.0f : a[4 * c + e];
                });
                *reinterpret_cast<float4*>(&sh.l11[p][4 * c]) = make_float4(v[0], v[1], v[2], v[3]);
            });
        } else if constexpr (P == 0) {
            dev::StaticFor<0, kNb / 4>([&](auto cc) {
                constexpr int c = decltype(cc)::value;
                *reinterpret_cast<float4*>(&sh.op.l[p - kNb][4 * c]) =
                    make_float4(4 * c < done ? -a[4 * c] : 0.0f, 4 * c + 1 < done ? -a[4 * c + 1] : 0.0f,
                                4 * c + 2 < done ? -a[4 * c + 2] : 0.0f, 4 * c + 3 < done ? -a[4 * c + 3] : 0.0f);
            });
        } else {
            dev::StaticFor<0, kNb>([&](auto jc) {
                constexpr int j = decltype(jc)::value;
                uint16_t h[P];
                SplitPlanes<P>(j < done ? -a[j] : 0.0f, h);
                dev::StaticFor<0, P>([&](auto pc) {
                    constexpr int q = decltype(pc)::value;
                    sh.op.l[q][p - kNb][j] = h[q];
                });
            });
        }
    }
    __syncthreads();
    return zeroCol;
}

// inv(L11) by wave 0, one column per lane, by column substitution, where L11
// holds the panel's first wl multiplier columns (the rest of the unit lower
// block is the identity's: the panel wrote zeros there; rows and columns past
// the panel's width are the identity's too).  It
// goes to LDS (for U12) and, when g is given, to g (32 x 32, column-major).
template <int P>
__device__ __forceinline__ void PanelInverses(const PanelView& pv, int k0, int wl, float* g, FactorShared<P>& sh)
{
    const int lane = dev::LaneId();
    const int wave = static_cast<int>(threadIdx.x) / kWave;
    const int w = min(kNb, pv.n - k0);
    if (wave == 0 && lane < kNb) {
        const int j = lane;
        float x[kNb];
        dev::StaticFor<0, kNb>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            float v = q == j ? 1.0f : 0.0f;
            if (q < w) {
                // row q of L11, four multipliers per read (columns past a zero pivot hold zeros)
                dev::StaticFor<0, (q + 3) / 4>([&](auto cc) {
                    constexpr int c = decltype(cc)::value;
                    const float4 rv = *reinterpret_cast<const float4*>(&sh.l11[q][4 * c]);
                    const float r4[4] = {rv.x, rv.y, rv.z, rv.w};
                    dev::StaticFor<0, 4>([&](auto ec) {
                        constexpr int i = 4 * decltype(cc)::value + decltype(ec)::value;
                        if constexpr (i < q) {
                            v = fmaf(-r4[i - 4 * c], x[i], v);
                        }
                    });
                });
            }
            x[q] = v;
        });
        dev::StaticFor<0, kNb>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            sh.linv[q][j] = x[q];
            if (g != nullptr) {
                dev::G(g)[q + kNb * j] = x[q];
            }
        });
    }
}

// One 16-column tile of the trailing matrix in accumulators: element j of
// accumulator t of a lane holds row 16t + (lane & 15) (panel position) and
// column c0 + 4 (lane >> 4) + j.
typedef f32x4 SweepAcc[kMaxTiles];

// Gathers the tile at column c0 through the panel's row map.
template <int P>
__device__ __forceinline__ void GatherTile(const PanelView& pv, int k0, int c0, const FactorShared<P>& sh,
                                           SweepAcc& acc)
{
    const int n = pv.n;
    const size_t ld = static_cast<size_t>(pv.srcLd);
    const int lane = dev::LaneId();
    const int rows = n - k0;
    const int r16 = lane & 15;
    const int cq = (lane >> 4) * 4;
    const auto LU = dev::G(pv.src);
    const int tiles = (rows + 15) / 16;
    dev::StaticFor<0, kMaxTiles>([&](auto tc) {
        constexpr int t = decltype(tc)::value;
        if (t < tiles) {
            const int row = 16 * t + r16;
            const bool ok = row < rows;
            const size_t src = static_cast<size_t>(k0 + (ok ? sh.rowmap[row] : 0));
            dev::StaticFor<0, 4>([&](auto jc) {
                constexpr int j = decltype(jc)::value;
                const int col = c0 + cq + j;
                acc[t][j] = (ok && col < n) ? LU[src + static_cast<size_t>(col) * ld] : 0.0f;
            });
        }
    });
}

// The gathered tile: U12 = inv(L11) A12, update, store in position order.
template <int P>
__device__ __forceinline__ void UpdateTile(const PanelView& pv, int k0, int c0, FactorShared<P>& sh, SweepAcc& acc)
{
    const int n = pv.n;
    const size_t ld = static_cast<size_t>(pv.ld);
    const int lane = dev::LaneId();
    const int wave = static_cast<int>(threadIdx.x) / kWave;
    const int rows = n - k0;
    const int r16 = lane & 15;
    const int grp = lane >> 4;
    const int cq = grp * 4;
    const auto LU = dev::G(pv.LU);
    const int tiles = (rows + 15) / 16;

    // U12 = inv(L11) A12 (top 32 rows) on the FP32 matrix instruction, as
    // U12^T = A12^T inv(L11)^T: A12 through the per-wave LDS tile
    float (*ut)[kTileN + 1] = sh.ut[wave];
    dev::StaticFor<0, 2>([&](auto tc) {
        constexpr int t = decltype(tc)::value;
        dev::StaticFor<0, 4>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            ut[16 * t + r16][cq + j] = acc[t][j];
        });
    });
    WaveSync();
    dev::StaticFor<0, 2>([&](auto bc) {
        constexpr int b = decltype(bc)::value;
        f32x4 v = {0.0f, 0.0f, 0.0f, 0.0f};
        dev::StaticFor<0, kNb / 4>([&](auto sc) {
            constexpr int st = decltype(sc)::value;
            v = __builtin_amdgcn_mfma_f32_16x16x4f32(ut[4 * st + grp][r16], sh.linv[16 * b + r16][4 * st + grp], v,
                                                     0, 0, 0);
        });
        acc[b] = v;
    });

    // U12 in the update's operand format (U12^T: [col][k])
    dev::StaticFor<0, 2>([&](auto bc) {
        constexpr int b = decltype(bc)::value;
        dev::StaticFor<0, 4>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            const int col = cq + j;
            const int k = 16 * b + r16;
            if constexpr (P == 0) {
                sh.op.u[wave][col][k] = acc[b][j];
            } else {
                uint16_t h[P];
                SplitPlanes<P>(acc[b][j], h);
                dev::StaticFor<0, P>([&](auto pc) {
                    constexpr int q = decltype(pc)::value;
                    sh.op.u[wave][q][col][k] = h[q];
                });
            }
        });
    });
    WaveSync();

    // A22 += (-L21) U12, as acc^T += U12^T (-L21)^T, small terms first
    if constexpr (P == 0) {
        float ua[kNb / 4];
        dev::StaticFor<0, kNb / 4>([&](auto sc) {
            constexpr int st = decltype(sc)::value;
            ua[st] = sh.op.u[wave][r16][4 * st + grp];
        });
        dev::StaticFor<2, kMaxTiles>([&](auto tc) {
            constexpr int t = decltype(tc)::value;
            if (t < tiles) {
                const int lr = 16 * t - kNb + r16;
                f32x4 v = acc[t];
                dev::StaticFor<0, kNb / 4>([&](auto sc) {
                    constexpr int st = decltype(sc)::value;
                    v = __builtin_amdgcn_mfma_f32_16x16x4f32(ua[st], sh.op.l[lr][4 * st + grp], v, 0, 0, 0);
                });
                acc[t] = v;
            }
        });
    } else {
        const int kq = grp * 8;
        bf16x8 ub[P];
        dev::StaticFor<0, P>([&](auto pc) {
            constexpr int q = decltype(pc)::value;
            ub[q] = *reinterpret_cast<const bf16x8*>(&sh.op.u[wave][q][r16][kq]);
        });
        dev::StaticFor<2, kMaxTiles>([&](auto tc) {
            constexpr int t = decltype(tc)::value;
            if (t < tiles) {
                const int lr = 16 * t - kNb + r16;
                bf16x8 la[P];
                dev::StaticFor<0, P>([&](auto pc) {
                    constexpr int q = decltype(pc)::value;
                    la[q] = *reinterpret_cast<const bf16x8*>(&sh.op.l[q][lr][kq]);
                });
                // every plane in its own registers before the products start
                __builtin_amdgcn_sched_barrier(0);
                f32x4 v = acc[t];
                if constexpr (P == 3) {
                    v = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ub[0], la[2], v, 0, 0, 0);
                    v = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ub[2], la[0], v, 0, 0, 0);
                    v = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ub[1], la[1], v, 0, 0, 0);
                }
                if constexpr (P >= 2) {
                    v = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ub[0], la[1], v, 0, 0, 0);
                    v = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ub[1], la[0], v, 0, 0, 0);
                }
                v = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ub[0], la[0], v, 0, 0, 0);
                acc[t] = v;
            }
        });
    }

    dev::StaticFor<0, kMaxTiles>([&](auto tc) {
        constexpr int t = decltype(tc)::value;
        if (t < tiles) {
            const int row = 16 * t + r16;
            if (row < rows) {
                dev::StaticFor<0, 4>([&](auto jc) {
                    constexpr int j = decltype(jc)::value;
                    const int col = c0 + cq + j;
                    if (col < n) {
                        LU[(k0 + row) + static_cast<size_t>(col) * ld] = acc[t][j];
                    }
                });
            }
        }
    });
    WaveSync();
}

// The trailing columns k0 + width .. n-1 of the panel at k0, in 16-column
// tiles dealt to the waves in turn.
template <int P>
__device__ __forceinline__ void SweepTrailing(const PanelView& pv, int k0, FactorShared<P>& sh)
{
    const int wave = static_cast<int>(threadIdx.x) / kWave;
    for (int c0 = k0 + kNb + wave * kTileN; c0 < pv.n; c0 += kWaves * kTileN) {
        SweepAcc acc;
        GatherTile<P>(pv, k0, c0, sh, acc);
        UpdateTile<P>(pv, k0, c0, sh, acc);
    }
}

// onlyMarked: factor only the systems the refinement marked for FP32 updates.
}  // namespace lub
}  // namespace aclblas
