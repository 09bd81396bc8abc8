// aclblasSgesvBatchedMixed: batched dense solves of mixed orders (n <= 256,
// nrhs <= 16) with an FP32-accurate LU whose trailing updates run on BF16
// matrix cores, and FP64-residual iterative refinement.
//
// The reference library has no such entry.  Its building blocks follow the
// reference's getrf/getrs semantics (partial pivoting, LAPACK-style pivot
// sequence; blas/getrf_batched/arch35/sgetrf_batched_kernel.cpp:52-150,
// blas/getrs_batched/arch35/sgetrs_batched_kernel.cpp:83-259); the algorithm is
// blocked and mixed-precision, so its results are compared with an FP64
// solver through the backward error, not bit by bit.
//
// Factorisation (GesvFactorKernel, persistent, one 256-thread workgroup per
// system, systems in order of decreasing n).  The system is copied to an FP32
// workspace and factored in panels of 32 columns:
//   1. wave 0 holds the panel (up to 256 rows x 32 columns, four rows per
//      lane) in registers and factors it with partial pivoting: rows keep their
//      place and carry their position, the pivot search is a DPP reduction of
//      |a| keys (lowest position on exact ties), the pivot row reaches the other
//      lanes through LDS, multipliers are a * (1/p);
//   2. the panel is written back in pivoted row order; the panel's row map
//      (new position -> old row) is kept for the solves, and -L21 is split into
//      three BF16 planes (hi + mid + lo = the FP32 value) in LDS;
//   3. all four waves sweep the trailing column tiles (16 columns each): the
//      tile's rows are gathered through the row map straight into MFMA
//      accumulators, the top 32 rows are solved with L11 (U12), U12 is split
//      into BF16 planes, and A22 += (-L21) U12 runs as six
//      v_mfma_f32_16x16x32_bf16 per 16 x 16 block (the three cross terms below
//      2^-24 relative are dropped), then the tile is stored back.
// Columns left of a panel are not permuted by later panels; the solves apply
// each panel's row map just before that panel's columns, which is equivalent.
//
// Solve and refinement (GesvRefineKernel, persistent, one workgroup per
// system): x0 from the FP32 factors, then until every right-hand side meets
// |b - A x|_inf / (|A|_inf |x|_inf + |b|_inf) <= tol (residual in FP64 from
// the original FP32 A and b, x kept in FP64) or maxIter corrections have been
// applied: d = LU \ r in FP32, x += d.
#include <algorithm>
#include <vector>

#include "cann_ops_blas.h"
#include "device_common.h"
#include "internal.h"
#include "tuning.h"

namespace aclblas {
namespace {

using dev::kWave;

constexpr int kMaxN = 256;
constexpr int kMaxRhs = 16;
constexpr int kNb = 32;                  // panel width
constexpr int kThreads = 256;            // four waves
constexpr int kWaves = kThreads / kWave;
constexpr int kTileN = 16;               // sweep tile width
constexpr int kMaxTiles = kMaxN / 16;    // 16-row accumulator tiles per column tile
constexpr int kLStride = kNb;            // bf16 elements per L21 plane row (64 bytes, 16-byte aligned)
constexpr int kUStride = kNb;            // bf16 elements per U12 plane row

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

struct MixedSystem {
    const float* A;
    const float* B;
    float* X;
    float* LU;       // workspace: n x n, leading dimension n
    int* rowmap;     // workspace: per-panel row maps, panel p at p*n - 16*p*(p-1)
    int* status;     // workspace: 0, or k+1 for a zero pivot at column k
    int n, nrhs, lda, ldb, ldx, index;
};

__device__ __forceinline__ int RowMapOffset(int n, int p) { return p * n - 16 * p * (p - 1); }

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

// x = hi + mid + lo exactly for finite x of normal range.
__device__ __forceinline__ void Split3(float x, uint16_t& hi, uint16_t& mid, uint16_t& lo)
{
    hi = Bf16Rne(x);
    const float r1 = x - Bf16ToF(hi);
    mid = Bf16Rne(r1);
    const float r2 = r1 - Bf16ToF(mid);
    lo = Bf16Rne(r2);
}

struct FactorShared {
    alignas(16) float cand[2][kWaves][kNb];                  // each wave's best row, by column parity
    uint32_t candKey[2][kWaves];
    int candPos[2][kWaves];
    alignas(16) uint16_t lplanes[3][kMaxN - kNb][kLStride];  // -L21 split into BF16 planes
    alignas(16) uint16_t uplanes[kWaves][3][kTileN][kUStride];
    int rowmap[kMaxN];                                       // panel row map
    float l11[kNb][kNb + 1];                                 // panel rows 0..31 (FP32)
    float ut[kWaves][kNb][kTileN + 1];                       // per-wave TRSM tile
};

__device__ __forceinline__ void WaveSync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Panel factorisation by the whole workgroup: wave v holds rows 64v..64v+63 of
// the panel (one row per lane) in registers.  Per column, each wave reduces
// its candidates with DPP, the wave's best lane publishes its row, |a| key and
// position in LDS (double-buffered by column parity), one barrier, and every
// lane picks the same winner among the four.  Returns the zero-pivot column
// + 1, or 0 (uniform over the workgroup).
__device__ __forceinline__ int FactorPanel(const MixedSystem& s, int k0, FactorShared& sh)
{
    const int n = s.n;
    const int lane = dev::LaneId();
    const int wave = static_cast<int>(threadIdx.x) / kWave;
    const int rows = n - k0;
    const int w = min(kNb, rows);
    const int r = wave * kWave + lane;
    const bool valid = r < rows;
    float* LU = s.LU;

    float a[kNb];
    dev::StaticFor<0, kNb>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        a[j] = (valid && j < w) ? LU[(k0 + r) + static_cast<size_t>(k0 + j) * n] : 0.0f;
    });
    int pos = r;
    int zeroCol = 0;

    dev::StaticFor<0, kNb>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        if (j >= w || zeroCol != 0) {
            return;
        }
        constexpr int buf = j & 1;
        const bool cand = valid && pos >= j;
        const uint32_t key = cand ? dev::MagKey(a[j]) : 0u;
        // wave best: largest key, then lowest position (non-candidates lose)
        const uint32_t best = dev::SegReduce<kWave, true>(key);
        const bool tied = cand && key == best;
        const uint32_t lowest = dev::SegReduce<kWave, false>(tied ? static_cast<uint32_t>(pos) : 0xFFFFu);
        const bool mine = tied && static_cast<uint32_t>(pos) == lowest;
        if (mine) {
            dev::StaticFor<0, kNb / 4>([&](auto cc) {
                constexpr int c = decltype(cc)::value;
                *reinterpret_cast<float4*>(&sh.cand[buf][wave][4 * c]) =
                    make_float4(a[4 * c], a[4 * c + 1], a[4 * c + 2], a[4 * c + 3]);
            });
            sh.candKey[buf][wave] = key;
            sh.candPos[buf][wave] = pos;
        } else if (lane == 0 && lowest == 0xFFFFu) {
            sh.candPos[buf][wave] = 0xFFFF;  // no candidate in this wave
            sh.candKey[buf][wave] = 0;
        }
        __syncthreads();
        int wb = -1;
        uint32_t bk = 0;
        int bp = 0xFFFF;
        dev::StaticFor<0, kWaves>([&](auto vc) {
            constexpr int v = decltype(vc)::value;
            const int p = sh.candPos[buf][v];
            const uint32_t k = sh.candKey[buf][v];
            if (p != 0xFFFF && (wb < 0 || k > bk || (k == bk && p < bp))) {
                wb = v;
                bk = k;
                bp = p;
            }
        });
        wb = dev::Uniform(wb);
        const int pw = dev::Uniform(bp);
        float u[kNb];
        dev::StaticFor<j / 4, kNb / 4>([&](auto cc) {
            constexpr int c = decltype(cc)::value;
            const float4 v = *reinterpret_cast<const float4*>(&sh.cand[buf][wb][4 * c]);
            u[4 * c] = v.x;
            u[4 * c + 1] = v.y;
            u[4 * c + 2] = v.z;
            u[4 * c + 3] = v.w;
        });
        const float pivot = u[j];
        if (pivot == 0.0f) {
            zeroCol = k0 + j + 1;
            return;
        }
        const bool winner = pos == pw && cand;
        if (pos == j) {
            pos = pw;
        }
        if (winner) {
            pos = j;
        }
        const float rcp = 1.0f / pivot;
        if (valid && pos > j) {
            const float l = a[j] * rcp;
            a[j] = l;
            dev::StaticFor<j + 1, kNb>([&](auto cc) {
                constexpr int c = decltype(cc)::value;
                a[c] = fmaf(-l, u[c], a[c]);
            });
        }
    });
    if (zeroCol != 0) {
        return zeroCol;
    }

    // write back in pivoted order; row map; L11 (FP32) and -L21 planes (BF16)
    int* gmap = s.rowmap + RowMapOffset(n, k0 / kNb);
    if (valid) {
        const int p = pos;
        sh.rowmap[p] = r;
        gmap[p] = r;
        dev::StaticFor<0, kNb>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            if (j < w) {
                LU[(k0 + p) + static_cast<size_t>(k0 + j) * n] = a[j];
            }
        });
        if (p < kNb) {
            dev::StaticFor<0, kNb>([&](auto jc) {
                constexpr int j = decltype(jc)::value;
                sh.l11[p][j] = a[j];
            });
        } else {
            dev::StaticFor<0, kNb>([&](auto jc) {
                constexpr int j = decltype(jc)::value;
                uint16_t h, m, l;
                Split3(-a[j], h, m, l);
                sh.lplanes[0][p - kNb][j] = h;
                sh.lplanes[1][p - kNb][j] = m;
                sh.lplanes[2][p - kNb][j] = l;
            });
        }
    }
    __syncthreads();
    return 0;
}

// One 16-column tile of the trailing matrix: gather, TRSM, BF16x3 update, store.
__device__ __forceinline__ void SweepTile(const MixedSystem& s, int k0, int c0, FactorShared& sh)
{
    const int n = s.n;
    const int lane = dev::LaneId();
    const int wave = static_cast<int>(threadIdx.x) / kWave;
    const int rows = n - k0;
    const int col = c0 + (lane & 15);
    const bool colOk = col < n;
    const int rq = (lane >> 4) * 4;
    float* LU = s.LU;
    const int tiles = (rows + 15) / 16;

    f32x4 acc[kMaxTiles];
    dev::StaticFor<0, kMaxTiles>([&](auto tc) {
        constexpr int t = decltype(tc)::value;
        if (t < tiles) {
            dev::StaticFor<0, 4>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                const int row = 16 * t + rq + i;
                acc[t][i] = (row < rows && colOk) ? LU[(k0 + sh.rowmap[row]) + static_cast<size_t>(col) * n] : 0.0f;
            });
        }
    });

    // U12 = L11^-1 (top 32 rows): through the per-wave LDS tile
    float (*ut)[kTileN + 1] = sh.ut[wave];
    dev::StaticFor<0, 2>([&](auto tc) {
        constexpr int t = decltype(tc)::value;
        dev::StaticFor<0, 4>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            ut[16 * t + rq + i][lane & 15] = acc[t][i];
        });
    });
    WaveSync();
    if (lane < kTileN) {
        float x[kNb];
        dev::StaticFor<0, kNb>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            float v = ut[q][lane];
            dev::StaticFor<0, q>([&](auto jc) {
                constexpr int j = decltype(jc)::value;
                v = fmaf(-sh.l11[q][j], x[j], v);
            });
            x[q] = v;
            ut[q][lane] = v;
        });
        dev::StaticFor<0, kNb>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            uint16_t h, m, l;
            Split3(x[q], h, m, l);
            sh.uplanes[wave][0][lane][q] = h;
            sh.uplanes[wave][1][lane][q] = m;
            sh.uplanes[wave][2][lane][q] = l;
        });
    }
    WaveSync();
    dev::StaticFor<0, 2>([&](auto tc) {
        constexpr int t = decltype(tc)::value;
        dev::StaticFor<0, 4>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            acc[t][i] = ut[16 * t + rq + i][lane & 15];
        });
    });

    // A22 += (-L21) U12 with split BF16 operands (small terms first)
    const int kq = (lane >> 4) * 8;
    bf16x8 ub[3];
    dev::StaticFor<0, 3>([&](auto pc) {
        constexpr int p = decltype(pc)::value;
        ub[p] = *reinterpret_cast<const bf16x8*>(&sh.uplanes[wave][p][lane & 15][kq]);
    });
    dev::StaticFor<2, kMaxTiles>([&](auto tc) {
        constexpr int t = decltype(tc)::value;
        if (t < tiles) {
            const int lr = 16 * t - kNb + (lane & 15);
            bf16x8 la[3];
            dev::StaticFor<0, 3>([&](auto pc) {
                constexpr int p = decltype(pc)::value;
                la[p] = *reinterpret_cast<const bf16x8*>(&sh.lplanes[p][lr][kq]);
            });
            f32x4 v = acc[t];
            v = __builtin_amdgcn_mfma_f32_16x16x32_bf16(la[2], ub[0], v, 0, 0, 0);
            v = __builtin_amdgcn_mfma_f32_16x16x32_bf16(la[0], ub[2], v, 0, 0, 0);
            v = __builtin_amdgcn_mfma_f32_16x16x32_bf16(la[1], ub[1], v, 0, 0, 0);
            v = __builtin_amdgcn_mfma_f32_16x16x32_bf16(la[1], ub[0], v, 0, 0, 0);
            v = __builtin_amdgcn_mfma_f32_16x16x32_bf16(la[0], ub[1], v, 0, 0, 0);
            v = __builtin_amdgcn_mfma_f32_16x16x32_bf16(la[0], ub[0], v, 0, 0, 0);
            acc[t] = v;
        }
    });

    dev::StaticFor<0, kMaxTiles>([&](auto tc) {
        constexpr int t = decltype(tc)::value;
        if (t < tiles) {
            dev::StaticFor<0, 4>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                const int row = 16 * t + rq + i;
                if (row < rows && colOk) {
                    LU[(k0 + row) + static_cast<size_t>(col) * n] = acc[t][i];
                }
            });
        }
    });
    WaveSync();
}

__global__ __launch_bounds__(kThreads, 2) void GesvFactorKernel(const MixedSystem* __restrict__ systems, int count)
{
    __shared__ __attribute__((aligned(16))) FactorShared sh;
    __shared__ int failed;
    const int t = threadIdx.x;
    const int wave = t / kWave;
    for (int idx = blockIdx.x; idx < count; idx += gridDim.x) {
        const MixedSystem s = systems[idx];
        const int n = s.n;
        // working copy of A
        if (t < n) {
            for (int j0 = 0; j0 < n; j0 += 8) {
                float v[8];
                dev::StaticFor<0, 8>([&](auto jc) {
                    constexpr int j = decltype(jc)::value;
                    v[j] = j0 + j < n ? s.A[t + static_cast<size_t>(j0 + j) * s.lda] : 0.0f;
                });
                dev::StaticFor<0, 8>([&](auto jc) {
                    constexpr int j = decltype(jc)::value;
                    if (j0 + j < n) {
                        s.LU[t + static_cast<size_t>(j0 + j) * n] = v[j];
                    }
                });
            }
        }
        if (t == 0) {
            failed = 0;
        }
        __syncthreads();
        for (int k0 = 0; k0 < n; k0 += kNb) {
            const int z = FactorPanel(s, k0, sh);
            if (z != 0) {
                if (t == 0) {
                    failed = z;
                }
                break;
            }
            const int first = k0 + kNb;  // first trailing column (panels before the last are full)
            for (int c0 = first + wave * kTileN; c0 < n; c0 += kWaves * kTileN) {
                SweepTile(s, k0, c0, sh);
            }
            __syncthreads();
        }
        if (t == 0) {
            *s.status = failed;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Solve and refinement
// ---------------------------------------------------------------------------
struct RefineShared {
    float bs[kMaxN][kMaxRhs];
    double x[kMaxN][kMaxRhs];
    double red[3][kWaves][kMaxRhs];  // per-wave max |r|, |x|, |b|
    double anorm[kWaves];            // per-wave max row sum of |A|
    double berr[kMaxRhs];
};

__device__ __forceinline__ double WaveMaxD(double v)
{
    for (int off = 32; off >= 1; off >>= 1) {
        const double o = __shfl_xor(v, off, kWave);
        v = o > v ? o : v;
    }
    return v;
}

// bs <- LU^-1 bs (forward with per-panel row maps, then backward).
__device__ __forceinline__ void FactoredSolve(const MixedSystem& s, RefineShared& sh)
{
    const int n = s.n;
    const int nrhs = s.nrhs;
    const int t = threadIdx.x;
    const int lane = dev::LaneId();
    const int wave = t / kWave;
    const float* LU = s.LU;
    for (int k0 = 0, p = 0; k0 < n; k0 += kNb, ++p) {
        const int rows = n - k0;
        const int w = min(kNb, rows);
        const int* map = s.rowmap + RowMapOffset(n, p);
        {
            // gather through the row map in registers, then write back in place
            float g[kMaxRhs];
            const int src = t < rows ? map[t] : 0;
            dev::StaticFor<0, kMaxRhs>([&](auto cc) {
                constexpr int c = decltype(cc)::value;
                g[c] = (t < rows && c < nrhs) ? sh.bs[k0 + src][c] : 0.0f;
            });
            __syncthreads();
            dev::StaticFor<0, kMaxRhs>([&](auto cc) {
                constexpr int c = decltype(cc)::value;
                if (t < rows && c < nrhs) {
                    sh.bs[k0 + t][c] = g[c];
                }
            });
        }
        __syncthreads();
        {
            // diagonal block: lane q holds row q; wave v solves right-hand sides c = v, v+4, ...
            const bool valid = lane < w;
            float l[kNb];
            float x[kMaxRhs / kWaves];
            dev::StaticFor<0, kNb>([&](auto jc) {
                constexpr int j = decltype(jc)::value;
                l[j] = (valid && j < lane) ? LU[(k0 + lane) + static_cast<size_t>(k0 + j) * n] : 0.0f;
            });
            dev::StaticFor<0, kMaxRhs / kWaves>([&](auto cc) {
                constexpr int ci = decltype(cc)::value;
                const int c = ci * kWaves + wave;
                x[ci] = (valid && c < nrhs) ? sh.bs[k0 + lane][c] : 0.0f;
            });
            dev::StaticFor<0, kNb>([&](auto jc) {
                constexpr int j = decltype(jc)::value;
                if (j < w) {
                    dev::StaticFor<0, kMaxRhs / kWaves>([&](auto cc) {
                        constexpr int ci = decltype(cc)::value;
                        if (ci * kWaves + wave < nrhs) {
                            const float xj = dev::ReadLane(x[ci], j);
                            if (lane > j) {
                                x[ci] = fmaf(-l[j], xj, x[ci]);
                            }
                        }
                    });
                }
            });
            if (valid) {
                dev::StaticFor<0, kMaxRhs / kWaves>([&](auto cc) {
                    constexpr int ci = decltype(cc)::value;
                    const int c = ci * kWaves + wave;
                    if (c < nrhs) {
                        sh.bs[k0 + lane][c] = x[ci];
                    }
                });
            }
        }
        __syncthreads();
        const int i = k0 + w + t;
        if (i < n) {
            float acc[kMaxRhs];
            dev::StaticFor<0, kMaxRhs>([&](auto cc) {
                constexpr int c = decltype(cc)::value;
                acc[c] = c < nrhs ? sh.bs[i][c] : 0.0f;
            });
            float lv[kNb];
            dev::StaticFor<0, kNb>([&](auto jc) {
                constexpr int j = decltype(jc)::value;
                lv[j] = j < w ? LU[i + static_cast<size_t>(k0 + j) * n] : 0.0f;
            });
            dev::StaticFor<0, kNb>([&](auto jc) {
                constexpr int j = decltype(jc)::value;
                if (j < w) {
                    dev::StaticFor<0, kMaxRhs>([&](auto cc) {
                        constexpr int c = decltype(cc)::value;
                        if (c < nrhs) {
                            acc[c] = fmaf(-lv[j], sh.bs[k0 + j][c], acc[c]);
                        }
                    });
                }
            });
            dev::StaticFor<0, kMaxRhs>([&](auto cc) {
                constexpr int c = decltype(cc)::value;
                if (c < nrhs) {
                    sh.bs[i][c] = acc[c];
                }
            });
        }
        __syncthreads();
    }
    // backward: 32-row blocks from the bottom
    const int blocks = (n + kNb - 1) / kNb;
    for (int bi = blocks - 1; bi >= 0; --bi) {
        const int r0 = bi * kNb;
        const int w = min(kNb, n - r0);
        {
            const bool valid = lane < w;
            float u[kNb];
            float x[kMaxRhs / kWaves];
            dev::StaticFor<0, kNb>([&](auto jc) {
                constexpr int j = decltype(jc)::value;
                u[j] = (valid && j >= lane && j < w) ? LU[(r0 + lane) + static_cast<size_t>(r0 + j) * n] : 0.0f;
            });
            dev::StaticFor<0, kMaxRhs / kWaves>([&](auto cc) {
                constexpr int ci = decltype(cc)::value;
                const int c = ci * kWaves + wave;
                x[ci] = (valid && c < nrhs) ? sh.bs[r0 + lane][c] : 0.0f;
            });
            dev::StaticFor<0, kNb>([&](auto jc) {
                constexpr int j = kNb - 1 - decltype(jc)::value;
                if (j < w) {
                    dev::StaticFor<0, kMaxRhs / kWaves>([&](auto cc) {
                        constexpr int ci = decltype(cc)::value;
                        if (ci * kWaves + wave < nrhs) {
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
                    const int c = ci * kWaves + wave;
                    if (c < nrhs) {
                        sh.bs[r0 + lane][c] = x[ci];
                    }
                });
            }
        }
        __syncthreads();
        if (t < r0) {
            float acc[kMaxRhs];
            dev::StaticFor<0, kMaxRhs>([&](auto cc) {
                constexpr int c = decltype(cc)::value;
                acc[c] = c < nrhs ? sh.bs[t][c] : 0.0f;
            });
            float uv[kNb];
            dev::StaticFor<0, kNb>([&](auto jc) {
                constexpr int j = decltype(jc)::value;
                uv[j] = j < w ? LU[t + static_cast<size_t>(r0 + j) * n] : 0.0f;
            });
            dev::StaticFor<0, kNb>([&](auto jc) {
                constexpr int j = decltype(jc)::value;
                if (j < w) {
                    dev::StaticFor<0, kMaxRhs>([&](auto cc) {
                        constexpr int c = decltype(cc)::value;
                        if (c < nrhs) {
                            acc[c] = fmaf(-uv[j], sh.bs[r0 + j][c], acc[c]);
                        }
                    });
                }
            });
            dev::StaticFor<0, kMaxRhs>([&](auto cc) {
                constexpr int c = decltype(cc)::value;
                if (c < nrhs) {
                    sh.bs[t][c] = acc[c];
                }
            });
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(kThreads, 3) void GesvRefineKernel(const MixedSystem* __restrict__ systems, int count,
                                                            double tol, int maxIter, int* __restrict__ iterOut,
                                                            float* __restrict__ berrOut, int* __restrict__ infoOut)
{
    __shared__ __attribute__((aligned(16))) RefineShared sh;
    const int t = threadIdx.x;
    for (int idx = blockIdx.x; idx < count; idx += gridDim.x) {
        const MixedSystem s = systems[idx];
        const int n = s.n;
        const int nrhs = s.nrhs;
        const int status = *s.status;
        if (status != 0) {
            for (int c = 0; c < nrhs; ++c) {
                for (int i = t; i < n; i += kThreads) {
                    s.X[i + static_cast<size_t>(c) * s.ldx] = __builtin_nanf("");
                }
            }
            if (t == 0) {
                if (iterOut != nullptr) iterOut[s.index] = 0;
                if (berrOut != nullptr) berrOut[s.index] = __builtin_inff();
                if (infoOut != nullptr) infoOut[s.index] = status;
            }
            continue;
        }
        // x0 = LU \ b
        if (t < n) {
            for (int c = 0; c < nrhs; ++c) {
                sh.bs[t][c] = s.B[t + static_cast<size_t>(c) * s.ldb];
            }
        }
        __syncthreads();
        FactoredSolve(s, sh);
        if (t < n) {
            for (int c = 0; c < nrhs; ++c) {
                sh.x[t][c] = static_cast<double>(sh.bs[t][c]);
            }
        }
        __syncthreads();
        int iter = 0;
        double worst = 0.0;
        for (;;) {
            // r = b - A x in FP64 (kept in registers for the correction solve);
            // norms reduced through LDS
            double r[kMaxRhs];
            double rs = 0.0;
            dev::StaticFor<0, kMaxRhs>([&](auto cc) {
                constexpr int c = decltype(cc)::value;
                r[c] = (t < n && c < nrhs) ? static_cast<double>(s.B[t + static_cast<size_t>(c) * s.ldb]) : 0.0;
            });
            if (t < n) {
                for (int j0 = 0; j0 < n; j0 += 8) {
                    float av[8];
                    dev::StaticFor<0, 8>([&](auto jc) {
                        constexpr int j = decltype(jc)::value;
                        av[j] = j0 + j < n ? s.A[t + static_cast<size_t>(j0 + j) * s.lda] : 0.0f;
                    });
                    dev::StaticFor<0, 8>([&](auto jc) {
                        constexpr int j = decltype(jc)::value;
                        if (j0 + j < n) {
                            const double a = static_cast<double>(av[j]);
                            rs += fabs(a);
                            dev::StaticFor<0, kMaxRhs>([&](auto cc) {
                                constexpr int c = decltype(cc)::value;
                                if (c < nrhs) {
                                    r[c] = fma(-a, sh.x[j0 + j][c], r[c]);
                                }
                            });
                        }
                    });
                }
            }
            const int lane = dev::LaneId();
            const int wave = t / kWave;
            {
                const double m = WaveMaxD(rs);
                if (lane == 0) {
                    sh.anorm[wave] = m;
                }
            }
            dev::StaticFor<0, kMaxRhs>([&](auto cc) {
                constexpr int c = decltype(cc)::value;
                if (c < nrhs) {
                    const double rv = WaveMaxD(fabs(r[c]));
                    const double xv = WaveMaxD(t < n ? fabs(sh.x[t][c]) : 0.0);
                    const double bv = WaveMaxD(t < n ? fabs(static_cast<double>(s.B[t + static_cast<size_t>(c) *
                                                                                      s.ldb])) : 0.0);
                    if (lane == 0) {
                        sh.red[0][wave][c] = rv;
                        sh.red[1][wave][c] = xv;
                        sh.red[2][wave][c] = bv;
                    }
                }
            });
            __syncthreads();
            if (t < nrhs) {
                double an = 0.0, rn = 0.0, xn = 0.0, bn = 0.0;
                for (int q = 0; q < kWaves; ++q) {
                    an = sh.anorm[q] > an ? sh.anorm[q] : an;
                    rn = sh.red[0][q][t] > rn ? sh.red[0][q][t] : rn;
                    xn = sh.red[1][q][t] > xn ? sh.red[1][q][t] : xn;
                    bn = sh.red[2][q][t] > bn ? sh.red[2][q][t] : bn;
                }
                const double den = an * xn + bn;
                sh.berr[t] = den > 0.0 ? rn / den : (rn > 0.0 ? __builtin_inf() : 0.0);
            }
            __syncthreads();
            worst = 0.0;
            for (int c = 0; c < nrhs; ++c) {
                const double be = sh.berr[c];
                worst = (be > worst || !(be == be)) ? be : worst;
            }
            __syncthreads();
            if (worst <= tol || iter >= maxIter || !(worst == worst)) {
                break;
            }
            // d = LU \ r (FP32), x += d
            dev::StaticFor<0, kMaxRhs>([&](auto cc) {
                constexpr int c = decltype(cc)::value;
                if (t < n && c < nrhs) {
                    sh.bs[t][c] = static_cast<float>(r[c]);
                }
            });
            __syncthreads();
            FactoredSolve(s, sh);
            if (t < n) {
                for (int c = 0; c < nrhs; ++c) {
                    sh.x[t][c] += static_cast<double>(sh.bs[t][c]);
                }
            }
            __syncthreads();
            ++iter;
        }
        if (t < n) {
            for (int c = 0; c < nrhs; ++c) {
                s.X[t + static_cast<size_t>(c) * s.ldx] = static_cast<float>(sh.x[t][c]);
            }
        }
        if (t == 0) {
            if (iterOut != nullptr) iterOut[s.index] = iter;
            if (berrOut != nullptr) berrOut[s.index] = static_cast<float>(worst);
            if (infoOut != nullptr) infoOut[s.index] = 0;
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
                              int maxIter)
{
    if (batch < 0 || !(tol >= 0.0) || maxIter < 0) {
        return ACLBLAS_STATUS_INVALID_VALUE;
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

}  // namespace
}  // namespace aclblas

extern "C" aclblasStatus_t aclblasSgesvBatchedMixed(
    aclblasHandle_t handle, int batchCount, const int* nArray, const int* nrhsArray, const float* const Aarray[],
    const int* ldaArray, const float* const Barray[], const int* ldbArray, float* const Xarray[],
    const int* ldxArray, double tol, int maxIter, int* iterArray, float* berrArray, int* infoArray)
{
    using namespace aclblas;
    if (handle == nullptr) {
        return ACLBLAS_STATUS_HANDLE_IS_NULLPTR;
    }
    aclblasStatus_t st = ValidateMixed(batchCount, nArray, nrhsArray, Aarray, ldaArray, Barray, ldbArray, Xarray,
                                       ldxArray, tol, maxIter);
    if (st != ACLBLAS_STATUS_SUCCESS) {
        LogError("aclblasSgesvBatchedMixed", "invalid argument (status %d)", static_cast<int>(st));
        return st;
    }
    if (batchCount == 0) {
        return ACLBLAS_STATUS_SUCCESS;
    }
    if (!HandleHasDevice(handle)) {
        return ACLBLAS_STATUS_EXECUTION_FAILED;
    }
    // largest systems first; workspace: table, then per system LU, row maps, status
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
        offs[static_cast<size_t>(i)] = total;
        total += AlignUp(sizeof(float) * static_cast<size_t>(n) * n) + AlignUp(sizeof(int) * RowMapInts(n)) + kAlign;
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
        MixedSystem& s = table[static_cast<size_t>(q)];
        char* p = base + offs[static_cast<size_t>(i)];
        s.A = Aarray[i];
        s.B = Barray[i];
        s.X = Xarray[i];
        s.LU = reinterpret_cast<float*>(p);
        p += AlignUp(sizeof(float) * static_cast<size_t>(n) * n);
        s.rowmap = reinterpret_cast<int*>(p);
        p += AlignUp(sizeof(int) * RowMapInts(n));
        s.status = reinterpret_cast<int*>(p);
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
    // persistent grids, workgroups per CU from the tuning table (built-in: two
    // factor and three solve workgroups, what LDS and registers let one CU hold)
    const tune::Choice c = tune::Get(handle, tune::kOpMixed, tune::MixedBucket(batchCount, cus), tune::Choice{2, 3});
    const unsigned gridF = static_cast<unsigned>(std::min(batchCount, c.a * cus));
    const unsigned gridS = static_cast<unsigned>(std::min(batchCount, c.b * cus));
    const MixedSystem* sys = static_cast<const MixedSystem*>(dev);
    hipLaunchKernelGGL(GesvFactorKernel, dim3(gridF), dim3(kThreads), 0, handle->stream, sys, batchCount);
    hipLaunchKernelGGL(GesvRefineKernel, dim3(gridS), dim3(kThreads), 0, handle->stream, sys, batchCount, tol,
                       maxIter, iterArray, berrArray, infoArray);
    return LaunchStatus("aclblasSgesvBatchedMixed");
}
