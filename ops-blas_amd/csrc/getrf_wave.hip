// Batched LU for n <= 64 (see getrf.hip for the algorithm and its reference).
// S is n rounded up to 8, 16, 32 or 64; a matrix occupies an aligned segment of
// L = S / R lanes and each lane holds R rows in registers: lane r of a
// segment holds rows r, r + L, ... (a[i][0..S-1]).  Rows >= n and columns
// j >= n are padding: zero, never pivot candidates, never updated, never
// stored.  Per column k:
//   1. pivot search: a 32-bit |a_rk| key per row, the lane's best row, then a
//      reduction over the segment with DPP row operations (and
//      v_permlane16/32_swap across 16-lane rows); when several candidate rows
//      share the maximal key (or every candidate is 0 / NaN) a second reduction
//      takes the lowest current position;
//   2. the winning row and its position go to a per-segment LDS slot and every
//      lane of the segment reads them back (broadcast reads; no barrier: the
//      producer and the consumers are lanes of one wave);
//   3. multipliers l = a_rk / p by division, rank-1 update by fmaf.
// Rows never move; each row tracks its logical position and is stored once, at
// its final position, at the end.  R = 2 halves the segment width, so a wave
// carries twice the matrices and the reduction, broadcast and bookkeeping
// instructions per matrix halve, for twice the update work per lane.
// OOP (the batched inverse's LU): the factors go to a packed n x n block per
// matrix (out + b n^2, leading dimension n) instead of back into A, the block's
// address into table[b], and every pivot entry is written (0 after a zero
// pivot), so the input is read once and needs no separate copy.
#include "device_common.h"
#include "internal.h"
#include "lapack_kernels.h"

namespace aclblas {
namespace {

using dev::kWave;

constexpr int kWavesPerBlock = 4;

template <int S, int R, bool PIVOT, bool OOP>
__global__ __launch_bounds__(kWave* kWavesPerBlock) void GetrfWaveKernel(
    float* const* __restrict__ Aarray, int n, int lda, int* __restrict__ ipiv, int* __restrict__ info, int batch,
    float* __restrict__ out, float** __restrict__ table)
{
    constexpr int L = S / R;            // lanes per matrix
    constexpr int kMpw = kWave / L;     // matrices per wave
    constexpr int kSlot = S + 4;        // LDS floats per segment: row + position, 16-byte multiple
    static_assert(L >= 8 && L <= kWave, "segment of 8 to 64 lanes");
    __shared__ __attribute__((aligned(16))) float srow[kWavesPerBlock][kMpw][kSlot];

    const int lane = dev::LaneId();
    const int wave = static_cast<int>(threadIdx.x) / kWave;
    const int seg = lane / L;
    const int r = lane % L;
    const int b = (blockIdx.x * kWavesPerBlock + wave) * kMpw + seg;
    const bool live = b < batch;
    bool real[R];
    dev::StaticFor<0, R>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        real[i] = live && r + i * L < n;
    });
    float* slot = static_cast<float*>(__builtin_assume_aligned(srow[wave][seg], 16));

    // global address space: plain global loads and stores instead of flat ones
    typedef __attribute__((address_space(1))) float GlobalFloat;
    GlobalFloat* A = live ? (GlobalFloat*)(Aarray[b]) : nullptr;
    // Every load is unconditional (element 0 of the first matrix where the
    // element is padding) and the padding is applied afterwards: a load under
    // a lane-dependent condition is issued inside a branch and waited for
    // there, which made the S loads of a row one round trip each.
    float a0[R][S];
    {
        GlobalFloat* A0 = (GlobalFloat*)(Aarray[live ? b : 0]);
        dev::StaticFor<0, R>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            dev::StaticFor<0, S>([&](auto jc) {
                constexpr int j = decltype(jc)::value;
                const bool need = real[i] && j < n;
                a0[i][j] = A0[need ? r + i * L + static_cast<size_t>(j) * lda : 0];
            });
        });
    }
    __builtin_amdgcn_sched_barrier(0);
    float a[R][S];
    dev::StaticFor<0, R>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        dev::StaticFor<0, S>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            a[i][j] = (real[i] && j < n) ? a0[i][j] : 0.0f;
        });
    });

    // number of live segments in this wave: each has exactly one pivot
    // candidate row with the maximal key unless keys tie
    const uint64_t liveMask = __ballot(live && r == 0);
    const int liveSegs = __popcll(liveMask);

    int pos[R];           // current logical position of each of this lane's rows
    int myPiv[R];         // ipiv entry for step k == r + i L
    dev::StaticFor<0, R>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        pos[i] = r + i * L;
        myPiv[i] = 0;
    });
    int infoVal = 0;
    bool running = true;  // false once this matrix met a zero pivot

    dev::StaticFor<0, S>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        if (k >= n) {
            return;
        }
        bool winner[R];
        if (PIVOT) {
            uint32_t key[R];
            bool cand[R];
            uint32_t lbest = 0;
            dev::StaticFor<0, R>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                cand[i] = real[i] && pos[i] >= k;
                key[i] = cand[i] ? dev::MagKey(a[i][k]) : 0u;
                lbest = key[i] > lbest ? key[i] : lbest;
            });
            const uint32_t best = dev::SegReduce<L, true>(lbest);
            int ntie = 0;
            dev::StaticFor<0, R>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                winner[i] = cand[i] && key[i] == best;
                ntie += winner[i] ? 1 : 0;
            });
            const bool multi = (R > 1 && __ballot(ntie > 1) != 0) || __popcll(__ballot(ntie > 0)) != liveSegs;
            if (multi) {
                // several candidate rows with the maximal |a| (or every
                // candidate is 0 / NaN): the lowest position wins
                uint32_t llow = 0xFFFFu;
                dev::StaticFor<0, R>([&](auto ic) {
                    constexpr int i = decltype(ic)::value;
                    const uint32_t pk = winner[i] ? static_cast<uint32_t>(pos[i]) : 0xFFFFu;
                    llow = pk < llow ? pk : llow;
                });
                const uint32_t lowest = dev::SegReduce<L, false>(llow);
                dev::StaticFor<0, R>([&](auto ic) {
                    constexpr int i = decltype(ic)::value;
                    winner[i] = winner[i] && static_cast<uint32_t>(pos[i]) == lowest;
                });
            }
        } else {
            dev::StaticFor<0, R>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                winner[i] = real[i] && pos[i] == k;
            });
        }

        // pivot row (columns k4..S-1) and the winner's position through LDS
        constexpr int k4 = k & ~3;
        dev::StaticFor<0, R>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            if (winner[i]) {
                dev::StaticFor<k4 / 4, S / 4>([&](auto qc) {
                    constexpr int q = decltype(qc)::value;
                    *reinterpret_cast<float4*>(slot + 4 * q) =
                        make_float4(a[i][4 * q], a[i][4 * q + 1], a[i][4 * q + 2], a[i][4 * q + 3]);
                });
                slot[S] = __int_as_float(pos[i]);
            }
        });
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        float u[S];
        dev::StaticFor<k4 / 4, S / 4>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            const float4 v = *reinterpret_cast<const float4*>(slot + 4 * q);
            u[4 * q] = v.x;
            u[4 * q + 1] = v.y;
            u[4 * q + 2] = v.z;
            u[4 * q + 3] = v.w;
        });
        const int pw = __float_as_int(slot[S]);
        // keep the slot stable until every lane has read it
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

        if (!running) {
            return;
        }
        if (PIVOT) {
            dev::StaticFor<0, R>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                if (r + i * L == k) {
                    myPiv[i] = pw + 1;
                }
            });
        }
        const float pivot = u[k];
        if (pivot == 0.0f) {
            infoVal = k + 1;
            running = false;
            return;
        }
        dev::StaticFor<0, R>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            if (PIVOT) {
                const int old = pos[i];
                if (old == k) {
                    pos[i] = pw;
                }
                if (winner[i]) {
                    pos[i] = k;
                }
            }
            if (real[i] && pos[i] > k) {
                const float l = a[i][k] / pivot;
                a[i][k] = l;
                dev::StaticFor<k + 1, S>([&](auto jc) {
                    constexpr int j = decltype(jc)::value;
                    a[i][j] = fmaf(-l, u[j], a[i][j]);
                });
            }
        });
    });

    GlobalFloat* O = A;
    size_t ldo = static_cast<size_t>(lda);
    if (OOP && live) {
        O = (GlobalFloat*)(out + static_cast<size_t>(b) * n * n);
        ldo = static_cast<size_t>(n);
        if (r == 0) {
            table[b] = out + static_cast<size_t>(b) * n * n;
        }
    }
    dev::StaticFor<0, R>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        if (real[i]) {
            dev::StaticFor<0, S>([&](auto jc) {
                constexpr int j = decltype(jc)::value;
                if (j < n) {
                    O[pos[i] + static_cast<size_t>(j) * ldo] = a[i][j];
                }
            });
        }
    });
    if (!live) {
        return;
    }
    if (PIVOT) {
        const int written = infoVal != 0 ? infoVal : n;
        dev::StaticFor<0, R>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            if (r + i * L < written) {
                ipiv[static_cast<size_t>(b) * n + r + i * L] = myPiv[i];
            } else if (OOP && r + i * L < n) {
                ipiv[static_cast<size_t>(b) * n + r + i * L] = 0;
            }
        });
    }
    if (info != nullptr && r == 0) {
        info[b] = infoVal;
    }
}

template <int S, int R, bool PIVOT, bool OOP = false>
void Launch(hipStream_t st, float* const* A, int n, int lda, int* ipiv, int* info, int batch,
            float* out = nullptr, float** table = nullptr)
{
    constexpr int kPerBlock = kWavesPerBlock * (kWave * R / S);
    const dim3 grid((batch + kPerBlock - 1) / kPerBlock);
    const dim3 block(kWave * kWavesPerBlock);
    hipLaunchKernelGGL((GetrfWaveKernel<S, R, PIVOT, OOP>), grid, block, 0, st, A, n, lda, ipiv, info, batch, out,
                       table);
}

}  // namespace

template <bool PIVOT>
void LaunchGetrfWave(hipStream_t st, float* const* A, int n, int lda, int* ipiv, int* info, int batch, int rows)
{
    if (n <= 8) {
        Launch<8, 1, PIVOT>(st, A, n, lda, ipiv, info, batch);
    } else if (n <= 16) {
        if (rows == 2) {
            Launch<16, 2, PIVOT>(st, A, n, lda, ipiv, info, batch);
        } else {
            Launch<16, 1, PIVOT>(st, A, n, lda, ipiv, info, batch);
        }
    } else if (n <= 32) {
        if (rows == 2) {
            Launch<32, 2, PIVOT>(st, A, n, lda, ipiv, info, batch);
        } else {
            Launch<32, 1, PIVOT>(st, A, n, lda, ipiv, info, batch);
        }
    } else {
        if (rows == 2) {
            Launch<64, 2, PIVOT>(st, A, n, lda, ipiv, info, batch);
        } else {
            Launch<64, 1, PIVOT>(st, A, n, lda, ipiv, info, batch);
        }
    }
}

void LaunchGetrfWaveOutOfPlace(hipStream_t st, const float* const* A, int n, int lda, float* out, float** table,
                               int* ipiv, int* info, int batch, int rows)
{
    float* const* in = const_cast<float* const*>(A);  // read only in this mode
    if (n <= 8) {
        Launch<8, 1, true, true>(st, in, n, lda, ipiv, info, batch, out, table);
    } else if (n <= 16) {
        if (rows == 2) {
            Launch<16, 2, true, true>(st, in, n, lda, ipiv, info, batch, out, table);
        } else {
            Launch<16, 1, true, true>(st, in, n, lda, ipiv, info, batch, out, table);
        }
    } else if (rows == 2) {
        Launch<32, 2, true, true>(st, in, n, lda, ipiv, info, batch, out, table);
    } else {
        Launch<32, 1, true, true>(st, in, n, lda, ipiv, info, batch, out, table);
    }
}

template void LaunchGetrfWave<true>(hipStream_t, float* const*, int, int, int*, int*, int, int);
template void LaunchGetrfWave<false>(hipStream_t, float* const*, int, int, int*, int*, int, int);

}  // namespace aclblas
