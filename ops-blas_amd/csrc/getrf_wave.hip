// Batched LU for n <= 64: one lane per row, 64/S matrices per wave, where S is
// n rounded up to 8, 16, 32 or 64 (see getrf.hip for the algorithm and its
// reference).
//
// Each matrix occupies an aligned segment of S lanes.  Lane r of a segment
// holds row r of its matrix in registers (a[0..S-1]); rows r >= n and columns
// j >= n are padding: zero, never pivot candidates, never updated, never
// stored.  Per column k:
//   1. pivot search: a 32-bit |a_rk| key, reduced over the segment with DPP row
//      operations (plus v_permlane16/32_swap across 16-lane rows); exact
//      ties between candidates fall back to a second reduction that takes the
//      lowest current position;
//   2. the winning lane writes its row and its position to a per-segment LDS
//      slot, and every lane of the segment reads them back (broadcast reads,
//      no barrier: the producer and the consumers are lanes of one wave);
//   3. multipliers l = a_rk / p by division, rank-1 update by fmaf.
// Rows never move; each lane tracks its row's logical position and the row is
// stored once, at its final position, at the end.
#include "device_common.h"
#include "internal.h"
#include "lapack_kernels.h"

namespace aclblas {
namespace {

using dev::kWave;

constexpr int kWavesPerBlock = 4;

template <int S, bool PIVOT>
__global__ __launch_bounds__(kWave* kWavesPerBlock) void GetrfWaveKernel(
    float* const* __restrict__ Aarray, int n, int lda, int* __restrict__ ipiv, int* __restrict__ info, int batch)
{
    constexpr int kMpw = kWave / S;     // matrices per wave
    constexpr int kSlot = S + 4;        // LDS floats per segment: row + position, 16-byte multiple
    __shared__ __attribute__((aligned(16))) float srow[kWavesPerBlock][kMpw][kSlot];

    const int lane = dev::LaneId();
    const int wave = static_cast<int>(threadIdx.x) / kWave;
    const int seg = lane / S;
    const int r = lane % S;
    const int b = (blockIdx.x * kWavesPerBlock + wave) * kMpw + seg;
    const bool live = b < batch;
    const bool real = live && r < n;
    float* slot = static_cast<float*>(__builtin_assume_aligned(srow[wave][seg], 16));

    // global address space: plain global loads and stores instead of flat ones
    typedef __attribute__((address_space(1))) float GlobalFloat;
    GlobalFloat* A = live ? (GlobalFloat*)(Aarray[b]) : nullptr;
    float a[S];
    {
        const int jlim = real ? n : 0;
        dev::StaticFor<0, S>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            a[j] = j < jlim ? A[r + static_cast<size_t>(j) * lda] : 0.0f;
        });
    }

    // number of live segments in this wave: each has exactly one pivot
    // candidate with the maximal key unless keys tie
    const uint64_t liveMask = __ballot(live && r == 0);
    const int liveSegs = __popcll(liveMask);

    int pos = r;          // current logical position of this lane's row
    int myPiv = 0;        // ipiv entry for step k == r
    int infoVal = 0;
    bool running = true;  // false once this matrix met a zero pivot

    dev::StaticFor<0, S>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        if (k >= n) {
            return;
        }
        bool winner;
        if (PIVOT) {
            const bool cand = real && pos >= k;
            const uint32_t key = cand ? dev::MagKey(a[k]) : 0u;
            const uint32_t best = dev::SegReduce<S, true>(key);
            bool tie = cand && key == best;
            if (__popcll(__ballot(tie)) != liveSegs) {
                // some segment has several candidates with the maximal |a|
                // (or every candidate is 0 / NaN): lowest position wins
                const uint32_t lowest = dev::SegReduce<S, false>(tie ? static_cast<uint32_t>(pos) : 0xFFFFu);
                tie = tie && static_cast<uint32_t>(pos) == lowest;
            }
            winner = tie;
        } else {
            winner = real && pos == k;
        }

        // pivot row (columns k4..S-1) and the winner's position through LDS
        constexpr int k4 = k & ~3;
        if (winner) {
            dev::StaticFor<k4 / 4, S / 4>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                *reinterpret_cast<float4*>(slot + 4 * q) =
                    make_float4(a[4 * q], a[4 * q + 1], a[4 * q + 2], a[4 * q + 3]);
            });
            slot[S] = __int_as_float(pos);
        }
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
        if (PIVOT && r == k) {
            myPiv = pw + 1;
        }
        const float pivot = u[k];
        if (pivot == 0.0f) {
            infoVal = k + 1;
            running = false;
            return;
        }
        if (PIVOT) {
            const int old = pos;
            if (old == k) {
                pos = pw;
            }
            if (winner) {
                pos = k;
            }
        }
        if (real && pos > k) {
            const float l = a[k] / pivot;
            a[k] = l;
            dev::StaticFor<k + 1, S>([&](auto jc) {
                constexpr int j = decltype(jc)::value;
                a[j] = fmaf(-l, u[j], a[j]);
            });
        }
    });

    if (real) {
        dev::StaticFor<0, S>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            if (j < n) {
                A[pos + static_cast<size_t>(j) * lda] = a[j];
            }
        });
    }
    if (!live) {
        return;
    }
    if (PIVOT) {
        const int written = infoVal != 0 ? infoVal : n;
        if (r < written) {
            ipiv[static_cast<size_t>(b) * n + r] = myPiv;
        }
    }
    if (info != nullptr && r == 0) {
        info[b] = infoVal;
    }
}

template <int S, bool PIVOT>
void Launch(hipStream_t st, float* const* A, int n, int lda, int* ipiv, int* info, int batch)
{
    constexpr int kPerBlock = kWavesPerBlock * (kWave / S);
    const dim3 grid((batch + kPerBlock - 1) / kPerBlock);
    const dim3 block(kWave * kWavesPerBlock);
    hipLaunchKernelGGL((GetrfWaveKernel<S, PIVOT>), grid, block, 0, st, A, n, lda, ipiv, info, batch);
}

}  // namespace

template <bool PIVOT>
void LaunchGetrfWave(hipStream_t st, float* const* A, int n, int lda, int* ipiv, int* info, int batch)
{
    if (n <= 8) {
        Launch<8, PIVOT>(st, A, n, lda, ipiv, info, batch);
    } else if (n <= 16) {
        Launch<16, PIVOT>(st, A, n, lda, ipiv, info, batch);
    } else if (n <= 32) {
        Launch<32, PIVOT>(st, A, n, lda, ipiv, info, batch);
    } else {
        Launch<64, PIVOT>(st, A, n, lda, ipiv, info, batch);
    }
}

template void LaunchGetrfWave<true>(hipStream_t, float* const*, int, int, int*, int*, int);
template void LaunchGetrfWave<false>(hipStream_t, float* const*, int, int, int*, int*, int);

}  // namespace aclblas
