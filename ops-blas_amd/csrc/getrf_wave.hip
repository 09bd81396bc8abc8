// Batched LU for n <= 64: one wave per matrix, one lane per row (see getrf.hip
// for the algorithm and its reference).
#include "device_common.h"
#include "internal.h"
#include "lapack_kernels.h"

namespace aclblas {
namespace {

using dev::kWave;

// ---------------------------------------------------------------------------
// n <= N <= 64: one wave per matrix
// ---------------------------------------------------------------------------
constexpr int kWavesPerBlock = 4;

template <int N, bool PIVOT>
__global__ __launch_bounds__(kWave* kWavesPerBlock) void GetrfWaveKernel(
    float* const* __restrict__ Aarray, int n, int lda, int* __restrict__ ipiv, int* __restrict__ info, int batch)
{
    const int lane = dev::LaneId();
    const int b = blockIdx.x * kWavesPerBlock + static_cast<int>(threadIdx.x) / kWave;
    if (b >= batch) {
        return;
    }
    float* A = Aarray[b];
    const bool valid = lane < n;

    float a[N];
    dev::StaticFor<0, N>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        a[j] = (valid && j < n) ? A[lane + static_cast<size_t>(j) * lda] : 0.0f;
    });

    int pos = lane;      // current logical row position of this lane's row
    int myPiv = 0;       // ipiv entry for step == lane
    int infoVal = 0;
    bool running = true;

    dev::StaticFor<0, N>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        if (!running || k >= n) {
            return;
        }
        int w = k;
        int pw = k;
        if (PIVOT) {
            uint64_t key = (valid && pos >= k) ? dev::PivotKey(a[k], pos, lane) : 0ull;
            key = dev::WaveMaxKey(key);
            w = dev::Uniform(dev::KeyRow(key));
            pw = dev::Uniform(dev::KeyPos(key));
            if (lane == k) {
                myPiv = pw + 1;
            }
        }
        const float pivot = dev::ReadLane(a[k], w);
        if (pivot == 0.0f) {
            infoVal = k + 1;
            running = false;
            return;
        }
        if (PIVOT) {
            if (pos == k) {
                pos = pw;
            }
            if (lane == w) {
                pos = k;
            }
        }
        const bool act = valid && pos > k;
        const float l = a[k] / pivot;
        if (act) {
            a[k] = l;
        }
        dev::StaticFor<k + 1, N>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            if (j < n) {
                const float u = dev::ReadLane(a[j], w);
                if (act) {
                    a[j] = fmaf(-l, u, a[j]);
                }
            }
        });
    });

    if (valid) {
        dev::StaticFor<0, N>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            if (j < n) {
                A[pos + static_cast<size_t>(j) * lda] = a[j];
            }
        });
    }
    if (PIVOT) {
        const int written = infoVal != 0 ? infoVal : n;
        if (lane < written) {
            ipiv[static_cast<size_t>(b) * n + lane] = myPiv;
        }
    }
    if (info != nullptr && lane == 0) {
        info[b] = infoVal;
    }
}

}  // namespace

template <bool PIVOT>
void LaunchGetrfWave(hipStream_t st, float* const* A, int n, int lda, int* ipiv, int* info, int batch)
{
    const dim3 grid((batch + kWavesPerBlock - 1) / kWavesPerBlock);
    const dim3 block(kWave * kWavesPerBlock);
    if (n <= 8) {
        hipLaunchKernelGGL((GetrfWaveKernel<8, PIVOT>), grid, block, 0, st, A, n, lda, ipiv, info, batch);
    } else if (n <= 16) {
        hipLaunchKernelGGL((GetrfWaveKernel<16, PIVOT>), grid, block, 0, st, A, n, lda, ipiv, info, batch);
    } else if (n <= 32) {
        hipLaunchKernelGGL((GetrfWaveKernel<32, PIVOT>), grid, block, 0, st, A, n, lda, ipiv, info, batch);
    } else {
        hipLaunchKernelGGL((GetrfWaveKernel<64, PIVOT>), grid, block, 0, st, A, n, lda, ipiv, info, batch);
    }
}

template void LaunchGetrfWave<true>(hipStream_t, float* const*, int, int, int*, int*, int);
template void LaunchGetrfWave<false>(hipStream_t, float* const*, int, int, int*, int*, int);

}  // namespace aclblas
