// Batched inverse from LU factors for n <= 64 (aclblasSgetriBatched, and the
// inverse step of aclblasSmatinvBatched).  See getrs.hip for the reference and
// the general kernel.
//
// One wave per matrix and one lane per column of the inverse: lane c carries
// column c of C = U^-1 L^-1 P I in registers through the whole computation.
// The factors are the same for every column, so they sit in LDS (zero-padded
// to S x S, S = n rounded up to 16, 32 or 64) and each substitution step reads
// its column of L or U with broadcast ds_read_b128 (every lane the same
// address: no bank conflicts, no cross-lane moves); the divisions by u_jj run
// in all columns at once.  Per element the operations and their order are
// those of the general kernel (column-oriented substitution,
// c_i = fma(-t_ij, c_j, c_i) for j in solve order, c_j = c_j / u_jj), so both
// give identical results.  Rows >= n are padding: their factors are zero and
// they are never read by a row < n nor stored.  With a zero u_jj the info is
// the largest such j + 1 and the U solve is skipped, as in the general kernel
// (reference getri_batched_kernel.cpp:142-151).  The result goes back through
// LDS so that the global stores are column-contiguous.
#include "device_common.h"
#include "internal.h"
#include "lapack_kernels.h"

namespace aclblas {
namespace {

using dev::kWave;

// waves (matrices) per workgroup: LDS per wave is S (S + 1) floats
template <int S>
constexpr int kWavesPerBlock = S == 64 ? 2 : 4;

// LDS writes of this wave visible to all of its lanes
__device__ __forceinline__ void WaveSync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int S>
__global__ __launch_bounds__(kWave* kWavesPerBlock<S>) void GetriColumnKernel(
    const float* const* __restrict__ Aarray, int n, int lda, const int* __restrict__ ipiv,
    float* const* __restrict__ Carray, int ldc, int* __restrict__ info, int batch)
{
    constexpr int kLd = S + 1;  // row length of the output transpose (odd: conflict-free)
    __shared__ __attribute__((aligned(16))) float lds[kWavesPerBlock<S>][S * kLd];
    const int lane = dev::LaneId();
    const int wave = static_cast<int>(threadIdx.x) / kWave;
    const int b = blockIdx.x * kWavesPerBlock<S> + wave;
    if (b >= batch) {
        return;  // the whole wave: nothing below waits for other waves
    }
    float* t = static_cast<float*>(__builtin_assume_aligned(lds[wave], 16));

    // factors, zero-padded, column-major with leading dimension S
    // (all loads unconditional, at element 0 where the entry is padding, so
    // that they issue together; the padding is applied when storing to LDS)
    {
        const auto A = dev::G(Aarray[b]);
        float v[S];
        dev::StaticFor<0, S>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            v[j] = A[(lane < n && j < n) ? lane + static_cast<size_t>(j) * lda : 0];
        });
        __builtin_amdgcn_sched_barrier(0);
        dev::StaticFor<0, S>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            if (lane < S) {
                t[j * S + lane] = (lane < n && j < n) ? v[j] : 0.0f;
            }
        });
    }

    // row at which column c's unit entry lands after the interchanges k <-> piv[k] - 1,
    // applied for k = 0 .. n-1 (entries outside [1, n] skipped)
    int pos = lane;
    if (ipiv != nullptr) {
        const int pv = dev::G(ipiv)[static_cast<size_t>(b) * n + (lane < n ? lane : 0)];
        const int myp = lane < n ? pv - 1 : -1;
        for (int k = 0; k < n; ++k) {
            const int pk = __builtin_amdgcn_readlane(myp, k);
            if (pk >= 0 && pk < n && pk != k) {
                pos = pos == k ? pk : (pos == pk ? k : pos);
            }
        }
    }
    float x[S];
    dev::StaticFor<0, S>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        x[i] = i == pos ? 1.0f : 0.0f;
    });
    WaveSync();

    const float diag = lane < n ? t[lane * S + lane] : 1.0f;
    const uint64_t zmask = __ballot(diag == 0.0f);
    const int zeroDiag = zmask != 0 ? 64 - __builtin_clzll(zmask) : 0;

    // L (unit lower): c_i -= l_ij c_j for i > j, j ascending
    dev::StaticFor<0, S - 1>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        if (j < n) {
            const float xj = x[j];
            dev::StaticFor<(j + 1) / 4, S / 4>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                const float4 v = *reinterpret_cast<const float4*>(t + j * S + 4 * q);
                const float l[4] = {v.x, v.y, v.z, v.w};
                dev::StaticFor<0, 4>([&](auto ec) {
                    constexpr int i = 4 * q + decltype(ec)::value;
                    if (i > j) {
                        x[i] = fmaf(-l[i - 4 * q], xj, x[i]);
                    }
                });
            });
        }
    });
    // U: c_j /= u_jj, then c_i -= u_ij c_j for i < j, j descending
    if (zeroDiag == 0) {
        dev::StaticFor<0, S>([&](auto jc) {
            constexpr int j = S - 1 - decltype(jc)::value;
            if (j < n) {
                x[j] = x[j] / t[j * S + j];
                const float xj = x[j];
                dev::StaticFor<0, (j + 3) / 4>([&](auto qc) {
                    constexpr int q = decltype(qc)::value;
                    const float4 v = *reinterpret_cast<const float4*>(t + j * S + 4 * q);
                    const float u[4] = {v.x, v.y, v.z, v.w};
                    dev::StaticFor<0, 4>([&](auto ec) {
                        constexpr int i = 4 * q + decltype(ec)::value;
                        if (i < j) {
                            x[i] = fmaf(-u[i - 4 * q], xj, x[i]);
                        }
                    });
                });
            }
        });
    }

    // transpose through LDS (every factor read is done), then column-contiguous stores
    WaveSync();
    if (lane < S) {
        dev::StaticFor<0, S>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            t[lane * kLd + i] = x[i];
        });
    }
    WaveSync();
    float* C = Carray[b];
    for (int c = 0; c < n; ++c) {
        if (lane < n) {
            C[lane + static_cast<size_t>(c) * ldc] = t[c * kLd + lane];
        }
    }
    if (lane == 0) {
        info[b] = zeroDiag;
    }
}

template <int S>
void Launch(hipStream_t st, const float* const* A, int n, int lda, const int* piv, float* const* C, int ldc,
            int* info, int batch)
{
    const dim3 grid((batch + kWavesPerBlock<S> - 1) / kWavesPerBlock<S>);
    hipLaunchKernelGGL((GetriColumnKernel<S>), grid, dim3(kWave * kWavesPerBlock<S>), 0, st, A, n, lda, piv, C, ldc,
                       info, batch);
}

}  // namespace

bool LaunchGetriSmall(hipStream_t st, int n, const float* const* A, int lda, const int* piv, float* const* C,
                      int ldc, int* info, int batch)
{
    if (n <= 0 || n > kWave) {
        return false;
    }
    if (n <= 16) {
        Launch<16>(st, A, n, lda, piv, C, ldc, info, batch);
    } else if (n <= 32) {
        Launch<32>(st, A, n, lda, piv, C, ldc, info, batch);
    } else {
        Launch<64>(st, A, n, lda, piv, C, ldc, info, batch);
    }
    return true;
}

}  // namespace aclblas
