// Batched LU for 64 < n <= 256: blocked right-looking LU, one 256-thread
// workgroup per matrix (persistent over the batch), in place in the caller's
// matrices, LAPACK output format.
//
// Reference: blas/getrf_batched/arch35/sgetrf_batched_kernel.cpp:106-150
// (pivot rule, division multipliers, stop at the first exact zero pivot) with
// LAPACK sgetrf's blocked algorithm, the reference tests' own golden
// (test/getrf_batched/sgetrf_batched/sgetrf_batched_golden.h:77).
//
// Per 32-column panel (lu_blocked.h): the panel is factored in registers by the
// four waves (one barrier per column), written back in pivoted order, and its
// pivot sequence goes to ipiv; inv(L11) is formed once; the trailing columns
// are swept in 16-column tiles (U12 = inv(L11) A12 and A22 -= L21 U12 on the
// FP32 matrix instruction, rows gathered through the panel's row map); the
// columns left of the panel get the panel's row interchanges.  At an exact
// zero pivot the steps before it are applied to the trailing columns, the
// panel's pivots up to and including that column are written, info = column
// + 1, and the matrix is left there (the reference's stopping state).
// Without a pivot array the same kernel runs with the row at position k as
// the only candidate.  Same multipliers (a / p) as the unblocked kernels; the
// blocked trailing update sums in a different order, so factors agree with
// the unblocked order and LAPACK to FP32 accuracy, pivots exactly on tie-free
// columns.
#include "device_common.h"
#include "internal.h"
#include "lapack_kernels.h"
#include "lu_blocked.h"

namespace aclblas {
namespace {

using lub::kNb;
using lub::kThreads;
using lub::kWaves;

// Rows k0.. of columns 0..k0-1 take the panel's interchanges: position q
// receives the row the panel's map names for it.  Wave v permutes columns
// v, v + 4, ...; a lane reads all its rows of a column before writing any.
__device__ __forceinline__ void PermuteLeft(float* A, int ld, int n, int k0, const int* map)
{
    const int lane = dev::LaneId();
    const int wave = static_cast<int>(threadIdx.x) / dev::kWave;
    const int rows = n - k0;
    for (int c = wave; c < k0; c += kWaves) {
        float* col = A + static_cast<size_t>(c) * ld + k0;
        float v[4];
        dev::StaticFor<0, 4>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            const int q = lane + 64 * i;
            v[i] = q < rows ? col[map[q]] : 0.0f;
        });
        dev::StaticFor<0, 4>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            const int q = lane + 64 * i;
            if (q < rows) {
                col[q] = v[i];
            }
        });
    }
}

template <bool PIVOT>
__global__ __launch_bounds__(kThreads, 2) void GetrfBlockedKernel(float* const* __restrict__ Aarray, int n, int lda,
                                                                 int* __restrict__ ipiv, int* __restrict__ info,
                                                                 int batch)
{
    __shared__ __attribute__((aligned(16))) lub::FactorShared<0> sh;
    const int t = threadIdx.x;
    for (int b = blockIdx.x; b < batch; b += gridDim.x) {
        float* const Ab = Aarray[b];
        const lub::PanelView view{Ab, lda, n, Ab, lda};
        int infoVal = 0;
        for (int k0 = 0; k0 < n; k0 += kNb) {
            int wl = 0;
            const int z = lub::FactorPanel<0, PIVOT, true>(view, k0, nullptr, sh, &wl);
            if (PIVOT) {
                const int count = z != 0 ? wl + 1 : wl;
                if (t < count) {
                    ipiv[static_cast<size_t>(b) * n + k0 + t] = k0 + sh.pw[t] + 1;
                }
            }
            lub::PanelInverses<0>(view, k0, wl, nullptr, sh);
            __syncthreads();
            lub::SweepTrailing<0>(view, k0, sh);
            if (PIVOT) {
                PermuteLeft(view.LU, lda, n, k0, sh.rowmap);
            }
            __syncthreads();
            if (z != 0) {
                infoVal = z;
                break;
            }
        }
        if (info != nullptr && t == 0) {
            info[b] = infoVal;
        }
    }
}

}  // namespace

template <bool PIVOT>
void LaunchGetrfBlocked(hipStream_t st, float* const* A, int n, int lda, int* ipiv, int* info, int batch, int grid)
{
    hipLaunchKernelGGL((GetrfBlockedKernel<PIVOT>), dim3(grid), dim3(kThreads), 0, st, A, n, lda, ipiv, info, batch);
}

template void LaunchGetrfBlocked<true>(hipStream_t, float* const*, int, int, int*, int*, int, int);
template void LaunchGetrfBlocked<false>(hipStream_t, float* const*, int, int, int*, int*, int, int);

}  // namespace aclblas
