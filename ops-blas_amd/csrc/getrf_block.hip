// Batched LU for 64 < n <= 256: one workgroup per matrix, block-cyclic
// register ownership (see getrf.hip for the algorithm and its reference).
#include "device_common.h"
#include "internal.h"
#include "lapack_kernels.h"

namespace aclblas {
namespace {

using dev::kWave;

// ---------------------------------------------------------------------------
// 64 < n <= 256: one workgroup per matrix, block-cyclic registers
// ---------------------------------------------------------------------------
// Thread t owns row group g = t % RG (rows g + RG*r, r < RPT) and column group
// s = t / RG (columns s + TPR*c, c < CPT).  RG is a multiple of 64 so each wave
// belongs to a single column group.
template <int RG, int RPT, int TPR, int CPT, bool PIVOT>
__global__ __launch_bounds__(RG* TPR) void GetrfBlockKernel(
    float* const* __restrict__ Aarray, int n, int lda, int* __restrict__ ipiv, int* __restrict__ info, int batch)
{
    constexpr int kRows = RG * RPT;
    constexpr int kCols = TPR * CPT;
    constexpr int kWavesPerGroup = RG / kWave;
    static_assert(RG % kWave == 0, "row groups must fill whole waves");

    __shared__ float ubuf[kCols];           // pivot row, column-group-major
    __shared__ float lbuf[kRows];           // multipliers by row index
    __shared__ uint64_t pkey[2][kWavesPerGroup];
    __shared__ float pval[2][kWavesPerGroup];

    const int b = blockIdx.x;
    if (b >= batch) {
        return;
    }
    float* A = Aarray[b];
    const int t = threadIdx.x;
    const int g = t % RG;
    const int s = t / RG;
    const int lane = dev::LaneId();
    const int waveInGroup = g / kWave;

    float a[RPT][CPT];
    int pos[RPT];
    dev::StaticFor<0, RPT>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        const int i = g + RG * r;
        pos[r] = i;
        dev::StaticFor<0, CPT>([&](auto cc) {
            constexpr int c = decltype(cc)::value;
            const int j = s + TPR * c;
            a[r][c] = (i < n && j < n) ? A[i + static_cast<size_t>(j) * lda] : 0.0f;
        });
    });

    int infoVal = 0;
    int parity = 0;
    bool running = true;

    dev::StaticFor<0, CPT>([&](auto cc) {
        constexpr int c = decltype(cc)::value;
        for (int sk = 0; sk < TPR; ++sk) {
            const int k = c * TPR + sk;
            if (!running || k >= n) {
                return;
            }
            // ---- phase 1: pivot candidates of column k ----
            if (s == sk) {
                if (PIVOT) {
                    uint64_t key = 0;
                    float val = 0.0f;
                    dev::StaticFor<0, RPT>([&](auto rc) {
                        constexpr int r = decltype(rc)::value;
                        const int i = g + RG * r;
                        if (i < n && pos[r] >= k) {
                            const uint64_t kk = dev::PivotKey(a[r][c], pos[r], i);
                            if (kk > key) {
                                key = kk;
                                val = a[r][c];
                            }
                        }
                    });
                    const uint64_t best = dev::WaveMaxKey(key);
                    // the lane holding the winner publishes the signed value
                    if (key == best && best != 0ull) {
                        pval[parity][waveInGroup] = val;
                    }
                    if (lane == 0) {
                        pkey[parity][waveInGroup] = best;
                    }
                } else {
                    // no pivoting: the row at position k (row k) is the pivot
                    dev::StaticFor<0, RPT>([&](auto rc) {
                        constexpr int r = decltype(rc)::value;
                        if (g + RG * r == k) {
                            pval[parity][0] = a[r][c];
                        }
                    });
                }
            }
            __syncthreads();

            // ---- phase 2: winner, pivot row and multipliers ----
            int w = k;
            int pw = k;
            float pivot;
            if (PIVOT) {
                uint64_t best = pkey[parity][0];
                int bi = 0;
                for (int q = 1; q < kWavesPerGroup; ++q) {
                    const uint64_t o = pkey[parity][q];
                    if (o > best) {
                        best = o;
                        bi = q;
                    }
                }
                pivot = pval[parity][bi];
                w = dev::KeyRow(best);
                pw = dev::KeyPos(best);
                if (t == 0) {
                    ipiv[static_cast<size_t>(b) * n + k] = pw + 1;
                }
            } else {
                pivot = pval[parity][0];
            }
            parity ^= 1;
            if (pivot == 0.0f) {
                infoVal = k + 1;
                running = false;
                return;
            }
            dev::StaticFor<0, RPT>([&](auto rc) {
                constexpr int r = decltype(rc)::value;
                const int i = g + RG * r;
                if (PIVOT) {
                    if (pos[r] == k) {
                        pos[r] = pw;
                    }
                    if (i == w) {
                        pos[r] = k;
                    }
                }
                if (i == w) {
                    dev::StaticFor<c, CPT>([&](auto c2c) {
                        constexpr int c2 = decltype(c2c)::value;
                        ubuf[s * CPT + c2] = a[r][c2];
                    });
                }
                if (s == sk && i < n && pos[r] > k) {
                    const float l = a[r][c] / pivot;
                    a[r][c] = l;
                    lbuf[i] = l;
                }
            });
            __syncthreads();

            // ---- phase 3: rank-1 update of the owned trailing elements ----
            dev::StaticFor<0, RPT>([&](auto rc) {
                constexpr int r = decltype(rc)::value;
                const int i = g + RG * r;
                if (i < n && pos[r] > k) {
                    const float l = lbuf[i];
                    if (s > sk) {
                        a[r][c] = fmaf(-l, ubuf[s * CPT + c], a[r][c]);
                    }
                    dev::StaticFor<c + 1, CPT>([&](auto c2c) {
                        constexpr int c2 = decltype(c2c)::value;
                        a[r][c2] = fmaf(-l, ubuf[s * CPT + c2], a[r][c2]);
                    });
                }
            });
        }
    });
    // All reads of A happened before the first barrier; rows go to their
    // final positions.
    dev::StaticFor<0, RPT>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        const int i = g + RG * r;
        if (i < n) {
            dev::StaticFor<0, CPT>([&](auto cc) {
                constexpr int c = decltype(cc)::value;
                const int j = s + TPR * c;
                if (j < n) {
                    A[pos[r] + static_cast<size_t>(j) * lda] = a[r][c];
                }
            });
        }
    });
    if (info != nullptr && t == 0) {
        info[b] = infoVal;
    }
}

}  // namespace

template <bool PIVOT>
void LaunchGetrfBlock(hipStream_t st, float* const* A, int n, int lda, int* ipiv, int* info, int batch)
{
    if (n <= 128) {
        // 64 row groups x 2 rows, 4 column groups x 32 columns: 256 threads.
        hipLaunchKernelGGL((GetrfBlockKernel<64, 2, 4, 32, PIVOT>), dim3(batch), dim3(256), 0, st, A, n, lda, ipiv,
                           info, batch);
    } else {
        // 128 row groups x 2 rows, 8 column groups x 32 columns: 1024 threads.
        hipLaunchKernelGGL((GetrfBlockKernel<128, 2, 8, 32, PIVOT>), dim3(batch), dim3(1024), 0, st, A, n, lda,
                           ipiv, info, batch);
    }
}

template void LaunchGetrfBlock<true>(hipStream_t, float* const*, int, int, int*, int*, int);
template void LaunchGetrfBlock<false>(hipStream_t, float* const*, int, int, int*, int*, int);

}  // namespace aclblas
