// Batched triangular solves with LU factors for n <= 64 (aclblasSgetrsBatched,
// trans N and T).  See getrs.hip for the reference and the general kernel.
//
// One lane per row, 64/S systems per wave, S = n rounded up to 16, 32 or 64.
// Lane r of a segment keeps, in registers, row r of the packed LU factors for
// trans N, or column r for trans T: exactly the factor entries that its
// unknown needs in a column-oriented substitution.  The right-hand sides of a
// chunk (RC columns of B) sit in registers too, so each substitution step is
// one broadcast of the solved value inside the segment (DPP row_newbcast for
// S = 16, plus v_permlane16_swap for S = 32, v_readlane for S = 64) and one
// fmaf per lane and right-hand side: no
// LDS, no barriers.  Row interchanges: every lane traces its final position
// back through the swap list to find which row of B it loads (trans N) or
// stores (trans T); for trans N each lane loads its own row and the rows are
// then permuted inside the segment with ds_bpermute, so that the load of B does
// not wait for the pivots.  Element by element the operations and their order are
// those of the general kernel, so both give identical results.
#include "device_common.h"
#include "internal.h"
#include "lapack_kernels.h"

namespace aclblas {
namespace {

using dev::kWave;

constexpr int kWavesPerBlock = 4;

// Value of lane j (compile-time) of this lane's S-lane segment.
template <int S, int J>
__device__ __forceinline__ float SegBroadcast(float v)
{
    if constexpr (S == 16) {
        return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x150 + J, 0xF, 0xF, false));
    } else if constexpr (S == 32) {
        // lane J % 16 of every 16-lane row, then the row holding J of each
        // 32-lane segment to both of its rows (v_permlane16_swap)
        const uint32_t t = static_cast<uint32_t>(
            __builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x150 + (J & 15), 0xF, 0xF, false));
        const auto p = __builtin_amdgcn_permlane16_swap(t, t, false, false);
        return __uint_as_float(J < 16 ? p[0] : p[1]);
    } else {
        return dev::ReadLane(v, J);
    }
}

template <int S, int RC, bool TRANS>
__global__ __launch_bounds__(kWave* kWavesPerBlock) void SolveSmallKernel(
    const float* const* __restrict__ Aarray, int n, int lda, const int* __restrict__ ipiv,
    float* const* __restrict__ Barray, int ldb, int nrhs, int batch)
{
    constexpr int kMpw = kWave / S;
    const int lane = dev::LaneId();
    const int wave = static_cast<int>(threadIdx.x) / kWave;
    const int r = lane % S;
    const int b = (blockIdx.x * kWavesPerBlock + wave) * kMpw + lane / S;
    const int c0 = blockIdx.y * RC;
    const bool live = b < batch;
    const bool real = live && r < n;
    const int rc = min(RC, nrhs - c0);

    // Every load below is unconditional, at element 0 of the lane's (or the
    // first) problem where the element is not needed, and its value is
    // replaced by the padding afterwards: the loads then issue together
    // instead of each under a branch with its own wait.
    // factor entries: row r (trans N) or column r (trans T), zero padding
    const auto A = dev::G(Aarray[live ? b : 0]);
    float a[S];
    dev::StaticFor<0, S>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        const size_t idx = TRANS ? j + static_cast<size_t>(r) * lda : r + static_cast<size_t>(j) * lda;
        a[j] = A[(real && j < n) ? idx : 0];
    });

    // this lane's own row of B (trans N: permuted below, inside the segment, so
    // the load does not wait for the pivots)
    const auto B = dev::G(Barray[live ? b : 0]);
    float x[RC];
    dev::StaticFor<0, RC>([&](auto cc) {
        constexpr int c = decltype(cc)::value;
        x[c] = B[(real && c < rc) ? r + static_cast<size_t>(c0 + c) * ldb : 0];
    });

    // pivots of the lane's problem
    int pv[S];
    if (ipiv != nullptr) {
        const auto p = dev::G(ipiv + static_cast<size_t>(live ? b : 0) * n);
        dev::StaticFor<0, S>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            pv[k] = p[k < n ? k : 0];
        });
    }
    __builtin_amdgcn_sched_barrier(0);
    dev::StaticFor<0, S>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        a[j] = (real && j < n) ? a[j] : 0.0f;
    });
    dev::StaticFor<0, RC>([&](auto cc) {
        constexpr int c = decltype(cc)::value;
        x[c] = (real && c < rc) ? x[c] : 0.0f;
    });

    // row of B that lands at position r after the interchanges
    int src = r;
    if (ipiv != nullptr && real) {
        dev::StaticFor<0, S>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            pv[k] = k < n ? pv[k] - 1 : -1;
        });
        dev::StaticFor<0, S>([&](auto kc) {
            constexpr int k = S - 1 - decltype(kc)::value;
            const int pk = pv[k];
            if (pk >= 0 && pk < n && pk != k) {
                src = src == k ? pk : (src == pk ? k : src);
            }
        });
    }

    if (!TRANS) {
        // x <- (P b): the value of lane src of this segment (every lane active here)
        const int from = (lane - r + src) * 4;
        dev::StaticFor<0, RC>([&](auto cc) {
            constexpr int c = decltype(cc)::value;
            x[c] = __int_as_float(__builtin_amdgcn_ds_bpermute(from, __float_as_int(x[c])));
        });
    }

    if (!TRANS) {
        // L (unit lower): x_i -= l_ij x_j for i > j, j ascending
        dev::StaticFor<0, S>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            if (j < n) {
                dev::StaticFor<0, RC>([&](auto cc) {
                    constexpr int c = decltype(cc)::value;
                    const float xj = SegBroadcast<S, j>(x[c]);
                    if (r > j) {
                        x[c] = fmaf(-a[j], xj, x[c]);
                    }
                });
            }
        });
        // U: x_j /= u_jj, then x_i -= u_ij x_j for i < j, j descending
        dev::StaticFor<0, S>([&](auto jc) {
            constexpr int j = S - 1 - decltype(jc)::value;
            if (j < n) {
                dev::StaticFor<0, RC>([&](auto cc) {
                    constexpr int c = decltype(cc)::value;
                    if (r == j) {
                        x[c] = x[c] / a[j];
                    }
                    const float xj = SegBroadcast<S, j>(x[c]);
                    if (r < j) {
                        x[c] = fmaf(-a[j], xj, x[c]);
                    }
                });
            }
        });
    } else {
        // U^T (lower): y_j /= u_jj, then y_i -= u_ji y_j for i > j, j ascending
        dev::StaticFor<0, S>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            if (j < n) {
                dev::StaticFor<0, RC>([&](auto cc) {
                    constexpr int c = decltype(cc)::value;
                    if (r == j) {
                        x[c] = x[c] / a[j];
                    }
                    const float xj = SegBroadcast<S, j>(x[c]);
                    if (r > j) {
                        x[c] = fmaf(-a[j], xj, x[c]);
                    }
                });
            }
        });
        // L^T (unit upper): z_i -= l_ji z_j for i < j, j descending
        dev::StaticFor<0, S>([&](auto jc) {
            constexpr int j = S - 1 - decltype(jc)::value;
            if (j < n) {
                dev::StaticFor<0, RC>([&](auto cc) {
                    constexpr int c = decltype(cc)::value;
                    const float xj = SegBroadcast<S, j>(x[c]);
                    if (r < j) {
                        x[c] = fmaf(-a[j], xj, x[c]);
                    }
                });
            }
        });
    }

    if (real) {
        const int row = TRANS ? src : r;
        dev::StaticFor<0, RC>([&](auto cc) {
            constexpr int c = decltype(cc)::value;
            if (c < rc) {
                B[row + static_cast<size_t>(c0 + c) * ldb] = x[c];
            }
        });
    }
}

template <int S, int RC, bool TRANS>
void Launch(hipStream_t st, const float* const* A, int n, int lda, const int* piv, float* const* B, int ldb,
            int nrhs, int batch)
{
    constexpr int kPerBlock = kWavesPerBlock * (kWave / S);
    const dim3 grid((batch + kPerBlock - 1) / kPerBlock, (nrhs + RC - 1) / RC);
    hipLaunchKernelGGL((SolveSmallKernel<S, RC, TRANS>), grid, dim3(kWave * kWavesPerBlock), 0, st, A, n, lda, piv,
                       B, ldb, nrhs, batch);
}

template <int S, bool TRANS>
void LaunchRc(hipStream_t st, const float* const* A, int n, int lda, const int* piv, float* const* B, int ldb,
              int nrhs, int batch)
{
    if (nrhs == 1) {
        Launch<S, 1, TRANS>(st, A, n, lda, piv, B, ldb, nrhs, batch);
    } else if (nrhs == 2) {
        Launch<S, 2, TRANS>(st, A, n, lda, piv, B, ldb, nrhs, batch);
    } else if (nrhs <= 4) {
        Launch<S, 4, TRANS>(st, A, n, lda, piv, B, ldb, nrhs, batch);
    } else {
        Launch<S, 8, TRANS>(st, A, n, lda, piv, B, ldb, nrhs, batch);
    }
}

template <bool TRANS>
void LaunchSmall(hipStream_t st, const float* const* A, int n, int lda, const int* piv, float* const* B, int ldb,
                 int nrhs, int batch)
{
    if (n <= 16) {
        LaunchRc<16, TRANS>(st, A, n, lda, piv, B, ldb, nrhs, batch);
    } else if (n <= 32) {
        LaunchRc<32, TRANS>(st, A, n, lda, piv, B, ldb, nrhs, batch);
    } else {
        LaunchRc<64, TRANS>(st, A, n, lda, piv, B, ldb, nrhs, batch);
    }
}

}  // namespace

bool LaunchGetrsSmall(hipStream_t st, bool trans, int n, int nrhs, const float* const* A, int lda, const int* piv,
                      float* const* B, int ldb, int batch)
{
    if (n > kWave || nrhs <= 0) {
        return false;
    }
    if (trans) {
        LaunchSmall<true>(st, A, n, lda, piv, B, ldb, nrhs, batch);
    } else {
        LaunchSmall<false>(st, A, n, lda, piv, B, ldb, nrhs, batch);
    }
    return true;
}

}  // namespace aclblas
