// Launchers of the batched LAPACK kernels, shared between translation units
// (kernels are split across files so they compile in parallel).
#pragma once

#include <hip/hip_runtime.h>

namespace aclblas {

// LU, n <= 64 (getrf_wave.hip); rows: matrix rows per lane (1 or 2; 1 for n <= 8)
template <bool PIVOT>
void LaunchGetrfWave(hipStream_t st, float* const* A, int n, int lda, int* ipiv, int* info, int batch, int rows);

// Pivoted LU, n <= 32, out of place (getrf_wave.hip): reads A, writes the
// factors to out + b n^2 (leading dimension n), &out[b n^2] to table[b] and
// all n pivots of each matrix (0 after a zero pivot).
void LaunchGetrfWaveOutOfPlace(hipStream_t st, const float* const* A, int n, int lda, float* out, float** table,
                               int* ipiv, int* info, int batch, int rows);

// LU, 64 < n <= 256, blocked (getrf_blocked.hip); grid: persistent workgroups
template <bool PIVOT>
void LaunchGetrfBlocked(hipStream_t st, float* const* A, int n, int lda, int* ipiv, int* info, int batch, int grid);

// Triangular solves, n <= 64 (getrs_small.hip); false when the shape is not covered.
bool LaunchGetrsSmall(hipStream_t st, bool trans, int n, int nrhs, const float* const* A, int lda, const int* piv,
                      float* const* B, int ldb, int batch);

// Inverse C = A^-1 from LU factors, n <= 64 (getrs_small.hip): info[b] = j+1
// for the zero u_jj with the largest j; false when the shape is not covered.
bool LaunchGetriSmall(hipStream_t st, int n, const float* const* A, int lda, const int* piv, float* const* C,
                      int ldc, int* info, int batch);

}  // namespace aclblas
