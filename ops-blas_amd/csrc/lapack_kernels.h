// Launchers of the batched LAPACK kernels, shared between translation units
// (kernels are split across files so they compile in parallel).
#pragma once

#include <hip/hip_runtime.h>

namespace aclblas {

// LU, n <= 64 (getrf_wave.hip)
template <bool PIVOT>
void LaunchGetrfWave(hipStream_t st, float* const* A, int n, int lda, int* ipiv, int* info, int batch);

// LU, 64 < n <= 256 (getrf_block.hip)
template <bool PIVOT>
void LaunchGetrfBlock(hipStream_t st, float* const* A, int n, int lda, int* ipiv, int* info, int batch);

}  // namespace aclblas
