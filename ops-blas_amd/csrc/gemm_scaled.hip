// FP8 GEMM with per-tensor scales, and the per-tensor FP8 quantiser.
//
// Neither entry exists in the reference library.  The scale semantics follow
// its Lt interface (include/cann_ops_blasLt.h:160-200,
// ACLBLASLT_MATMUL_DESC_A_SCALE_POINTER / B_SCALE_POINTER: a device float that
// multiplies the operand), the FP8 formats and saturation follow
// test/frame/dtype_utils.h:39-160 (E4M3FN largest finite 448, E5M2 57344,
// round to nearest even, saturating).
//
// aclblasGemmExScaled runs the MFMA tile kernel with the native FP8 matrix
// instruction (v_mfma_f32_16x16x128_f8f6f4, E4M3/E5M2 in any combination) and
// folds scaleA * scaleB * alpha into the fused epilogue.
//
// aclblasQuantizeFp8 computes scale = amax(|x|) / fp8_max (1 for an all-zero
// x) and q = fp8(x * (fp8_max / amax)) in two launches: per-block maxima into
// the handle's workspace, then every block re-reduces those maxima (so no
// global atomics or zeroed counters are needed), quantises its part, and block
// 0 stores the scale.  Aligned operands move as float4 loads and packed
// 4-byte stores.
#include <mutex>

#include "cann_ops_blas.h"
#include "gemm_kernel.h"
#include "internal.h"
#include "tuning.h"

namespace aclblas {
namespace {

using namespace gemm;

constexpr int kQThreads = 256;
constexpr int kQBlocks = 1024;
constexpr int kMaxDevices = 64;

// Per-block maxima of aclblasQuantizeFp8NoAlloc: 4 KiB on every GPU the
// library runs on, held until the process ends.  Host code never passes this
// name to a kernel: it passes the device address that hipGetSymbolAddress
// returns for the current device (QuantScratch).
static __device__ __align__(256) float scratch[kQBlocks];

// Device address of `scratch` on the current device, looked up once per
// device; nullptr if the lookup fails.
float* QuantScratch(int device)
{
    static std::mutex mu;
    static float* ptrs[kMaxDevices] = {};
    static bool done[kMaxDevices] = {};
    if (device < 0 || device >= kMaxDevices) {
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(mu);
    if (!done[device]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        ptrs[device] = static_cast<float*>(p);
        done[device] = true;
    }
    return ptrs[device];
}

// float -> FP8 (E4M3FN when E4M3, else E5M2), round to nearest even,
// saturating to the largest finite value; NaN -> the format's NaN.
template <bool E4M3>
__device__ __forceinline__ uint8_t ToFp8(float x)
{
    constexpr int kMant = E4M3 ? 3 : 2;
    constexpr int kBias = E4M3 ? 7 : 15;
    constexpr uint32_t kMaxCode = E4M3 ? 0x7Eu : 0x7Bu;
    constexpr float kMax = E4M3 ? 448.0f : 57344.0f;
    const uint32_t bits = __float_as_uint(x);
    const uint32_t sign = (bits >> 24) & 0x80u;
    const float a = __uint_as_float(bits & 0x7FFFFFFFu);
    if (!(a == a)) {
        return static_cast<uint8_t>(sign | 0x7Fu);
    }
    if (a >= kMax) {
        return static_cast<uint8_t>(sign | kMaxCode);
    }
    int e;
    (void)frexpf(a, &e);
    e -= 1;  // a = 1.f * 2^e
    uint32_t code;
    if (a == 0.0f) {
        code = 0;
    } else if (e < 1 - kBias) {
        // subnormal: multiples of 2^(1 - bias - mant)
        code = static_cast<uint32_t>(__builtin_rintf(ldexpf(a, kBias - 1 + kMant)));
    } else {
        uint32_t m = static_cast<uint32_t>(__builtin_rintf(ldexpf(a, kMant - e)));  // in [2^mant, 2^(mant+1)]
        if (m == (1u << (kMant + 1))) {
            m >>= 1;
            e += 1;
        }
        code = (static_cast<uint32_t>(e + kBias) << kMant) | (m - (1u << kMant));
    }
    code = code > kMaxCode ? kMaxCode : code;
    return static_cast<uint8_t>(sign | code);
}

__device__ __forceinline__ float BlockMaxF(float v, float* red)
{
    for (int off = 32; off >= 1; off >>= 1) {
        const float o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    if (lane == 0) {
        red[wave] = v;
    }
    __syncthreads();
    float m = red[0];
    for (int q = 1; q < kQThreads / 64; ++q) {
        m = red[q] > m ? red[q] : m;
    }
    __syncthreads();
    return m;
}

// |x| maximum per block (NaN ignored) -> partial[blockIdx].  VEC: x is
// 16-byte aligned and is read four floats at a time (the count % 4 tail singly).
template <bool VEC>
__global__ __launch_bounds__(kQThreads) void AmaxKernel(const float* __restrict__ x, long long count,
                                                        float* __restrict__ partial)
{
    __shared__ float red[kQThreads / 64];
    float m = 0.0f;
    const long long stride = static_cast<long long>(kQThreads) * gridDim.x;
    const long long first = blockIdx.x * static_cast<long long>(kQThreads) + threadIdx.x;
    long long scalarFrom = 0;
    if (VEC) {
        const long long quads = count / 4;
        const float4* x4 = reinterpret_cast<const float4*>(x);
        for (long long i = first; i < quads; i += stride) {
            const float4 v = x4[i];
            const float a0 = fabsf(v.x), a1 = fabsf(v.y), a2 = fabsf(v.z), a3 = fabsf(v.w);
            m = a0 > m ? a0 : m;  // NaN compares false: ignored
            m = a1 > m ? a1 : m;
            m = a2 > m ? a2 : m;
            m = a3 > m ? a3 : m;
        }
        scalarFrom = quads * 4;
    }
    for (long long i = scalarFrom + first; i < count; i += stride) {
        const float a = fabsf(x[i]);
        m = a > m ? a : m;
    }
    m = BlockMaxF(m, red);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = m;
    }
}

// VEC: x 16-byte and q 4-byte aligned: four values per thread and step, one
// packed 4-byte store.
template <bool E4M3, bool VEC>
__global__ __launch_bounds__(kQThreads) void QuantizeKernel(const float* __restrict__ x, long long count,
                                                            const float* __restrict__ partial, int parts,
                                                            uint8_t* __restrict__ q, float* __restrict__ scale)
{
    __shared__ float red[kQThreads / 64];
    constexpr float kMax = E4M3 ? 448.0f : 57344.0f;
    float m = 0.0f;
    for (int i = threadIdx.x; i < parts; i += kQThreads) {
        m = partial[i] > m ? partial[i] : m;
    }
    const float amax = BlockMaxF(m, red);
    const float inv = amax > 0.0f ? kMax / amax : 1.0f;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *scale = amax > 0.0f ? amax / kMax : 1.0f;
    }
    const long long stride = static_cast<long long>(kQThreads) * gridDim.x;
    const long long first = blockIdx.x * static_cast<long long>(kQThreads) + threadIdx.x;
    long long scalarFrom = 0;
    if (VEC) {
        const long long quads = count / 4;
        const float4* x4 = reinterpret_cast<const float4*>(x);
        uint32_t* q4 = reinterpret_cast<uint32_t*>(q);
        for (long long i = first; i < quads; i += stride) {
            const float4 v = x4[i];
            q4[i] = static_cast<uint32_t>(ToFp8<E4M3>(v.x * inv)) |
                    (static_cast<uint32_t>(ToFp8<E4M3>(v.y * inv)) << 8) |
                    (static_cast<uint32_t>(ToFp8<E4M3>(v.z * inv)) << 16) |
                    (static_cast<uint32_t>(ToFp8<E4M3>(v.w * inv)) << 24);
        }
        scalarFrom = quads * 4;
    }
    for (long long i = scalarFrom + first; i < count; i += stride) {
        q[i] = ToFp8<E4M3>(x[i] * inv);
    }
}

template <bool E4M3, bool VEC>
void LaunchQuantize(hipStream_t st, dim3 grid, const float* x, long long count, float* partial, int parts,
                    uint8_t* q, float* scale)
{
    hipLaunchKernelGGL((AmaxKernel<VEC>), grid, dim3(kQThreads), 0, st, x, count, partial);
    hipLaunchKernelGGL((QuantizeKernel<E4M3, VEC>), grid, dim3(kQThreads), 0, st, x, count, partial, parts, q,
                       scale);
}

int ScaledKind(int t)
{
    if (t == ACL_FLOAT8_E4M3FN) return kE4M3S;
    if (t == ACL_FLOAT8_E5M2) return kE5M2S;
    return -1;
}

int OutKindScaled(int t)
{
    if (t == ACL_FLOAT) return kOutF32;
    if (t == ACL_BF16) return kOutBF16;
    if (t == ACL_FLOAT16) return kOutF16;
    return -1;
}

template <int KA, int KB>
void LaunchOut(hipStream_t st, int ok, const GemmArgs& p, int tile, int group)
{
    if (ok == kOutF32) {
        LaunchGemmTyped<KA, KB, kOutF32>(st, p, 1, tile, group);
    } else if (ok == kOutBF16) {
        LaunchGemmTyped<KA, KB, kOutBF16>(st, p, 1, tile, group);
    } else {
        LaunchGemmTyped<KA, KB, kOutF16>(st, p, 1, tile, group);
    }
}

}  // namespace
}  // namespace aclblas

extern "C" aclblasStatus_t aclblasGemmExScaled(
    aclblasHandle_t handle, aclblasOperation_t transA, aclblasOperation_t transB, int m, int n, int k,
    const float* alpha, const void* A, aclDataType Atype, int lda, const float* scaleA, const void* B,
    aclDataType Btype, int ldb, const float* scaleB, const float* beta, void* C, aclDataType Ctype, int ldc)
{
    using namespace aclblas;
    using namespace aclblas::gemm;
    if (handle == nullptr) {
        return ACLBLAS_STATUS_HANDLE_IS_NULLPTR;
    }
    if (!IsValidOp(transA) || !IsValidOp(transB) || m < 0 || n < 0 || k < 0 || alpha == nullptr || beta == nullptr) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    const bool ta = transA != ACLBLAS_OP_N;
    const bool tb = transB != ACLBLAS_OP_N;
    if (lda < MaxInt(1, ta ? k : m) || ldb < MaxInt(1, tb ? n : k) || ldc < MaxInt(1, m)) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    const int ka = ScaledKind(EnumBits(Atype));
    const int kb = ScaledKind(EnumBits(Btype));
    const int ok = OutKindScaled(EnumBits(Ctype));
    if (ka < 0 || kb < 0 || ok < 0) {
        return ACLBLAS_STATUS_NOT_SUPPORTED;
    }
    if (m == 0 || n == 0) {
        return ACLBLAS_STATUS_SUCCESS;
    }
    if (C == nullptr || (k > 0 && (A == nullptr || B == nullptr))) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (!HandleHasDevice(handle)) {
        return ACLBLAS_STATUS_EXECUTION_FAILED;
    }
    GemmArgs p{};
    p.A = A;
    p.B = B;
    p.C = C;
    p.m = m;
    p.n = n;
    p.k = k;
    p.lda = lda;
    p.ldb = ldb;
    p.ldc = ldc;
    p.transA = ta ? 1 : 0;
    p.transB = tb ? 1 : 0;
    p.alpha = *alpha;
    p.beta = *beta;
    p.scaleA = scaleA;
    p.scaleB = scaleB;
    p.vecA = (reinterpret_cast<uintptr_t>(A) % 16 == 0 && lda % 16 == 0) ? 1 : 0;
    p.vecB = (reinterpret_cast<uintptr_t>(B) % 16 == 0 && ldb % 16 == 0) ? 1 : 0;
    {
        const int cb = ok == kOutF32 ? 4 : 2;
        p.vecC = (reinterpret_cast<uintptr_t>(C) % (4 * cb) == 0 && (static_cast<long long>(ldc) * cb) % (4 * cb) == 0)
                     ? 1 : 0;
    }
    const int cus = handle->cu_count > 0 ? handle->cu_count : 256;
    const long long t128 = static_cast<long long>((m + 127) / 128) * ((n + 127) / 128);
    const tune::Choice c = tune::Get(handle, tune::kOpGemm, tune::GemmBucket(tune::kClassF8Scaled, t128, k, cus),
                                     tune::Choice{DefaultGemmTile(1, m, n, 1, cus), 8});
    hipStream_t st = handle->stream;
    if (ka == kE4M3S && kb == kE4M3S) {
        LaunchOut<kE4M3S, kE4M3S>(st, ok, p, c.a, c.b);
    } else if (ka == kE4M3S) {
        LaunchOut<kE4M3S, kE5M2S>(st, ok, p, c.a, c.b);
    } else if (kb == kE4M3S) {
        LaunchOut<kE5M2S, kE4M3S>(st, ok, p, c.a, c.b);
    } else {
        LaunchOut<kE5M2S, kE5M2S>(st, ok, p, c.a, c.b);
    }
    return LaunchStatus("aclblasGemmExScaled");
}

namespace aclblas {
namespace {

aclblasStatus_t CheckQuantizeArgs(aclblasHandle_t handle, long long count, const float* x, aclDataType qtype,
                                  const void* q, const float* scale)
{
    if (handle == nullptr) {
        return ACLBLAS_STATUS_HANDLE_IS_NULLPTR;
    }
    if (count < 0 || (qtype != ACL_FLOAT8_E4M3FN && qtype != ACL_FLOAT8_E5M2)) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (scale == nullptr || (count > 0 && (x == nullptr || q == nullptr))) {
        return ACLBLAS_STATUS_INVALID_VALUE;
    }
    if (!HandleHasDevice(handle)) {
        return ACLBLAS_STATUS_EXECUTION_FAILED;
    }
    return ACLBLAS_STATUS_SUCCESS;
}

// Both launches of the quantiser with the per-block maxima in `partial`
// (kQBlocks floats of device memory).
void LaunchQuantizer(hipStream_t stream, long long count, const float* x, aclDataType qtype, void* q, float* scale,
                     float* partial)
{
    long long blocks = (count / 4 + kQThreads - 1) / kQThreads;
    blocks = blocks < 1 ? 1 : (blocks > kQBlocks ? kQBlocks : blocks);
    const dim3 grid(static_cast<unsigned>(blocks));
    const bool vec = reinterpret_cast<uintptr_t>(x) % 16 == 0 && reinterpret_cast<uintptr_t>(q) % 4 == 0;
    uint8_t* q8 = static_cast<uint8_t*>(q);
    const int parts = static_cast<int>(blocks);
    if (qtype == ACL_FLOAT8_E4M3FN) {
        if (vec) {
            LaunchQuantize<true, true>(stream, grid, x, count, partial, parts, q8, scale);
        } else {
            LaunchQuantize<true, false>(stream, grid, x, count, partial, parts, q8, scale);
        }
    } else {
        if (vec) {
            LaunchQuantize<false, true>(stream, grid, x, count, partial, parts, q8, scale);
        } else {
            LaunchQuantize<false, false>(stream, grid, x, count, partial, parts, q8, scale);
        }
    }
}

}  // namespace
}  // namespace aclblas

extern "C" aclblasStatus_t aclblasQuantizeFp8(aclblasHandle_t handle, long long count, const float* x,
                                              aclDataType qtype, void* q, float* scale)
{
    using namespace aclblas;
    aclblasStatus_t st = CheckQuantizeArgs(handle, count, x, qtype, q, scale);
    if (st != ACLBLAS_STATUS_SUCCESS) {
        return st;
    }
    st = EnsureWorkspace(handle, kQBlocks * sizeof(float));
    if (st != ACLBLAS_STATUS_SUCCESS) {
        return st;
    }
    LaunchQuantizer(handle->stream, count, x, qtype, q, scale, static_cast<float*>(handle->workspace));
    return LaunchStatus("aclblasQuantizeFp8");
}

// The quantiser without the handle's workspace: the per-block maxima live in
// the file-scope device array `scratch` (kQBlocks floats per GPU), reached
// through its device address.  Nothing is allocated, so the call can be made
// where allocation is not allowed.  Every call on one GPU shares that array
// (the first launch writes every entry the second reads, so nothing needs
// clearing); calls that overlap on one GPU race.
extern "C" aclblasStatus_t aclblasQuantizeFp8NoAlloc(aclblasHandle_t handle, long long count, const float* x,
                                                     aclDataType qtype, void* q, float* scale)
{
    using namespace aclblas;
    aclblasStatus_t st = CheckQuantizeArgs(handle, count, x, qtype, q, scale);
    if (st != ACLBLAS_STATUS_SUCCESS) {
        return st;
    }
    int device = -1;
    if (hipGetDevice(&device) != hipSuccess) {
        (void)hipGetLastError();
        return ACLBLAS_STATUS_EXECUTION_FAILED;
    }
    float* partial = QuantScratch(device);
    if (partial == nullptr) {
        LogError("aclblasQuantizeFp8NoAlloc", "device address of the quantiser's scratch array not found on device %d",
                 device);
        return ACLBLAS_STATUS_EXECUTION_FAILED;
    }
    LaunchQuantizer(handle->stream, count, x, qtype, q, scale, partial);
    return LaunchStatus("aclblasQuantizeFp8NoAlloc");
}
