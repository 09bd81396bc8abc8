// Device-side helpers shared by the batched LAPACK kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace aclblas {
namespace dev {

constexpr int kWave = 64;

// Partial-pivoting selection key.  The pivot of column k is the candidate row
// with the largest |a| (NaN counts as 0, so NaN never wins), ties resolved to
// the lowest current row position; when every candidate is 0 or NaN the row
// already at position k is kept.  This is LAPACK isamax order applied to the
// permuted column and agrees with the reference's strict-'>' scan on every
// tie-free input (reference blas/getrf_batched/arch35/sgetrf_batched_kernel.cpp:52-79).
//
// key = bits(|a|) << 32 | (0xFFFF - pos) << 16 | row ; 0 marks "not a candidate".
// For non-negative floats the IEEE bit pattern orders like the value, so the
// unsigned maximum of the keys is the pivot.
__device__ __forceinline__ uint64_t PivotKey(float v, int pos, int row)
{
    float a = fabsf(v);
    uint32_t hi = (a == a) ? __float_as_uint(a) : 0u;
    uint32_t lo = (static_cast<uint32_t>(0xFFFF - pos) << 16) | static_cast<uint32_t>(row);
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

__device__ __forceinline__ int KeyRow(uint64_t key) { return static_cast<int>(key & 0xFFFFu); }
__device__ __forceinline__ int KeyPos(uint64_t key) { return 0xFFFF - static_cast<int>((key >> 16) & 0xFFFFu); }

__device__ __forceinline__ uint64_t ShflXor64(uint64_t v, int mask)
{
    uint32_t lo = static_cast<uint32_t>(v);
    uint32_t hi = static_cast<uint32_t>(v >> 32);
    lo = static_cast<uint32_t>(__shfl_xor(static_cast<int>(lo), mask, kWave));
    hi = static_cast<uint32_t>(__shfl_xor(static_cast<int>(hi), mask, kWave));
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

// Butterfly max over the 64 lanes of a wave; every lane receives the result.
__device__ __forceinline__ uint64_t WaveMaxKey(uint64_t k)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        uint64_t o = ShflXor64(k, m);
        k = o > k ? o : k;
    }
    return k;
}

// DPP lane exchange inside 16-lane rows (no LDS traffic).  Controls used here:
// quad_perm [1,0,3,2] = 0xB1 (lane ^ 1), quad_perm [2,3,0,1] = 0x4E (lane ^ 2),
// row_half_mirror = 0x141 (lane i <-> 7 - i within each 8 lanes), row_mirror =
// 0x140 (lane i <-> 15 - i within each 16 lanes).  Every source lane exists.
template <int CTRL>
__device__ __forceinline__ uint32_t Dpp(uint32_t v)
{
    return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), CTRL, 0xF, 0xF, false));
}

// Maximum (MAX = true) or minimum of v over aligned groups of S lanes
// (S = 8, 16, 32 or 64); every lane of a group receives its group's result.
template <int S, bool MAX>
__device__ __forceinline__ uint32_t SegReduce(uint32_t v)
{
    static_assert(S == 8 || S == 16 || S == 32 || S == 64, "segment of 8, 16, 32 or 64 lanes");
    auto op = [](uint32_t x, uint32_t y) { return MAX ? (x > y ? x : y) : (x < y ? x : y); };
    v = op(v, Dpp<0xB1>(v));
    v = op(v, Dpp<0x4E>(v));
    v = op(v, Dpp<0x141>(v));
    if constexpr (S >= 16) {
        v = op(v, Dpp<0x140>(v));
    }
    if constexpr (S >= 32) {
        // v_permlane16_swap: the odd 16-lane rows of the first copy trade places
        // with the even rows of the second, so every lane sees its partner row
        const auto p = __builtin_amdgcn_permlane16_swap(v, v, false, false);
        v = op(p[0], p[1]);
    }
    if constexpr (S == 64) {
        // v_permlane32_swap: the same between the two 32-lane halves
        const auto p = __builtin_amdgcn_permlane32_swap(v, v, false, false);
        v = op(p[0], p[1]);
    }
    return v;
}

// |v| as an order-preserving integer key; NaN maps to 0 so that it never wins.
__device__ __forceinline__ uint32_t MagKey(float v)
{
    const uint32_t m = __float_as_uint(v) & 0x7FFFFFFFu;
    return m > 0x7F800000u ? 0u : m;
}

// Compile-time loop: f(std::integral_constant<int, I>) for I in [Begin, End).
// Register arrays indexed by I stay in registers whatever the loop size.
template <int Begin, int End, typename F>
__device__ __forceinline__ void StaticFor(F&& f)
{
    if constexpr (Begin < End) {
        f(std::integral_constant<int, Begin>{});
        StaticFor<Begin + 1, End>(f);
    }
}

// Global memory through the global address space.  Through a generic pointer
// the compiler issues flat instructions, which count on lgkmcnt as well as
// vmcnt, so every later wait for an LDS access would also wait for all flat
// loads in flight; global_load / global_store count on vmcnt only.  The
// pointers must point to device (global) memory.
#define ACL_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ const ACL_GLOBAL T* G(const T* p)
{
    return (const ACL_GLOBAL T*)p;
}
template <typename T>
__device__ __forceinline__ ACL_GLOBAL T* G(T* p)
{
    return (ACL_GLOBAL T*)p;
}
// Whole-object loads and stores of 1, 2, 4, 8 or 16 bytes.  They move the bits
// as a plain integer or integer vector: HIP's vector types (uint4, float4, ...)
// are classes whose copy constructor takes a generic reference, which would
// turn the access back into a flat one.
template <int N>
struct GlobalBits;
template <>
struct GlobalBits<1> { using T = uint8_t; };
template <>
struct GlobalBits<2> { using T = uint16_t; };
template <>
struct GlobalBits<4> { using T = uint32_t; };
template <>
struct GlobalBits<8> { typedef uint32_t T __attribute__((ext_vector_type(2))); };
template <>
struct GlobalBits<16> { typedef uint32_t T __attribute__((ext_vector_type(4))); };

template <typename T>
__device__ __forceinline__ T LoadG(const void* p)
{
    using B = typename GlobalBits<sizeof(T)>::T;
    const B b = *(const ACL_GLOBAL B*)p;
    return __builtin_bit_cast(T, b);
}
template <typename T>
__device__ __forceinline__ void StoreG(void* p, const T& v)
{
    using B = typename GlobalBits<sizeof(T)>::T;
    *(ACL_GLOBAL B*)p = __builtin_bit_cast(B, v);
}

// Makes a wave-uniform value provably scalar (lives in an SGPR).
__device__ __forceinline__ int Uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ float ReadLane(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

__device__ __forceinline__ int LaneId() { return static_cast<int>(threadIdx.x) & (kWave - 1); }

// Reads the b-th entry of a device array of device pointers.
template <typename T>
__device__ __forceinline__ T* BatchPtr(T* const* arr, int b)
{
    return arr[b];
}

}  // namespace dev
}  // namespace aclblas
