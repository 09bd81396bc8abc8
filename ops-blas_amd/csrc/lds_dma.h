// Direct global-to-LDS loads and transposing LDS reads (gfx950).
//
// The loads are issued with inline assembly, so the compiler neither tracks
// them nor inserts waits for them: the caller waits with WaitVmcnt (they are
// the only vector-memory operations in flight where it does) and publishes the
// LDS data to the workgroup with BarrierLds.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace aclblas {
namespace dev {

typedef short V4i16 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) V4i16 LdsV4i16;

// LDS byte address of a __shared__ object (the low 32 bits of its generic
// address are its LDS offset).
__device__ __forceinline__ uint32_t LdsAddr(const void* p)
{
    return static_cast<uint32_t>(reinterpret_cast<uintptr_t>(p));
}

// One wave instruction: lane l copies the 16 bytes at g (16-byte aligned,
// per lane) to LDS bytes [lds + 16 l, lds + 16 l + 16); lds is wave-uniform.
// The LDS base goes through M0, which needs one wait state between its write
// and the load that reads it.
// M0 is listed as clobbered; the compiler warns that it reserves M0.  In the
// kernels that use this load no other instruction reads M0 (no dynamic
// register indexing, no compiler-issued LDS DMA): in their gfx950 code every
// M0 reference is this function's own s_mov_b32.  The compiler's own form,
// __builtin_amdgcn_global_load_lds(g, lds, 16, 0, 0), is not used because the
// compiler then waits for every outstanding LDS DMA after each barrier of the
// main loop (an s_waitcnt vmcnt(0) after the s_barrier), which drains the
// ring of stages in flight on every stage.
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
__device__ __forceinline__ void LdsDma16(const void* g, uint32_t lds)
{
    asm volatile("s_mov_b32 m0, %1\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %0, off" : : "v"(g), "s"(lds)
                 : "memory", "m0");
}
#pragma clang diagnostic pop

// Waits until at most N vector-memory operations of this wave are in flight
// (they complete in issue order).
template <int N>
__device__ __forceinline__ void WaitVmcnt()
{
    static_assert(N >= 0 && N < 64, "vmcnt range");
    asm volatile("s_waitcnt vmcnt(%0)" : : "n"(N) : "memory");
}

// Issue priority of this wave relative to the other waves of its SIMD (0..3).
template <int P>
__device__ __forceinline__ void SetPriority()
{
    static_assert(P >= 0 && P <= 3, "priority range");
    asm volatile("s_setprio %0" : : "n"(P));
}

// Workgroup barrier after this wave's LDS writes have completed.
__device__ __forceinline__ void BarrierLds()
{
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" : : : "memory");
}

// ds_read_b64_tr_b16.  Within each group of 16 lanes, let M[j][e] (e < 4) be
// the four 16-bit values at lane j's address.  Lane i receives
// M[4 e + i / 4][i % 4] for e = 0..3: with lane j = 4 e + a pointing at
// values (row 4a .. 4a+3, column e) of a row-contiguous block, lane i gets
// row i, columns 0..3 - a 16 x 4 transpose per lane group.
__device__ __forceinline__ uint2 ReadTr16(const void* lds)
{
    const V4i16 v = __builtin_amdgcn_ds_read_tr16_b64_v4i16((LdsV4i16*)(lds));
    return __builtin_bit_cast(uint2, v);
}

typedef int V2i32 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) V2i32 LdsV2i32;

// ds_read_b64_tr_b8.  Within each group of 16 lanes, let M[j][e] (e < 8) be
// the eight bytes at lane j's address.  Lane i receives M[2 e + i / 8][i % 8]
// for e = 0..7: with lane j = 2 e + a pointing at bytes (row 8a .. 8a+7,
// column e) of a row-contiguous block, lane i gets row i, columns 0..7.
__device__ __forceinline__ uint2 ReadTr8(const void* lds)
{
    const V2i32 v = __builtin_amdgcn_ds_read_tr8_b64_v2i32((LdsV2i32*)(lds));
    return __builtin_bit_cast(uint2, v);
}

}  // namespace dev
}  // namespace aclblas
