// Small device helpers shared by the kernel translation units.
#pragma once

#include <hip/hip_runtime.h>

#include "scmath.h"

namespace sc {

__device__ __forceinline__ f3 xyz(float4 v) { return f3{v.x, v.y, v.z}; }
__device__ __forceinline__ float4 f4(f3 v, float w) { return make_float4(v.x, v.y, v.z, w); }

// Workgroup barrier that orders LDS only: workgroups that exchange data
// through LDS alone (the colours of a slab, the groups of a staged contact
// pass) may keep global loads and stores in flight across it (a plain
// __syncthreads waits for every outstanding global access).
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Wave-wide maximum (every lane of the wave active), returned on every lane;
// the reduction moves lanes with DPP (shifts inside rows of 16, then row
// broadcasts), so lane 63 ends with the maximum.
__device__ __forceinline__ unsigned wave_max(unsigned u)
{
    u = max(u, (unsigned)__builtin_amdgcn_update_dpp(0, (int)u, 0x111, 0xf, 0xf, false));  // row_shr:1
    u = max(u, (unsigned)__builtin_amdgcn_update_dpp(0, (int)u, 0x112, 0xf, 0xf, false));  // row_shr:2
    u = max(u, (unsigned)__builtin_amdgcn_update_dpp(0, (int)u, 0x114, 0xf, 0xf, false));  // row_shr:4
    u = max(u, (unsigned)__builtin_amdgcn_update_dpp(0, (int)u, 0x118, 0xf, 0xf, false));  // row_shr:8
    u = max(u, (unsigned)__builtin_amdgcn_update_dpp(0, (int)u, 0x142, 0xa, 0xf, false));  // row_bcast:15
    u = max(u, (unsigned)__builtin_amdgcn_update_dpp(0, (int)u, 0x143, 0xc, 0xf, false));  // row_bcast:31
    return (unsigned)__builtin_amdgcn_readlane((int)u, 63);
}

// The wave's maximum of non-negative, NaN-free floats (which order as their
// unsigned bit patterns), one atomic per wave.
__device__ __forceinline__ void wave_max_atomic(unsigned* dst, float v)
{
    const unsigned m = wave_max(__float_as_uint(v));
    if ((threadIdx.x & 63) == 0 && m > 0u) atomicMax(dst, m);
}

}  // namespace sc
