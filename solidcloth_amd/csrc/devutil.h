// Small device helpers shared by the kernel translation units.
#pragma once

#include <hip/hip_runtime.h>

#include "scmath.h"

namespace sc {

__device__ __forceinline__ f3 xyz(float4 v) { return f3{v.x, v.y, v.z}; }
__device__ __forceinline__ float4 f4(f3 v, float w) { return make_float4(v.x, v.y, v.z, w); }

// Wave-wide maximum of non-negative, NaN-free values, then one atomic per wave.
__device__ __forceinline__ void wave_max_atomic(unsigned* dst, float v)
{
    for (int o = 32; o > 0; o >>= 1) {
        float w = __shfl_xor(v, o, 64);
        v = v < w ? w : v;
    }
    if ((threadIdx.x & 63) == 0 && v > 0.f) atomicMax(dst, __float_as_uint(v));
}

}  // namespace sc
