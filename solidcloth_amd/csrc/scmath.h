// Vector arithmetic shared by the host set-up code and the gfx950 kernels.
//
// Every operation is written out component by component in the operation
// order of the reference's value types, so that FP32 results are identical on
// the CPU and the GPU when both are compiled without contraction
// (-ffp-contract=off) and with correctly rounded division and square root:
//   Math/Vector3.h:102-105  dot   = (x*x' + y*y') + z*z'
//   Math/Vector3.h:131-141  length = sqrt(dot), normalize = scale by 1/len
//   Math/Vector3.h:150-153  cross component order
//   Math/Vector3.h:237-247  scalar * vector = (s*x, s*y, s*z)
#pragma once

#if defined(__HIPCC__)
#define SC_HD __host__ __device__ __forceinline__
#else
#define SC_HD inline
#endif

#include <math.h>

namespace sc {

struct f3 {
    float x, y, z;
};

SC_HD f3 mkf3(float x, float y, float z) { return f3{x, y, z}; }
SC_HD f3 add(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
SC_HD f3 sub(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
SC_HD f3 neg(f3 a) { return f3{-a.x, -a.y, -a.z}; }
SC_HD f3 mul(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
SC_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
SC_HD f3 cross(f3 a, f3 b) { return f3{a.y * b.z - b.y * a.z, -a.x * b.z + a.z * b.x, a.x * b.y - b.x * a.y}; }
SC_HD float len(f3 a) { return sqrtf(dot(a, a)); }
SC_HD f3 normalized(f3 a)
{
    float l = len(a);
    return mul(1.f / l, a);
}
// std::max / std::min semantics (Math/Utils.h:9-10): the first argument wins
// ties and NaN comparisons.
SC_HD float smax(float a, float b) { return (a < b) ? b : a; }
SC_HD float smin(float a, float b) { return (b < a) ? b : a; }

// Arccosine from Abramowitz & Stegun 4.4.46 (|error| <= 2e-8 on [0,1] in
// exact arithmetic), reflected as pi - acos(-x) for x < 0, in Horner form.
// Host and device evaluate the identical sequence, so the bend angle agrees
// bit for bit everywhere it is computed.
SC_HD float acos_as(float x)
{
    const float a0 = 1.5707963050f, a1 = -0.2145988016f, a2 = 0.0889789874f, a3 = -0.0501743046f;
    const float a4 = 0.0308918810f, a5 = -0.0170881256f, a6 = 0.0066700901f, a7 = -0.0012624911f;
    const float pi = 3.14159265358979f;
    bool negx = x < 0.f;
    float ax = negx ? -x : x;
    float p = a7;
    p = p * ax + a6;
    p = p * ax + a5;
    p = p * ax + a4;
    p = p * ax + a3;
    p = p * ax + a2;
    p = p * ax + a1;
    p = p * ax + a0;
    float r = sqrtf(1.f - ax) * p;
    return negx ? pi - r : r;
}

}  // namespace sc
