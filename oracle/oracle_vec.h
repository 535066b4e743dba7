// TEST INFRASTRUCTURE ONLY. This directory is the CPU oracle that the GPU path
// is checked against; only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it. The product library never links it.
//
// Vector arithmetic with the exact operation order of the reference's value
// types, so that an FP32 instantiation rounds like the reference:
//   Math/Vector3.h:102-105  Dot = (x*x' + y*y') + z*z'
//   Math/Vector3.h:131-134  Length = sqrt(Dot) (correctly rounded sqrtf)
//   Math/Vector3.h:136-141  Normalize = Scale(1/len)
//   Math/Vector3.h:150-153  Cross component order
//   Math/Vector3.h:237-247  scalar * vector = (s*x, s*y, s*z)
#pragma once
#include <cmath>

namespace oracle {

template <class R>
struct V3 {
    R x, y, z;
};

template <class R> inline V3<R> mk(R x, R y, R z) { return V3<R>{x, y, z}; }
template <class R> inline V3<R> operator+(const V3<R>& a, const V3<R>& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline V3<R> operator-(const V3<R>& a, const V3<R>& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> operator-(const V3<R>& a) { return {-a.x, -a.y, -a.z}; }
template <class R> inline V3<R> operator*(R s, const V3<R>& a) { return {s * a.x, s * a.y, s * a.z}; }
template <class R> inline V3<R> operator*(const V3<R>& a, R s) { return {s * a.x, s * a.y, s * a.z}; }
template <class R> inline V3<R>& operator+=(V3<R>& a, const V3<R>& b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
template <class R> inline V3<R>& operator-=(V3<R>& a, const V3<R>& b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; return a; }
template <class R> inline void scale(V3<R>& a, R s) { a.x *= s; a.y *= s; a.z *= s; }
template <class R> inline R dot(const V3<R>& a, const V3<R>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> inline V3<R> cross(const V3<R>& a, const V3<R>& b)
{
    // Math/Vector3.h:152 — (y*vz - vy*z, -x*vz + z*vx, x*vy - vx*y)
    return {a.y * b.z - b.y * a.z, -a.x * b.z + a.z * b.x, a.x * b.y - b.x * a.y};
}
inline float rsqrt_len(float v) { return sqrtf(v); }
inline double rsqrt_len(double v) { return std::sqrt(v); }
template <class R> inline R length(const V3<R>& a) { return rsqrt_len(dot(a, a)); }
template <class R> inline void normalize(V3<R>& a)
{
    R len = length(a);
    scale(a, R(1) / len);
}
// std::max / std::min semantics (first argument returned on ties and NaN in
// the comparison), Math/Utils.h:9-10
template <class R> inline R smax(R a, R b) { return (a < b) ? b : a; }
template <class R> inline R smin(R a, R b) { return (b < a) ? b : a; }

// Project-written arccosine. Abramowitz & Stegun, Handbook of Mathematical
// Functions, formula 4.4.46: acos(x) = sqrt(1-x) * sum_{k=0..7} a_k x^k on
// [0,1] (|error| <= 2e-8), reflected with acos(x) = pi - acos(-x) for x < 0.
// Evaluated in Horner form without contraction. The GPU kernels evaluate the
// same sequence, so both sides agree bit for bit; the accuracy against libm
// is bounded in tests/test_oracle_math.py.
inline float acos_as(float x)
{
    const float a0 = 1.5707963050f, a1 = -0.2145988016f, a2 = 0.0889789874f, a3 = -0.0501743046f;
    const float a4 = 0.0308918810f, a5 = -0.0170881256f, a6 = 0.0066700901f, a7 = -0.0012624911f;
    const float pi = 3.14159265358979f;
    bool neg = x < 0.f;
    float ax = neg ? -x : x;
    float p = a7;
    p = p * ax + a6;
    p = p * ax + a5;
    p = p * ax + a4;
    p = p * ax + a3;
    p = p * ax + a2;
    p = p * ax + a1;
    p = p * ax + a0;
    float r = sqrtf(1.f - ax) * p;
    return neg ? pi - r : r;
}
inline double acos_as(double x) { return std::acos(x); }

}  // namespace oracle
