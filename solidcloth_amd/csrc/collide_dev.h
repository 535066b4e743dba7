// Device queries against analytic and grid-SDF collidables.
#pragma once

#include <float.h>

#include "devutil.h"
#include "geom.h"
#include "kernels.h"

namespace sc {

// Geometry/SDF.h:105-155 — trilinear value and gradient of a node-sampled
// grid. Returns false when the cell lies outside (value FLT_MAX, gradient 0).
__device__ __forceinline__ bool sdf_lookup(const CollDev& c, f3 p, float& val, f3& g)
{
    float rx = (p.x - c.ax) / c.bx;
    float ry = (p.y - c.ay) / c.by;
    float rz = (p.z - c.az) / c.bz;
    float flx = floorf(rx), fly = floorf(ry), flz = floorf(rz);
    if (!(flx >= 0.f && flx <= (float)c.nc0 && fly >= 0.f && fly <= (float)c.nc1 && flz >= 0.f &&
          flz <= (float)c.nc2))
        return false;
    int x = (int)flx, y = (int)fly, z = (int)flz;
    float fx = fabsf(rx - (float)x), fy = fabsf(ry - (float)y), fz = fabsf(rz - (float)z);
    float4 nd[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        // term order of the reference: (x,y,z) (x,y,z+1) (x,y+1,z) ... (x+1,y+1,z+1)
        int dx = (k >> 2) & 1, dy = (k >> 1) & 1, dz = k & 1;
        int X = x + dx, Y = y + dy, Z = z + dz;
        bool out = X > c.nc0 || Y > c.nc1 || Z > c.nc2;
        nd[k] = out ? make_float4(0.f, 0.f, 0.f, 0.f) : c.nodes[X + Y * c.nv0 + Z * c.nv0 * c.nv1];
    }
    const float ox = 1.f - fx, oy = 1.f - fy, oz = 1.f - fz;
    float w[8] = {ox * oy * oz, ox * oy * fz, ox * fy * oz, ox * fy * fz,
                  fx * oy * oz, fx * oy * fz, fx * fy * oz, fx * fy * fz};
    float v = w[0] * nd[0].w;
    f3 gr = mul(w[0], xyz(nd[0]));
#pragma unroll
    for (int k = 1; k < 8; k++) {
        v = v + w[k] * nd[k].w;
        gr = add(gr, mul(w[k], xyz(nd[k])));
    }
    val = v;
    g = gr;
    return true;
}

// Physics/ClothCollisionSDF.h:26-41 (SphereSDF::QuerySegment): distance of
// segment [p, q] to the sphere surface, surface point, radial normal and the
// segment coordinates (1 - t, t) of the closest segment point.
__device__ __forceinline__ float sphere_segment(const CollDev& c, f3 p, f3 q, f3& cp, f3& n, float& w1, float& w2)
{
    const f3 ctr{c.ax, c.ay, c.az};
    float t;
    f3 d = sub(closest_pt_segm(ctr, p, q, t), ctr);
    const float l = len(d);
    d = mul(1.f / l, d);
    n = d;
    cp = add(ctr, mul(c.r, d));
    w1 = 1.0f - t;
    w2 = t;
    return l - c.r;
}

// Physics/ClothCollisionSDF.h:43-55 (SphereSDF::QueryTriangle)
__device__ __forceinline__ float sphere_triangle(const CollDev& c, f3 a, f3 b, f3 cc, f3& cp, f3& n, f3& bar)
{
    const f3 ctr{c.ax, c.ay, c.az};
    f3 pt;
    closest_pt_point_triangle(ctr, a, b, cc, pt, bar);
    f3 d = sub(pt, ctr);
    const float l = len(d);
    d = mul(1.f / l, d);
    n = d;
    cp = add(ctr, mul(c.r, d));
    return l - c.r;
}

// Physics/ClothCollisionSDF.h:57-122 (CapsuleSDF): a capsule is its axis
// [P, Q] = (ax, ay, az)-(bx, by, bz) grown by radius r.
// QueryPoint: distance of x to the surface, surface point and normal.
__device__ __forceinline__ float capsule_point(const CollDev& c, f3 x, f3& cp, f3& n)
{
    const f3 P{c.ax, c.ay, c.az}, Q{c.bx, c.by, c.bz};
    float t;
    const f3 s = closest_pt_segm(x, P, Q, t);
    f3 d = sub(x, s);
    const float l = len(d);
    d = mul(1.f / l, d);
    n = d;
    cp = add(s, mul(c.r, d));
    return l - c.r;
}

// QuerySegment: segment [p, q] against the axis; coordinates (1 - s, s) on
// the segment.
__device__ __forceinline__ float capsule_segment(const CollDev& c, f3 p, f3 q, f3& cp, f3& n, float& w1, float& w2)
{
    const f3 P{c.ax, c.ay, c.az}, Q{c.bx, c.by, c.bz};
    float s, t;
    f3 c1, c2;
    closest_pt_segm_segm(p, q, P, Q, s, t, c1, c2);
    f3 d = sub(c1, c2);
    const float l = len(d);
    d = mul(1.f / l, d);
    n = d;
    cp = add(c2, mul(c.r, d));
    w1 = 1.0f - s;
    w2 = s;
    return l - c.r;
}

// QueryTriangle: the axis against triangle (a, b, cc), coordinates as
// closest_pt_segm_triangle leaves them.
__device__ __forceinline__ float capsule_triangle(const CollDev& c, f3 a, f3 b, f3 cc, f3& cp, f3& n, f3& bar)
{
    const f3 P{c.ax, c.ay, c.az}, Q{c.bx, c.by, c.bz};
    float t;
    f3 ce, ct;
    const float l = closest_pt_segm_triangle(P, Q, a, b, cc, t, bar, ce, ct);
    f3 d = sub(ct, ce);
    d = mul(1.f / l, d);
    n = d;
    cp = add(ce, mul(c.r, d));
    return l - c.r;
}

// Physics/ClothCollisionSDF.h:148-205 (GridSDF::QueryTriangle): Frank-Wolfe
// minimisation of the SDF over the triangle from 0.33 (a + b + c), step
// 2 / (i + 2), at most 99 steps, stopped at a relative change below 1e-4 or
// when the iterate leaves the grid (value FLT_MAX).
__device__ __forceinline__ float sdf_triangle(const CollDev& c, f3 a, f3 b, f3 cc, f3& closest, f3& n, f3& bar)
{
    f3 cp = mul(0.33f, add(add(a, b), cc));
    float distOld = 0.f;
    for (int i = 1; i < 100; i++) {
        float dist;
        f3 g;
        if (!sdf_lookup(c, cp, dist, g)) return FLT_MAX;
        const float dotA = dot(a, g), dotB = dot(b, g), dotC = dot(cc, g);
        float d = dotA;
        f3 s = a;
        if (dotB < d) {
            d = dotB;
            s = b;
        }
        if (dotC < d) {
            d = dotC;
            s = cc;
        }
        if (distOld > 0.f) {
            const float rel = fabsf(distOld - dist) / distOld;
            if (rel < 0.0001f) break;
        }
        distOld = dist;
        const float alpha = 2.f / ((float)i + 2.f);
        cp = add(cp, mul(alpha, sub(s, cp)));
    }
    float val;
    f3 g;
    if (!sdf_lookup(c, cp, val, g)) {
        val = FLT_MAX;
        g = f3{0.f, 0.f, 0.f};
    }
    g = normalized(g);
    n = g;
    closest = sub(cp, mul(val, g));
    bar = barycentric(a, b, cc, cp);
    return val;
}

}  // namespace sc
