// Closest-point primitives for the GPU kernels and the host, in the
// reference's operation order (Ericson, Real-Time Collision Detection
// §5.1.5 and §5.1.9, as written in Geometry/Collision3D.cpp:10-141,298-301).
#pragma once

#include "scmath.h"

namespace sc {

enum { TR_VERTEX_A, TR_VERTEX_B, TR_VERTEX_C, TR_EDGE_AB, TR_EDGE_AC, TR_EDGE_BC, TR_FACE_INTERIOR };

SC_HD int closest_pt_point_triangle(f3 p, f3 a, f3 b, f3 c, f3& cp, f3& bar)
{
    bar = f3{0.f, 0.f, 0.f};
    const f3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
    const float d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.f && d2 <= 0.f) {
        bar.x = 1.f;
        cp = a;
        return TR_VERTEX_A;
    }
    const f3 bp = sub(p, b);
    const float d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.f && d4 <= d3) {
        bar.y = 1.f;
        cp = b;
        return TR_VERTEX_B;
    }
    const float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.f && d1 >= 0.f && d3 <= 0.f) {
        bar.y = d1 / (d1 - d3);
        bar.x = 1.f - bar.y;
        cp = add(a, mul(bar.y, ab));
        return TR_EDGE_AB;
    }
    const f3 cq = sub(p, c);
    const float d5 = dot(ab, cq), d6 = dot(ac, cq);
    if (d6 >= 0.f && d5 <= d6) {
        bar.z = 1.f;
        cp = c;
        return TR_VERTEX_C;
    }
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.f && d2 >= 0.f && d6 <= 0.f) {
        bar.z = d2 / (d2 - d6);
        bar.x = 1.f - bar.z;
        cp = add(a, mul(bar.z, ac));
        return TR_EDGE_AC;
    }
    const float va = d3 * d6 - d5 * d4;
    const float d43 = d4 - d3, d56 = d5 - d6;
    if (va <= 0.f && d43 >= 0.f && d56 >= 0.f) {
        bar.z = d43 / (d43 + d56);
        bar.y = 1.f - bar.z;
        cp = add(b, mul(bar.z, sub(c, b)));
        return TR_EDGE_BC;
    }
    const float denom = 1.f / (va + vb + vc);
    bar.y = vb * denom;
    bar.z = vc * denom;
    bar.x = 1.f - bar.y - bar.z;
    cp = add(add(a, mul(bar.y, ab)), mul(bar.z, ac));
    return TR_FACE_INTERIOR;
}

// Math::clamp (Math/Utils.h:24-28): std::max(lo, std::min(hi, v))
SC_HD float clampf(float v, float lo, float hi) { return smax(lo, smin(hi, v)); }

SC_HD int closest_pt_segm_segm(f3 p1, f3 q1, f3 p2, f3 q2, float& s, float& t, f3& c1, f3& c2)
{
    const f3 d1 = sub(q1, p1), d2 = sub(q2, p2), r = sub(p1, p2);
    const float a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
    const float c = dot(d1, r), b = dot(d1, d2);
    const float denom = a * e - b * b;
    s = (denom != 0.f) ? clampf((b * f - c * e) / denom, 0.f, 1.f) : 0.f;
    int region = 0;
    t = (b * s + f) / e;
    if (t < 0.f) {
        region |= 4;
        t = 0.f;
        s = clampf(-c / a, 0.f, 1.f);
    } else if (t > 1.f) {
        region |= 8;
        t = 1.f;
        s = clampf((b - c) / a, 0.f, 1.f);
    }
    if (s == 0.f) region |= 1;
    if (s == 1.f) region |= 2;
    c1 = add(p1, mul(s, d1));
    c2 = add(p2, mul(t, d2));
    return region;
}

// Collision3D.h:125-141: closest point to p on segment [a, b], parameter t
SC_HD f3 closest_pt_segm(f3 p, f3 a, f3 b, float& t)
{
    const f3 u = sub(b, a), v = sub(p, a);
    t = dot(u, v) / dot(u, u);
    if (t <= 0.f) {
        t = 0.f;
        return a;
    }
    if (t >= 1.f) {
        t = 1.f;
        return b;
    }
    return add(a, mul(t, u));
}

// Geometry/Collision3D.cpp:143-210: the nearest pair of points of segment
// [p, q] and triangle (a, b, c) among the segment against each edge and each
// end point against the triangle, first strictly nearer kept. Coordinates as
// the reference leaves them: an edge sets (1 - t, t) on its two corners, end
// point p its barycentric coordinates; the query of end point q writes q's
// coordinates whatever the outcome, and when q is nearest they are then
// replaced by p's. Returns the distance.
SC_HD float closest_pt_segm_triangle(f3 p, f3 q, f3 a, f3 b, f3 c, float& param, f3& bar, f3& closestEdge,
                                     f3& closestTri)
{
    float minDist = 3.402823466e+38f;  // FLT_MAX
    float s, t;
    f3 c1, c2;
    closest_pt_segm_segm(p, q, a, b, s, t, c1, c2);
    const float dist1 = len(sub(c1, c2));
    if (dist1 < minDist) {
        minDist = dist1;
        closestEdge = c1;
        closestTri = c2;
        param = s;
        bar = f3{1.f - t, t, 0.f};
    }
    closest_pt_segm_segm(p, q, b, c, s, t, c1, c2);
    const float dist2 = len(sub(c1, c2));
    if (dist2 < minDist) {
        minDist = dist2;
        closestEdge = c1;
        closestTri = c2;
        param = s;
        bar = f3{0.f, 1.f - t, t};
    }
    closest_pt_segm_segm(p, q, c, a, s, t, c1, c2);
    const float dist3 = len(sub(c1, c2));
    if (dist3 < minDist) {
        minDist = dist3;
        closestEdge = c1;
        closestTri = c2;
        param = s;
        bar = f3{t, 0.f, 1.f - t};
    }
    f3 coords, cp4;
    closest_pt_point_triangle(p, a, b, c, cp4, coords);
    const float dist4 = len(sub(p, cp4));
    if (dist4 < minDist) {
        minDist = dist4;
        closestEdge = p;
        closestTri = cp4;
        param = 0.f;
        bar = coords;
    }
    f3 cp5;
    closest_pt_point_triangle(q, a, b, c, cp5, bar);
    const float dist5 = len(sub(q, cp5));
    if (dist5 < minDist) {
        minDist = dist5;
        closestEdge = q;
        closestTri = cp5;
        param = 1.f;
        bar = coords;
    }
    return minDist;
}

// Collision3D.h:42-58: barycentric coordinates of p in (a, b, c)
SC_HD f3 barycentric(f3 a, f3 b, f3 c, f3 p)
{
    const f3 v0 = sub(b, a), v1 = sub(c, a), v2 = sub(p, a);
    const float d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1);
    const float d20 = dot(v2, v0), d21 = dot(v2, v1);
    const float denom = d00 * d11 - d01 * d01;
    f3 r;
    r.y = (d11 * d20 - d01 * d21) / denom;
    r.z = (d00 * d21 - d01 * d20) / denom;
    r.x = 1.f - r.y - r.z;
    return r;
}

// triple(a1 - a0, a2 - a0, a3 - a0) = (a3 - a0) . ((a1 - a0) x (a2 - a0))
SC_HD float signed_volume(f3 a0, f3 a1, f3 a2, f3 a3)
{
    return dot(sub(a3, a0), cross(sub(a1, a0), sub(a2, a0)));
}

}  // namespace sc
