// Nearest-feature queries against a static triangle-mesh obstacle, on the GPU.
//
// The reference (Geometry/ClosestPointOnMesh.cpp:304-360, 454-480, 589-667)
// walks an AABB tree and keeps the first strictly nearer feature in its walk
// order. Here the answer is defined independently of any tree: the feature
// with the smallest FP32 distance, ties to the lowest feature index, among
// features at distance below the query radius (distance 0 skipped, as in the
// reference). The LBVH walk only prunes what cannot qualify: a node is
// skipped when its box lies farther than the current best plus a slack that
// covers the rounding of the box distance and of the closest-point
// computation (every closest point lies inside its feature's box up to
// rounding, since all clamped parameters stay in [0, 1]).
#pragma once

#include "devutil.h"
#include "geom.h"
#include "kernels.h"
#include "lbvh.h"

namespace sc {

// Squared distance between boxes [alo, ahi] and [blo, bhi] (0 if they overlap).
__device__ __forceinline__ float box_dist2(f3 alo, f3 ahi, float4 blo, float4 bhi)
{
    const float dx = fmaxf(fmaxf(blo.x - ahi.x, alo.x - bhi.x), 0.f);
    const float dy = fmaxf(fmaxf(blo.y - ahi.y, alo.y - bhi.y), 0.f);
    const float dz = fmaxf(fmaxf(blo.z - ahi.z, alo.z - bhi.z), 0.f);
    return dx * dx + dy * dy + dz * dz;
}

// Depth-first walk, nearer child first. `visit(k)` handles sorted leaf k and
// may shrink `r2`, the squared radius within which boxes are still opened.
// The Karras tree has depth at most 63 (prefix lengths strictly grow along a
// path and are below 64), so the stack of 64 never overflows.
template <class Visit>
__device__ __forceinline__ void bvh_walk(const BvhView& bv, f3 qlo, f3 qhi, float& r2, Visit visit)
{
    if (bv.n <= 0) return;
    if (bv.n == 1) {
        if (box_dist2(qlo, qhi, bv.leafLo[0], bv.leafHi[0]) <= r2) visit(0);
        return;
    }
    int stack[64];
    float sdist[64];
    int sp = 0;
    int node = 0;
    if (box_dist2(qlo, qhi, bv.lo[0], bv.hi[0]) > r2) return;
    while (true) {
        const float4 L = bv.lo[node], H = bv.hi[node];
        const int c[2] = {__float_as_int(L.w), __float_as_int(H.w)};
        float d[2];
        int next = -1;
        float dnext = 0.f;
#pragma unroll
        for (int k = 0; k < 2; k++) {
            if (c[k] < 0) {
                const int leaf = ~c[k];
                if (box_dist2(qlo, qhi, bv.leafLo[leaf], bv.leafHi[leaf]) <= r2) visit(leaf);
                d[k] = -1.f;
            } else {
                d[k] = box_dist2(qlo, qhi, bv.lo[c[k]], bv.hi[c[k]]);
            }
        }
        const bool in0 = d[0] >= 0.f && d[0] <= r2, in1 = d[1] >= 0.f && d[1] <= r2;
        if (in0 && in1) {
            const int nearK = d[1] < d[0] ? 1 : 0;
            next = c[nearK];
            dnext = d[nearK];
            if (sp < 64) {
                stack[sp] = c[1 - nearK];
                sdist[sp] = d[1 - nearK];
                sp++;
            }
        } else if (in0) {
            next = c[0];
            dnext = d[0];
        } else if (in1) {
            next = c[1];
            dnext = d[1];
        }
        (void)dnext;
        if (next < 0) {
            while (sp > 0) {
                sp--;
                if (sdist[sp] <= r2) {
                    next = stack[sp];
                    break;
                }
            }
            if (next < 0) return;
        }
        node = next;
    }
}

// Rounding allowance for pruning, proportional to the coordinates involved.
__device__ __forceinline__ float query_slack(f3 qlo, f3 qhi, float meshScale)
{
    const float m = fmaxf(fmaxf(fmaxf(fabsf(qlo.x), fabsf(qhi.x)), fmaxf(fabsf(qlo.y), fabsf(qhi.y))),
                          fmaxf(fabsf(qlo.z), fabsf(qhi.z)));
    return 1e-5f * (m + meshScale);
}

struct NearestResult {
    float dist;
    int feature;
    f3 point;   // closest point on the obstacle feature
    f3 normal;
    float param;  // edge query: parameter on the cloth edge; triangle query: unused
    f3 bary;      // triangle query: barycentric coordinates on the cloth triangle
};

// Physics/ClothCollision.cpp:108-135 with ClosestPointOnMeshToPointAcc
// (ClosestPointOnMesh.cpp:304-360): nearest obstacle triangle to point q;
// normal = (q - closest) / dist.
__device__ __forceinline__ bool nearest_triangle(const MeshGeo& g, f3 q, float radius, NearestResult& r)
{
    r.dist = radius;
    r.feature = -1;
    const float slack = query_slack(q, q, g.scale);
    float lim = radius + slack;
    float r2 = lim * lim;
    bvh_walk(g.triBvh, q, q, r2, [&](int leaf) {
        const int t = __float_as_int(g.triBvh.leafLo[leaf].w);
        const int4 ix = g.tri[t];
        f3 pt, bar;
        closest_pt_point_triangle(q, xyz(g.v[ix.x]), xyz(g.v[ix.y]), xyz(g.v[ix.z]), pt, bar);
        const f3 delta = sub(q, pt);
        const float d = len(delta);
        if (d == 0.f) return;
        if (d < r.dist || (d == r.dist && t < r.feature)) {
            r.dist = d;
            r.feature = t;
            r.point = pt;
            r.normal = mul(1.f / d, delta);
            lim = d + slack;
            r2 = lim * lim;
        }
    });
    return r.feature >= 0 && r.dist < radius;
}

// Physics/ClothCollision.cpp:137-166 with ClosestPointOnMeshToTriangleAcc
// (ClosestPointOnMesh.cpp:454-480): nearest obstacle vertex to the cloth
// triangle (a, b, c); normal = (closest on triangle - vertex) / dist.
__device__ __forceinline__ bool nearest_vertex(const MeshGeo& g, f3 a, f3 b, f3 c, float radius, NearestResult& r)
{
    r.dist = radius;
    r.feature = -1;
    const f3 lo{fminf(fminf(a.x, b.x), c.x), fminf(fminf(a.y, b.y), c.y), fminf(fminf(a.z, b.z), c.z)};
    const f3 hi{fmaxf(fmaxf(a.x, b.x), c.x), fmaxf(fmaxf(a.y, b.y), c.y), fmaxf(fmaxf(a.z, b.z), c.z)};
    const float slack = query_slack(lo, hi, g.scale);
    float lim = radius + slack;
    float r2 = lim * lim;
    bvh_walk(g.vtxBvh, lo, hi, r2, [&](int leaf) {
        const int v = __float_as_int(g.vtxBvh.leafLo[leaf].w);
        const f3 mv = xyz(g.v[v]);
        f3 cp, bar;
        closest_pt_point_triangle(mv, a, b, c, cp, bar);
        const f3 delta = sub(cp, mv);
        const float d = len(delta);
        if (d == 0.f) return;
        if (d < r.dist || (d == r.dist && v < r.feature)) {
            r.dist = d;
            r.feature = v;
            r.point = mv;
            r.bary = bar;
            r.normal = mul(1.f / d, delta);
            lim = d + slack;
            r2 = lim * lim;
        }
    });
    return r.feature >= 0 && r.dist < radius;
}

// Physics/ClothCollision.cpp:168-197 with ClosestPointOnMeshToSegmentEdgeRec
// (ClosestPointOnMesh.cpp:589-628): nearest obstacle edge to the cloth edge
// (p, q); normal = the edge's pseudo-normal, or a vertex normal when the
// closest obstacle point is an end point (ComputeEdgeNormal, 483-497).
__device__ __forceinline__ bool nearest_edge(const MeshGeo& g, f3 p, f3 q, float radius, NearestResult& r)
{
    r.dist = radius;
    r.feature = -1;
    const f3 lo{fminf(p.x, q.x), fminf(p.y, q.y), fminf(p.z, q.z)};
    const f3 hi{fmaxf(p.x, q.x), fmaxf(p.y, q.y), fmaxf(p.z, q.z)};
    const float slack = query_slack(lo, hi, g.scale);
    float lim = radius + slack;
    float r2 = lim * lim;
    bvh_walk(g.edgeBvh, lo, hi, r2, [&](int leaf) {
        const int e = __float_as_int(g.edgeBvh.leafLo[leaf].w);
        const int2 ev = g.edge[e];
        float ps, pm;
        f3 cs, cm;
        closest_pt_segm_segm(p, q, xyz(g.v[ev.x]), xyz(g.v[ev.y]), ps, pm, cs, cm);
        const f3 delta = sub(cm, cs);
        const float d = len(delta);
        if (d == 0.f) return;
        if (d < r.dist || (d == r.dist && e < r.feature)) {
            r.dist = d;
            r.feature = e;
            r.point = cm;
            r.param = ps;
            float4 nf = g.en[e];
            if (pm == 0.f) nf = g.vn[ev.x];
            if (pm == 1.f) nf = g.vn[ev.y];
            r.normal = xyz(nf);
            lim = d + slack;
            r2 = lim * lim;
        }
    });
    return r.feature >= 0 && r.dist < radius;
}

}  // namespace sc
