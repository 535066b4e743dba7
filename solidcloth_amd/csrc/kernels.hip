// gfx950 kernels for one position-based-dynamics cloth substep.
//
// Each kernel restates one loop of the reference solver; the per-element
// arithmetic follows the reference's operation order exactly (no
// contraction, correctly rounded division and square root), so a kernel that
// processes a set of vertex-disjoint constraints in parallel produces the
// same bits as the reference processing them one after another.
#include "kernels.h"

#include "bvh_query.h"
#include "collide_dev.h"
#include "ctl_dev.h"
#include "devutil.h"
#include "hiputil.h"

namespace sc {

static constexpr int kBlock = 256;

// Physics/ClothPBD.cpp:258-266 — symplectic Euler prediction. Thread 0 also
// resets the solver control block for the substep.
__global__ void k_predict(float4* __restrict__ pos, float4* __restrict__ prev, float4* __restrict__ vel, int n, float h,
                          f3 gdv, float drag, SolverCtl* ctl)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        begin_substep_ctl(ctl);
    }
    if (i < kErrParts) begin_substep_parts(ctl, i);
    if (i >= n) return;
    float4 P = pos[i];
    prev[i] = make_float4(P.x, P.y, P.z, 0.f);
    if (P.w == 0.f) return;
    float4 V = vel[i];
    f3 v = add(xyz(V), gdv);
    v = f3{v.x * drag, v.y * drag, v.z * drag};
    f3 p = add(xyz(P), mul(h, v));
    vel[i] = f4(v, 0.f);
    pos[i] = f4(p, P.w);
}

__device__ __forceinline__ void put_contact(ContactBuf cb, int v, int& cnt, f3 p, f3 n, int ci, int feat = -1)
{
    cb.P[cnt * cb.nv + v] = f4(p, 0.f);
    cb.N[cnt * cb.nv + v] = f4(n, 0.f);
    cb.src[cnt * cb.nv + v] = ci;
    cb.feat[cnt * cb.nv + v] = feat;
    cnt++;
}

// Physics/ClothCollision.cpp:50-135, 199-220, 430-470 — external vertex
// contacts of vertex v (predicted position p, previous position pv),
// collidables visited in world order. `meshRad` is the mesh query radius
// thickness + max(tolerance, 1.5 max_i |pos_i - prev_i|) (ClothCollision.cpp:95-104).
// Instantiated without the mesh branch for worlds that have no mesh.
template <bool kMesh>
__device__ __forceinline__ void detect_vertex(int v, const f3 p, const f3 pv, const CollSet& cs, ContactBuf cb,
                                              float thickness, float tolerance, int collFlags, float meshRad)
{
    int cnt = 0;
    const float tv = len(sub(pv, p));
    const float radTol = thickness + smax(tolerance, tv * 1.1f);
    for (int ci = 0; ci < cs.n; ci++) {
        const CollDev& c = cs.c[ci];
        if (c.type == CT_WALLS) {
            if (!(collFlags & CF_WALLS)) continue;
            const float r = thickness + tolerance;
            if (p.y - r <= c.ay) put_contact(cb, v, cnt, f3{p.x, c.ay, p.z}, f3{0.f, 1.f, 0.f}, ci);
            if (p.x - r <= c.ax) put_contact(cb, v, cnt, f3{c.ax, p.y, p.z}, f3{1.f, 0.f, 0.f}, ci);
            if (p.x + r >= c.bx) put_contact(cb, v, cnt, f3{c.bx, p.y, p.z}, f3{-1.f, 0.f, 0.f}, ci);
            if (p.z - r <= c.az) put_contact(cb, v, cnt, f3{p.x, p.y, c.az}, f3{0.f, 0.f, 1.f}, ci);
            if (p.z + r >= c.bz) put_contact(cb, v, cnt, f3{p.x, p.y, c.bz}, f3{0.f, 0.f, -1.f}, ci);
        } else if (c.type == CT_SPHERE) {
            if (!(collFlags & CF_VERTICES)) continue;
            // Physics/ClothCollisionSDF.h:14-24
            f3 ctr{c.ax, c.ay, c.az};
            f3 d = sub(pv, ctr);
            float l = len(d);
            d = mul(1.f / l, d);
            f3 cp = add(ctr, mul(c.r, d));
            float dist = l - c.r;
            if (dist < radTol && dist >= 0.f) put_contact(cb, v, cnt, cp, d, ci);
        } else if (c.type == CT_CAPSULE) {
            if (!(collFlags & CF_VERTICES)) continue;
            f3 cp, n;
            const float dist = capsule_point(c, pv, cp, n);
            if (dist < radTol && dist >= 0.f) put_contact(cb, v, cnt, cp, n, ci);
        } else if (c.type == CT_SDF) {
            if (!(collFlags & CF_VERTICES)) continue;
            // Physics/ClothCollisionSDF.h:130-139
            float val;
            f3 g;
            if (!sdf_lookup(c, pv, val, g)) continue;  // value FLT_MAX: never a contact
            g = normalized(g);
            f3 cp = sub(pv, mul(val, g));
            if (val < radTol && val >= 0.f) put_contact(cb, v, cnt, cp, g, ci);
        } else if (kMesh && c.type == CT_MESH) {
            if (!(collFlags & CF_VERTICES)) continue;
            NearestResult r;
            if (nearest_triangle(*c.mesh, pv, meshRad, r)) put_contact(cb, v, cnt, r.point, r.normal, ci, r.feature);
        }
    }
    cb.count[v] = cnt;
}

__global__ void k_detect(const float4* __restrict__ pos, const float4* __restrict__ prev, int n, CollSet cs,
                         ContactBuf cb, float thickness, float tolerance, int collFlags,
                         const SolverCtl* __restrict__ ctl)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const float meshRad = thickness + smax(tolerance, __uint_as_float(ctl->maxTv) * 1.5f);
    detect_vertex<true>(v, xyz(pos[v]), xyz(prev[v]), cs, cb, thickness, tolerance, collFlags, meshRad);
}

// Prediction (ClothPBD.cpp:258-266) and detection of the same vertex in one
// pass; detection of a vertex reads only that vertex's positions, so this
// equals the reference's two loops.
__global__ void k_predict_detect(float4* __restrict__ pos, float4* __restrict__ prev, float4* __restrict__ vel, int n,
                                 float h, f3 gdv, float drag, SolverCtl* ctl, CollSet cs, ContactBuf cb,
                                 float thickness, float tolerance, int collFlags)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        begin_substep_ctl(ctl);
    }
    if (i < kErrParts) begin_substep_parts(ctl, i);
    if (i >= n) return;
    float4 P = pos[i];
    const f3 pv = xyz(P);
    prev[i] = make_float4(P.x, P.y, P.z, 0.f);
    f3 p = pv;
    if (P.w != 0.f) {
        f3 v = add(xyz(vel[i]), gdv);
        v = f3{v.x * drag, v.y * drag, v.z * drag};
        p = add(p, mul(h, v));
        vel[i] = f4(v, 0.f);
        pos[i] = f4(p, P.w);
    }
    detect_vertex<false>(i, p, pv, cs, cb, thickness, tolerance, collFlags, 0.f);
}

// Physics/ClothPBD.cpp:52-108 — one FEM triangle projection (warp, weft,
// shear). P1..P3 hold position and inverse mass; only positions change.
__device__ __forceinline__ void fem_project(float4& P1, float4& P2, float4& P3, const float4 A, const float4 B,
                                            const float2 C, float h2, float epsUV)
{
    const float du1 = A.x, du2 = A.y, dv1 = A.z, dv2 = A.w;
    const float invDet = B.x, lu2 = B.y, lv2 = B.z, area = B.w;
    const float lu0 = C.x, lv0 = C.y;
    const float m1 = P1.w, m2 = P2.w, m3 = P3.w;
    f3 x1 = xyz(P1), x2 = xyz(P2), x3 = xyz(P3);
    f3 dx1 = sub(x2, x1);
    f3 dx2 = sub(x3, x1);
    f3 wu = mul(invDet, sub(mul(dv2, dx1), mul(dv1, dx2)));
    f3 wv = mul(invDet, add(mul(-du2, dx1), mul(du1, dx2)));
    float err, diag, lambda;
    f3 q1, q2, q3;
    // warp
    q2 = mul(invDet * dv2, wu);
    q3 = mul(-invDet * dv1, wu);
    q1 = sub(neg(q2), q3);
    diag = dot(mul(m1, q1), q1) + dot(mul(m2, q2), q2) + dot(mul(m3, q3), q3);
    err = 0.5f * (dot(wu, wu) - lu2);
    lambda = -err / diag;
    x1 = add(x1, mul(m1 * lambda, q1));
    x2 = add(x2, mul(m2 * lambda, q2));
    x3 = add(x3, mul(m3 * lambda, q3));
    // weft
    q2 = mul(-invDet * du2, wv);
    q3 = mul(invDet * du1, wv);
    q1 = sub(neg(q2), q3);
    diag = dot(mul(m1, q1), q1) + dot(mul(m2, q2), q2) + dot(mul(m3, q3), q3);
    err = 0.5f * (dot(wv, wv) - lv2);
    lambda = -err / diag;
    x1 = add(x1, mul(m1 * lambda, q1));
    x2 = add(x2, mul(m2 * lambda, q2));
    x3 = add(x3, mul(m3 * lambda, q3));
    // shear
    q2 = sub(mul(dv2, wv), mul(du2, wu));
    q3 = add(mul(-dv1, wv), mul(du1, wu));
    q1 = sub(neg(q2), q3);
    diag = invDet * (dot(mul(m1, q1), q1) + dot(mul(m2, q2), q2) + dot(mul(m3, q3), q3));
    err = dot(wu, wv);
    float s = 1.f / (1.f + lu0 * lv0 * epsUV / (diag * h2 * area));
    lambda = -s * err / diag;
    x1 = add(x1, mul(m1 * lambda, q1));
    x2 = add(x2, mul(m2 * lambda, q2));
    x3 = add(x3, mul(m3 * lambda, q3));
    P1 = f4(x1, m1);
    P2 = f4(x2, m2);
    P3 = f4(x3, m3);
}

// Physics/ClothPBD.cpp:179-224 with EvaluateAngle (110-139) and
// EvaluateGradient (141-177) — one dihedral bend projection. Returns the
// bend's error |s theta| when it was applied, else 0 (skipped bends and NaN
// errors do not count, as in the reference's `>` comparison).
__device__ __forceinline__ float bend_project(float4& P1, float4& P2, float4& P3, float4& P4, float sStiff)
{
    const float m1 = P1.w, m2 = P2.w, m3 = P3.w, m4 = P4.w;
    if (m1 + m2 + m3 + m4 == 0.f) return 0.f;
    const f3 x1 = xyz(P1), x2 = xyz(P2), x3 = xyz(P3), x4 = xyz(P4);
    f3 x12 = sub(x2, x1);
    float l = len(x12);
    f3 x13 = sub(x3, x1);
    f3 n1 = cross(x12, x13);
    float l1 = len(n1);
    f3 x14 = sub(x4, x1);
    f3 n2 = cross(x14, x12);
    float l2 = len(n2);
    // the reference computes 1/l1, 1/l2 and the unit normals in both its
    // angle and its gradient; they are the same values, computed once here
    // (unused when l1 or l2 is 0: the bend is then skipped)
    const float il1 = 1.f / l1, il2 = 1.f / l2;
    const f3 n1n = mul(il1, n1), n2n = mul(il2, n2);
    float cosine = 1.f, theta = 0.f;
    if (!(l1 == 0.f || l2 == 0.f)) {
        cosine = dot(n1n, n2n);
        cosine = smin(1.f, smax(-1.f, cosine));
        theta = acos_as(cosine);
    }
    if (fabsf(cosine) >= 1 - 0.001f) return 0.f;
    float h1 = l1 / l, h2 = l2 / l;
    float he = 0.5f * (h1 + h2);
    f3 n1p = sub(n1n, mul(cosine, n2n));
    f3 n2p = sub(n2n, mul(cosine, n1n));
    float sine = sqrtf(1 - cosine * cosine);
    // The reference's four factors (+-il1 / sine / he, +-il2 / sine / he) from
    // two division chains: rounding to nearest is symmetric in sign, so
    // (-a) / b is exactly -(a / b).
    const float f1 = il1 / sine / he, f2 = il2 / sine / he;
    f3 q3 = mul(f1, cross(x12, n2p));
    f3 q4 = mul(-f2, cross(x12, n1p));
    f3 q2 = add(mul(-f1, cross(x13, n2p)), mul(f2, cross(x14, n1p)));
    f3 q1 = sub(sub(neg(q2), q3), q4);
    float diag = m1 * dot(q1, q1) + m2 * dot(q2, q2) + m3 * dot(q3, q3) + m4 * dot(q4, q4);
    if (diag == 0.f) return 0.f;
    float e = fabsf(sStiff * theta);
    float lambda = -(sStiff * theta / he / diag);
    P1 = f4(add(x1, mul(lambda * m1, q1)), m1);
    P2 = f4(add(x2, mul(lambda * m2, q2)), m2);
    P3 = f4(add(x3, mul(lambda * m3, q3)), m3);
    P4 = f4(add(x4, mul(lambda * m4, q4)), m4);
    return e > 0.f ? e : 0.f;
}

// The value of the partner lane (lanes 2b and 2b + 1 trade; DPP quad_perm
// [1, 0, 3, 2]). Every lane of the wave must be active.
__device__ __forceinline__ float pair_swap(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xf, 0xf, false));
}
__device__ __forceinline__ f3 pair_swap(f3 v) { return f3{pair_swap(v.x), pair_swap(v.y), pair_swap(v.z)}; }
__device__ __forceinline__ f3 sel(bool c, f3 a, f3 b) { return f3{c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z}; }

// bend_project computed by a pair of adjacent lanes: lane A (side b = false)
// takes the triangle (x1, x2, x3), lane B the triangle (x1, x2, x4). Each lane
// forms its side's normal, height and gradient terms with the same
// instructions on its own operands, the pair trades them, and both evaluate
// the shared terms, so every value is computed by the same operations in the
// same order as in bend_project and the results agree to the bit. Returns
// the bend's error (as bend_project) and the lane's two updated vertices in
// Pa, Pb — (x1, x3) on lane A, (x2, x4) on lane B — with apply false when
// bend_project would leave the bend alone. Both lanes of every pair in the
// wave must be active.
__device__ __forceinline__ float bend_project_pair(const float4 P1, const float4 P2, const float4 P3,
                                                   const float4 P4, float sStiff, bool b, float4& Pa, float4& Pb,
                                                   bool& apply)
{
    const float m1 = P1.w, m2 = P2.w, m3 = P3.w, m4 = P4.w;
    const f3 x1 = xyz(P1), x2 = xyz(P2), x3 = xyz(P3), x4 = xyz(P4);
    const f3 x12 = sub(x2, x1);
    const float l = len(x12);
    // this lane's side: n1 = x12 x x13 (A) or n2 = x14 x x12 (B)
    const f3 xs = sub(b ? x4 : x3, x1);
    const f3 ns = cross(sel(b, xs, x12), sel(b, x12, xs));
    const float ls = len(ns);
    const float ils = 1.f / ls;
    const f3 nsn = mul(ils, ns);
    const float hs = ls / l;
    const f3 nso = pair_swap(nsn);
    const float lso = pair_swap(ls), hso = pair_swap(hs);
    // (IEEE products and sums are commutative, so dot(n1n, n2n) and h1 + h2
    // come out the same on both lanes from their own and the partner's terms)
    float cosine = 1.f, theta = 0.f;
    if (!(ls == 0.f || lso == 0.f)) {
        cosine = dot(nsn, nso);
        cosine = smin(1.f, smax(-1.f, cosine));
        theta = acos_as(cosine);
    }
    apply = m1 + m2 + m3 + m4 != 0.f && !(fabsf(cosine) >= 1 - 0.001f);
    const float he = 0.5f * (hs + hso);
    // the partner's n'p: n2p = n2n - cos n1n on lane A, n1p = n1n - cos n2n on lane B
    const f3 npo = sub(nso, mul(cosine, nsn));
    const float sine = sqrtf(1 - cosine * cosine);
    const float fs = ils / sine / he;  // f1 on lane A, f2 on lane B
    // q3 = f1 (x12 x n2p) and -f1 (x13 x n2p) on lane A; q4 = -f2 (x12 x n1p)
    // and f2 (x14 x n1p) on lane B
    const f3 qs = mul(b ? -fs : fs, cross(x12, npo));
    const f3 ts = mul(b ? fs : -fs, cross(xs, npo));
    const f3 qso = pair_swap(qs), tso = pair_swap(ts);
    const f3 q3 = sel(b, qso, qs), q4 = sel(b, qs, qso);
    const f3 q2 = add(ts, tso);
    const f3 q1 = sub(sub(neg(q2), q3), q4);
    const float diag = m1 * dot(q1, q1) + m2 * dot(q2, q2) + m3 * dot(q3, q3) + m4 * dot(q4, q4);
    if (diag == 0.f) apply = false;
    const float e = fabsf(sStiff * theta);
    const float lambda = -(sStiff * theta / he / diag);
    const float ma = b ? m2 : m1, mb = b ? m4 : m3;
    Pa = f4(add(b ? x2 : x1, mul(lambda * ma, b ? q2 : q1)), ma);
    Pb = f4(add(b ? x4 : x3, mul(lambda * mb, b ? q4 : q3)), mb);
    return apply && e > 0.f ? e : 0.f;
}

// Physics/ClothPBD.cpp:9-50 — one link projection (omega = 1). Returns
// |s err| when the link was projected (skipped links and NaN errors count 0,
// as in the reference's `>` comparison).
__device__ __forceinline__ float link_project(float4& P1, float4& P2, const float2 L, float k)
{
    const float m1 = P1.w, m2 = P2.w;
    if (m1 + m2 == 0.f) return 0.f;
    const float mu = 1.f / (m1 + m2);
    f3 delta = sub(xyz(P1), xyz(P2));
    const float len0 = L.x;
    const float l = len(delta);
    if (l == 0.f || isnan(l) || isinf(l)) return 0.f;
    const float err = l - len0;
    delta = mul(1.f / l, delta);
    const float epsilon = len0 / (k * L.y);
    const float s = 1.f / (1.f + mu * epsilon);
    const float lambda = err * mu * s;
    const float e = fabsf(s * err);
    const f3 fc = mul(1.f * lambda, delta);
    P1 = f4(sub(xyz(P1), mul(m1, fc)), m1);
    P2 = f4(add(xyz(P2), mul(m2, fc)), m2);
    return e > 0.f ? e : 0.f;
}

// Colour schedule: one thread per link of one colour.
__global__ void __launch_bounds__(kBlock) k_links(float4* __restrict__ pos, const int2* __restrict__ lidx,
                                                  const float2* __restrict__ lrest, int start, int count, float k,
                                                  SolverCtl* ctl)
{
    if (ctl->done) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    float err = 0.f;
    if (t < count) {
        const int2 id = lidx[start + t];
        float4 P1 = pos[id.x], P2 = pos[id.y];
        err = link_project(P1, P2, lrest[start + t], k);
        pos[id.x] = P1;
        pos[id.y] = P2;
    }
    wave_max_atomic(&ctl->cl, err);
}

// Colour schedule: one thread per triangle of one colour, positions in HBM.
__global__ void __launch_bounds__(kBlock) k_fem(float4* __restrict__ pos, TriStreams tr, int start, int count,
                                                float h2, float epsUV, const SolverCtl* __restrict__ ctl)
{
    if (ctl->done) return;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    t += start;
    const int4 id = tr.idx[t];
    float4 P1 = pos[id.x], P2 = pos[id.y], P3 = pos[id.z];
    fem_project(P1, P2, P3, tr.a[t], tr.b[t], tr.c[t], h2, epsUV);
    pos[id.x] = P1;
    pos[id.y] = P2;
    pos[id.z] = P3;
}

// Colour schedule: one thread per bend of one colour.
__global__ void __launch_bounds__(kBlock) k_bend(float4* __restrict__ pos, const int4* __restrict__ bidx, int start,
                                                 int count, float sStiff, SolverCtl* ctl)
{
    if (ctl->done) return;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    float errOut = 0.f;
    if (t < count) {
        const int4 b = bidx[start + t];
        float4 P1 = pos[b.x], P2 = pos[b.y], P3 = pos[b.z], P4 = pos[b.w];
        errOut = bend_project(P1, P2, P3, P4, sStiff);
        pos[b.x] = P1;
        pos[b.y] = P2;
        pos[b.z] = P3;
        pos[b.w] = P4;
    }
    wave_max_atomic(&ctl->cb, errOut);
}

// Contact slot s of vertex v (point CP, normal CN as stored) applied to its
// position p and velocity ve (ClothCollisionHandler.cpp:141-199); err keeps
// the largest depth.
__device__ __forceinline__ void vertex_contact_slot(int s, int v, float4 CP, float4 CN, f3& p, const f3 pp, f3& ve,
                                                    ContactBuf cb, float thickness, float mu, float invH, float& err)
{
    {
        if (cb.geo) {
            const int f = cb.feat[s * cb.nv + v];
            if (f >= 0) {
                // ClothCollisionHandler.cpp:145-160: re-project onto the same
                // obstacle triangle, normal interpolated from its vertex
                // normals (ComputeTriangleNormal, ClosestPointOnMesh.cpp:37-62)
                const MeshGeo& g = cb.geo[cb.src[s * cb.nv + v]];
                const int4 ix = g.tri[f];
                f3 cp, co;
                closest_pt_point_triangle(p, xyz(g.v[ix.x]), xyz(g.v[ix.y]), xyz(g.v[ix.z]), cp, co);
                f3 nn = add(add(mul(co.x, xyz(g.vn[ix.x])), mul(co.y, xyz(g.vn[ix.y]))), mul(co.z, xyz(g.vn[ix.z])));
                nn = normalized(nn);
                CP = f4(cp, CP.w);
                CN = f4(nn, CN.w);
            }
        }
        f3 cpt = xyz(CP), nn = xyz(CN);
        float lam = CP.w, lamF = CN.w;
        f3 delta = sub(p, cpt);
        float l = dot(delta, nn);
        if (l > thickness) {
            if (cb.geo) {
                cb.P[s * cb.nv + v] = CP;
                cb.N[s * cb.nv + v] = CN;
            }
            return;
        }
        float depth = thickness - l;
        if (fabsf(depth) > err) err = fabsf(depth);
        lam += depth;
        f3 disp = mul(depth, nn);
        f3 v12 = sub(p, pp);
        float vnrel = dot(nn, v12);
        f3 vt = sub(v12, mul(vnrel, nn));
        float vtrel = len(vt);
        const float limit = mu * lam;
        if (vtrel > 0.001f) {
            float dLF = vtrel;
            float lf0 = lamF;
            lamF = lf0 + dLF;
            if (lamF >= limit) lamF = limit;
            dLF = lamF - lf0;
            vt = mul(1.0f / vtrel, vt);
            disp = add(disp, mul(-dLF, vt));
        }
        p = add(p, disp);
        ve = add(ve, mul(invH, disp));
        cb.P[s * cb.nv + v] = make_float4(CP.x, CP.y, CP.z, lam);
        cb.N[s * cb.nv + v] = make_float4(CN.x, CN.y, CN.z, lamF);
    }
}

// The contacts of vertex v, in slot order, applied to its position p and
// velocity ve (ClothCollisionHandler.cpp:141-199). Returns the largest depth.
__device__ __forceinline__ float vertex_contacts(int v, int n, f3& p, const f3 pp, f3& ve, ContactBuf cb,
                                                 float thickness, float mu, float invH)
{
    float err = 0.f;
    for (int s = 0; s < n; s++)
        vertex_contact_slot(s, v, cb.P[s * cb.nv + v], cb.N[s * cb.nv + v], p, pp, ve, cb, thickness, mu, invH, err);
    return err;
}

// Slab schedule: one workgroup per slab of the phase. The slab's vertices
// are staged in LDS, its local colours are swept with a barrier between
// colours, and the vertices are written back. Slabs of one phase touch
// disjoint vertices, so workgroups never race. The constraint data of colour
// c+1 is loaded into registers while colour c is projected, so the sweep
// waits on global memory once per launch instead of once per colour.
// Slab vertex ids carry an owner flag in bit 31: the vertex's first slab in
// phase order. With `kc.enabled`, the owner applies the vertex's contacts of
// the previous iteration while staging it — after every constraint of that
// iteration (earlier launches) and before any constraint of this one, which
// is where the reference's sequential loop puts them.

// Keeps the loads above it above the early exit below it: without it the
// compiler tests the exit flag first and only then issues the loads, one
// memory latency after another.
__device__ __forceinline__ void issue_before_exit() { asm volatile("" ::: "memory"); }

// The slab's descriptor and its first order position (two independent
// loads): every later load of the launch depends only on these, so the
// kernels issue the first colour's constraint loads before staging the
// vertices and both are in flight together.
__device__ __forceinline__ int4 slab_info(const SlabSet& ss, int& first)
{
    first = ss.first[ss.base + blockIdx.x];
    return ss.info[ss.base + blockIdx.x];
}

// Stages the slab's vertices in LDS. The global ids of this thread's first
// kSlabBatch vertices are kept in `gk` for slab_end (all of a slab's ids when
// it has at most kSlabBatch * kSlabBlock vertices). The slab kernels are
// launched with kSlabBlock threads, and their index arithmetic uses that
// constant.
__device__ __forceinline__ int slab_begin(const SlabSet& ss, const int4& info, int* scol, float4* sv,
                                          const float4* pos, const SlabContacts& kc, int (&gk)[kSlabBatch])
{
    const int* vg = ss.verts + info.x;
    // a slab has at most kMaxSlabColours colours, fewer than the workgroup's
    // threads: one offset per thread, held in a register until the vertices
    // are staged so that its load does not hold up theirs
    static_assert(kSlabBlock > kMaxSlabColours, "one colour offset per thread");
    const bool hasCol = (int)threadIdx.x <= info.w;
    const int colOff = ss.colOff[info.z + min((int)threadIdx.x, info.w)];
    // vertex ids, then positions, kSlabBatch per thread at a time: the loads
    // of one batch are in flight together instead of one after another. Loads
    // past the slab's end repeat its last vertex (a slab has at least one)
    // and their stores go to the spare LDS slot sv[info.y] (the launch
    // allocates one slot more than the largest slab), so nothing is
    // conditional and the compiler keeps the loads together.
    const int last = info.y - 1;
#pragma unroll
    for (int u = 0; u < kSlabBatch; u++) gk[u] = 0;
    for (int v0 = threadIdx.x; v0 < info.y; v0 += kSlabBatch * kSlabBlock) {
        int raw[kSlabBatch];
        float4 P[kSlabBatch];
#pragma unroll
        for (int u = 0; u < kSlabBatch; u++) raw[u] = vg[min(v0 + u * kSlabBlock, last)];
#pragma unroll
        for (int u = 0; u < kSlabBatch; u++) P[u] = pos[raw[u] & kSlabIdMask];
#pragma unroll
        for (int u = 0; u < kSlabBatch; u++) {
            const int v = v0 + u * kSlabBlock;
            sv[v < info.y ? v : info.y] = P[u];
        }
        if (v0 == (int)threadIdx.x) {
#pragma unroll
            for (int u = 0; u < kSlabBatch; u++) gk[u] = raw[u];
        }
    }
    if (kc.enabled == 1) {
        // the owner applies the vertex's contacts (same thread-to-vertex map
        // as the staging loop, so no barrier is needed in between)
        for (int v = threadIdx.x; v < info.y; v += kSlabBlock) {
            const int raw = vg[v];
            if (raw >= 0) continue;
            const int g = raw & kSlabIdMask;
            const int n = kc.cb.count[g];
            if (n == 0) continue;
            const float4 P = sv[v];
            f3 p = xyz(P), ve = xyz(kc.vel[g]);
            vertex_contacts(g, n, p, xyz(kc.prev[g]), ve, kc.cb, kc.thickness, kc.mu, kc.invH);
            sv[v] = f4(p, P.w);
            kc.vel[g] = f4(ve, 0.f);
        }
    }
    if (hasCol) scol[threadIdx.x] = colOff;
    __syncthreads();
    return info.w;
}

// Writes the slab's vertices back (kOwnedOnly: only those whose id carries
// the owner flag, as a tile writes back its own vertices but not its halo).
template <bool kOwnedOnly = false>
__device__ __forceinline__ void slab_end(const SlabSet& ss, const int4& info, const float4* sv, float4* pos,
                                         const int (&gk)[kSlabBatch])
{
    __syncthreads();
    const int batch = kSlabBatch * kSlabBlock;
    if (info.y <= batch) {
        // the ids staged by this thread, kept in registers
#pragma unroll
        for (int u = 0; u < kSlabBatch; u++) {
            const int v = threadIdx.x + u * kSlabBlock;
            if (v < info.y && (!kOwnedOnly || gk[u] < 0)) pos[gk[u] & kSlabIdMask] = sv[v];
        }
        return;
    }
    const int* vg = ss.verts + info.x;
    const int last = info.y - 1;
    for (int v0 = threadIdx.x; v0 < info.y; v0 += batch) {
        int raw[kSlabBatch];
#pragma unroll
        for (int u = 0; u < kSlabBatch; u++) raw[u] = vg[min(v0 + u * kSlabBlock, last)];
#pragma unroll
        for (int u = 0; u < kSlabBatch; u++) {
            const int v = v0 + u * kSlabBlock;
            if (v < info.y && (!kOwnedOnly || raw[u] < 0)) pos[raw[u] & kSlabIdMask] = sv[v];
        }
    }
}

// The iteration's vertex contacts of the slab vertices whose id carries
// `flag`, applied after the slab's last colour: a tile's own vertices
// (kSlabOwner, FEM tiles, kc.enabled == 2) or each vertex in the last bend
// slab that touches it (kSlabContacts, kc.enabled == 3). No other workgroup
// of the launch, nor any later constraint of the iteration, touches these
// vertices, and no other workgroup reads their contact state.
__device__ __forceinline__ void slab_contacts_end(const SlabSet& ss, const int4& info, float4* sv,
                                                  const SlabContacts& kc, SolverCtl* ctl, unsigned flag)
{
    const int* vg = ss.verts + info.x;
    float err = 0.f;
    for (int v = threadIdx.x; v < info.y; v += kSlabBlock) {
        const int raw = vg[v];
        if (!((unsigned)raw & flag)) continue;
        const int g = raw & kSlabIdMask;
        const int n = kc.cb.count[g];
        if (n == 0) continue;
        const float4 P = sv[v];
        f3 p = xyz(P), ve = xyz(kc.vel[g]);
        const float e = vertex_contacts(g, n, p, xyz(kc.prev[g]), ve, kc.cb, kc.thickness, kc.mu, kc.invH);
        err = err < e ? e : err;
        sv[v] = f4(p, P.w);
        kc.vel[g] = f4(ve, 0.f);
    }
    wave_max_atomic(&ctl->cv, err);
}

// One colour's constraint for this thread, fetched a colour ahead. Indices
// past the colour are clamped to `last` (the slab's last constraint, or the
// schedule's for the first colour), so the loads are unconditional and stay
// in flight across the LDS barrier; the thread has a constraint in the
// colour when its position t is below the colour's end.
struct FemItem {
    ushort4 id;
    float4 A, B;
    float2 C;
    int t;
};

__device__ __forceinline__ FemItem fem_fetch(TriStreams tr, const ushort4* __restrict__ lidx, int t, int last)
{
    FemItem f;
    const int u = min(t, last);
    f.id = lidx[u];
    f.A = tr.a[u];
    f.B = tr.b[u];
    f.C = tr.c[u];
    f.t = t;
    return f;
}

// The thread's constraints of one colour ending at `e`: the prefetched one,
// then (a colour wider than the workgroup) the rest, in order.
__device__ __forceinline__ void fem_colour(const FemItem& f, int e, float4* sv, TriStreams tr,
                                           const ushort4* __restrict__ lidx, float h2, float epsUV)
{
    if (f.t < e) {
        float4 P1 = sv[f.id.x], P2 = sv[f.id.y], P3 = sv[f.id.z];
        fem_project(P1, P2, P3, f.A, f.B, f.C, h2, epsUV);
        sv[f.id.x] = P1;
        sv[f.id.y] = P2;
        sv[f.id.z] = P3;
    }
    for (int t2 = f.t + kSlabBlock; t2 < e; t2 += kSlabBlock) {
        const ushort4 i2 = lidx[t2];
        float4 P1 = sv[i2.x], P2 = sv[i2.y], P3 = sv[i2.z];
        fem_project(P1, P2, P3, tr.a[t2], tr.b[t2], tr.c[t2], h2, epsUV);
        sv[i2.x] = P1;
        sv[i2.y] = P2;
        sv[i2.z] = P3;
    }
}

// kTiles: the slabs are tiles of a tile schedule (extended by their halos):
// only the tile's own vertices are written back, and with kc.enabled == 2
// their contacts of this iteration are applied after the last colour.
template <bool kTiles>
__global__ void __launch_bounds__(kSlabBlock) k_fem_slab(float4* __restrict__ pos, TriStreams tr,
                                                         const ushort4* __restrict__ lidx, SlabSet ss, float h2,
                                                         float epsUV, SolverCtl* __restrict__ ctl, SlabContacts kc)
{
    extern __shared__ float4 sv[];
    __shared__ int scol[kMaxSlabColours + 1];
    const int done = ctl->done;
    int first;
    const int4 info = slab_info(ss, first);
    // two colours per trip with the two items in turn: no register copies
    // of fetched data (a copy would wait for the load before the barrier)
    FemItem a = fem_fetch(tr, lidx, first + threadIdx.x, ss.count - 1), b;
    issue_before_exit();
    if (done) return;
    int gk[kSlabBatch];
    const int nc = slab_begin(ss, info, scol, sv, pos, kc, gk);
    const int last = scol[nc] - 1;
    for (int c = 0; c < nc; c += 2) {
        const bool n1 = c + 1 < nc;
        b = fem_fetch(tr, lidx, scol[c + 1] + threadIdx.x, last);
        fem_colour(a, scol[c + 1], sv, tr, lidx, h2, epsUV);
        lds_barrier();
        if (!n1) break;
        a = fem_fetch(tr, lidx, scol[c + 2] + threadIdx.x, last);
        fem_colour(b, scol[c + 2], sv, tr, lidx, h2, epsUV);
        lds_barrier();
    }
    if (kTiles && kc.enabled == 2) slab_contacts_end(ss, info, sv, kc, ctl, kSlabOwner);
    slab_end<kTiles>(ss, info, sv, pos, gk);
}

struct QuadItem {
    ushort4 id;
    int t;
};

__device__ __forceinline__ QuadItem quad_fetch(const ushort4* __restrict__ lidx, int t, int last)
{
    QuadItem f;
    f.id = lidx[min(t, last)];
    f.t = t;
    return f;
}

__device__ __forceinline__ void bend_colour(const QuadItem& f, int e, float4* sv, const ushort4* __restrict__ lidx,
                                            float sStiff, float& err)
{
    if (f.t < e) {
        float4 P1 = sv[f.id.x], P2 = sv[f.id.y], P3 = sv[f.id.z], P4 = sv[f.id.w];
        const float e1 = bend_project(P1, P2, P3, P4, sStiff);
        err = err < e1 ? e1 : err;
        sv[f.id.x] = P1;
        sv[f.id.y] = P2;
        sv[f.id.z] = P3;
        sv[f.id.w] = P4;
    }
    for (int t2 = f.t + kSlabBlock; t2 < e; t2 += kSlabBlock) {
        const ushort4 i2 = lidx[t2];
        float4 P1 = sv[i2.x], P2 = sv[i2.y], P3 = sv[i2.z], P4 = sv[i2.w];
        const float e1 = bend_project(P1, P2, P3, P4, sStiff);
        err = err < e1 ? e1 : err;
        sv[i2.x] = P1;
        sv[i2.y] = P2;
        sv[i2.z] = P3;
        sv[i2.w] = P4;
    }
}

// With kc.enabled == 3, the vertices whose last bend slab this is get the
// iteration's vertex contacts after the last colour.
__global__ void __launch_bounds__(kSlabBlock) k_bend_slab(float4* __restrict__ pos, const ushort4* __restrict__ lidx,
                                                          SlabSet ss, float sStiff, SolverCtl* ctl, SlabContacts kc)
{
    extern __shared__ float4 sv[];
    __shared__ int scol[kMaxSlabColours + 1];
    const int done = ctl->done;
    int first;
    const int4 info = slab_info(ss, first);
    QuadItem a = quad_fetch(lidx, first + threadIdx.x, ss.count - 1), b;
    issue_before_exit();
    if (done) return;
    SlabContacts none;
    none.enabled = 0;
    int gk[kSlabBatch];
    const int nc = slab_begin(ss, info, scol, sv, pos, none, gk);
    const int last = scol[nc] - 1;
    float err = 0.f;
    for (int c = 0; c < nc; c += 2) {
        const bool n1 = c + 1 < nc;
        b = quad_fetch(lidx, scol[c + 1] + threadIdx.x, last);
        bend_colour(a, scol[c + 1], sv, lidx, sStiff, err);
        lds_barrier();
        if (!n1) break;
        a = quad_fetch(lidx, scol[c + 2] + threadIdx.x, last);
        bend_colour(b, scol[c + 2], sv, lidx, sStiff, err);
        lds_barrier();
    }
    if (kc.enabled == 3) slab_contacts_end(ss, info, sv, kc, ctl, kSlabContacts);
    slab_end(ss, info, sv, pos, gk);
    wave_max_atomic(&ctl->cb, err);
}

struct LinkItem {
    ushort4 id;
    float2 L;
    int t;
};

__device__ __forceinline__ LinkItem link_fetch(const ushort4* __restrict__ lidx, const float2* __restrict__ lrest,
                                               int t, int last)
{
    LinkItem f;
    const int u = min(t, last);
    f.id = lidx[u];
    f.L = lrest[u];
    f.t = t;
    return f;
}

__device__ __forceinline__ void link_colour(const LinkItem& f, int e, float4* sv, const ushort4* __restrict__ lidx,
                                            const float2* __restrict__ lrest, float k, float& err)
{
    if (f.t < e) {
        float4 P1 = sv[f.id.x], P2 = sv[f.id.y];
        const float e1 = link_project(P1, P2, f.L, k);
        err = err < e1 ? e1 : err;
        sv[f.id.x] = P1;
        sv[f.id.y] = P2;
    }
    for (int t2 = f.t + kSlabBlock; t2 < e; t2 += kSlabBlock) {
        const ushort4 i2 = lidx[t2];
        float4 P1 = sv[i2.x], P2 = sv[i2.y];
        const float e1 = link_project(P1, P2, lrest[t2], k);
        err = err < e1 ? e1 : err;
        sv[i2.x] = P1;
        sv[i2.y] = P2;
    }
}

// Slab schedule for links: as k_bend_slab with two vertices per constraint.
__global__ void __launch_bounds__(kSlabBlock) k_link_slab(float4* __restrict__ pos, const ushort4* __restrict__ lidx,
                                                          const float2* __restrict__ lrest, SlabSet ss, float k,
                                                          SolverCtl* ctl)
{
    extern __shared__ float4 sv[];
    __shared__ int scol[kMaxSlabColours + 1];
    const int done = ctl->done;
    int first;
    const int4 info = slab_info(ss, first);
    LinkItem a = link_fetch(lidx, lrest, first + threadIdx.x, ss.count - 1), b;
    issue_before_exit();
    if (done) return;
    SlabContacts none;
    none.enabled = 0;
    int gk[kSlabBatch];
    const int nc = slab_begin(ss, info, scol, sv, pos, none, gk);
    const int last = scol[nc] - 1;
    float err = 0.f;
    for (int c = 0; c < nc; c += 2) {
        const bool n1 = c + 1 < nc;
        b = link_fetch(lidx, lrest, scol[c + 1] + threadIdx.x, last);
        link_colour(a, scol[c + 1], sv, lidx, lrest, k, err);
        lds_barrier();
        if (!n1) break;
        a = link_fetch(lidx, lrest, scol[c + 2] + threadIdx.x, last);
        link_colour(b, scol[c + 2], sv, lidx, lrest, k, err);
        lds_barrier();
    }
    slab_end(ss, info, sv, pos, gk);
    wave_max_atomic(&ctl->cl, err);
}

// ---------------------------------------------------------------------------
// Packed sweeps. A slab launch above waits on memory once per colour: the
// colour's constraint data is requested one colour ahead, and a colour step
// (one projection and an LDS barrier) is shorter than a load's latency. With
// a packed schedule (PackedFem, PackedBend) every address of the launch
// follows from the workgroup and thread, so a launch requests its vertex ids,
// then the positions they name and the contact state of the vertices it
// applies contacts to, then the constraint data of every colour at once, and
// the colour steps run behind the data: two memory latencies per launch.
// Padding slots compute on the spare LDS entry, so nothing is conditional.

// Keeps the scheduler from moving loads across this point, so that loads are
// issued in the order the waits below rely on (memory returns in order).
__device__ __forceinline__ void issue_fence()
{
    asm volatile("" ::: "memory");
    __builtin_amdgcn_sched_barrier(0);
}

// A 16-byte load as four scalars (an aggregate copy would go through private
// memory across issue_fence).
__device__ __forceinline__ float4 load4(const float4* p)
{
    const float* f = reinterpret_cast<const float*>(p);
    return make_float4(f[0], f[1], f[2], f[3]);
}

// The first two contact slots of a vertex, its previous position and
// velocity, loaded ahead of the contact pass (slot 1 only when the buffer
// has it; a vertex rarely holds more than two contacts, and later slots are
// loaded when applied).
struct ContactPre {
    int cnt;
    float4 C0P, C0N, C1P, C1N, PV, VE;
};

// The rest of the state of a vertex with cnt > 0 contacts (its count is
// requested first, with the positions).
__device__ __forceinline__ void contact_pre(int g, ContactPre& c, const SlabContacts& kc)
{
    c.C0P = load4(kc.cb.P + g);
    c.C0N = load4(kc.cb.N + g);
    const int g1 = c.cnt > 1 ? kc.cb.nv + g : g;
    c.C1P = load4(kc.cb.P + g1);
    c.C1N = load4(kc.cb.N + g1);
    c.PV = load4(kc.prev + g);
    c.VE = load4(kc.vel + g);
}

// The vertex contacts of one staged vertex (position in LDS at sv[slot]) from
// its preloaded state (ClothCollisionHandler.cpp:141-199, slots in order).
__device__ __forceinline__ float packed_vertex_contacts(int g, const ContactPre& c, float4* sv, int slot,
                                                        const SlabContacts& kc)
{
    const float4 P = sv[slot];
    f3 p = xyz(P), ve = xyz(c.VE);
    const f3 pp = xyz(c.PV);
    float err = 0.f;
    if (c.cnt > 0) vertex_contact_slot(0, g, c.C0P, c.C0N, p, pp, ve, kc.cb, kc.thickness, kc.mu, kc.invH, err);
    if (c.cnt > 1) vertex_contact_slot(1, g, c.C1P, c.C1N, p, pp, ve, kc.cb, kc.thickness, kc.mu, kc.invH, err);
    for (int s = 2; s < c.cnt; s++)
        vertex_contact_slot(s, g, kc.cb.P[s * kc.cb.nv + g], kc.cb.N[s * kc.cb.nv + g], p, pp, ve, kc.cb,
                            kc.thickness, kc.mu, kc.invH, err);
    sv[slot] = f4(p, P.w);
    kc.vel[g] = f4(ve, 0.f);
    return err;
}

// One colour's FEM constraint slots of this thread.
struct FemSlot {
    uint4 a0;
    float4 a1;
};

__device__ __forceinline__ FemSlot fem_slot(const PackedFem& pk, size_t k)
{
    return FemSlot{pk.s0[k], load4(pk.s1 + k)};
}

// The rest data of a packed FEM slot as fem_project takes it: the slot holds
// du1, du2, dv1, dv2, invDet, lu0 and lv0; area = det / 2, lu2 = lu0^2 and
// lv2 = lv0^2 are recomputed with the operations the host uses
// (ClothModel.cpp:54-92), so they are the same values to the bit.
__device__ __forceinline__ void fem_slot_rest(const FemSlot& f, float4& A, float4& B, float2& C)
{
    A = f.a1;
    const float invDet = __uint_as_float(f.a0.y), lu0 = __uint_as_float(f.a0.z), lv0 = __uint_as_float(f.a0.w);
    const float det = A.x * A.w - A.y * A.z;
    B = make_float4(invDet, lu0 * lu0, lv0 * lv0, 0.5f * det);
    C = make_float2(lu0, lv0);
}

// FEM triangles on packed tiles (NC colours, KV vertex slots per thread).
// The constraint data streams kFemAhead colours ahead of the sweep: requested
// all at once it would queue behind the position gathers and hold the
// workgroup's first colour back. kContacts: the tile's own vertices get this
// iteration's vertex contacts after the last colour (as k_fem_slab<true> with
// kc.enabled == 2); their contact state is requested with the last colours.
constexpr int kFemAhead = 3;

template <int NC, int KV, bool kContacts>
__global__ void __launch_bounds__(kSlabBlock) k_fem_packed(float4* pos, PackedFem pk, float h2,
                                                           float epsUV, SolverCtl* __restrict__ ctl, SlabContacts kc)
{
    static_assert(NC > kFemAhead, "colours beyond the stream distance");
    constexpr int kSpare = KV * kSlabBlock;  // vertex slots; the last is padding and the spare entry
    __shared__ float4 sv[kSpare];
    const int tid = threadIdx.x;
    const int done = ctl->done;
    const int* vg = pk.verts + (size_t)blockIdx.x * kSpare;
    int raw[KV];
#pragma unroll
    for (int u = 0; u < KV; u++) raw[u] = vg[tid + u * kSlabBlock];
    issue_fence();
    if (done) return;
    float4 P[KV];
#pragma unroll
    for (int u = 0; u < KV; u++) P[u] = load4(pos + (raw[u] & kPackIdMask));
    // slot 0 holds the thread's own vertex, if any, whose contact count comes
    // with the positions and the rest of whose contact state comes later
    const int g0 = raw[0] & kPackIdMask;
    ContactPre cp;
    cp.cnt = 0;
    if (kContacts && raw[0] < 0) cp.cnt = kc.cb.count[g0];
    issue_fence();
    FemSlot f[NC];
    const size_t sb = (size_t)blockIdx.x * NC * kSlabBlock + tid;
#pragma unroll
    for (int c = 0; c < kFemAhead; c++) f[c] = fem_slot(pk, sb + c * kSlabBlock);
    issue_fence();
#pragma unroll
    for (int u = 0; u < KV; u++) sv[tid + u * kSlabBlock] = P[u];
    lds_barrier();
#pragma unroll
    for (int c = 0; c < NC; c++) {
        if (c + kFemAhead < NC) f[c + kFemAhead] = fem_slot(pk, sb + (c + kFemAhead) * kSlabBlock);
        // every lane requests (vertex 0's state where there is no contact): a
        // load under a branch makes the colour steps after it wait for it
        if (kContacts && c + kFemAhead == NC) contact_pre(cp.cnt > 0 || kc.velocity ? g0 : 0, cp, kc);
        issue_fence();
        // (a branch per colour keeps each colour's arithmetic, and the wait
        // for its data, in its own block)
        if (c < pk.nc) {
            const int i1 = f[c].a0.x & 0x3ff, i2 = (f[c].a0.x >> 10) & 0x3ff, i3 = f[c].a0.x >> 20;
            float4 P1 = sv[i1], P2 = sv[i2], P3 = sv[i3];
            float4 A, B;
            float2 C;
            fem_slot_rest(f[c], A, B, C);
            fem_project(P1, P2, P3, A, B, C, h2, epsUV);
            sv[i1] = P1;
            sv[i2] = P2;
            sv[i3] = P3;
        }
        lds_barrier();
    }
    const bool own = raw[0] < 0;  // kSlabOwner
    if (kContacts) {
        float err = 0.f;
        if (own && cp.cnt > 0) err = packed_vertex_contacts(g0, cp, sv, tid, kc);
        if (kc.report & 1) wave_max_atomic(err_part(ctl, 1), err);
    }
    if (own) {
        const float4 P = sv[tid];
        pos[g0] = P;
        // (after the contacts' own velocity update, as the velocity launch)
        if (kContacts && kc.velocity && P.w != 0.f) kc.vel[g0] = f4(mul(kc.invH, sub(xyz(P), xyz(cp.PV))), 0.f);
    }
    if (kContacts && kc.velocity && blockIdx.x == 0 && tid == 0) ctl->maxTv = 0;  // as the velocity launch
}

// Bends on packed slabs (KV vertex slots per thread, at most
// kPackMaxBendColours colours). kContacts: the vertices whose last bend slab
// this is get the iteration's vertex contacts after the last colour (as
// k_bend_slab with kc.enabled == 3).
// kPair: each bend on a pair of adjacent lanes (bend_project_pair; the
// phase's width is at most half the workgroup, rounded so that whole waves
// are active or idle).
template <int KV, bool kContacts, bool kPair>
__global__ void __launch_bounds__(kSlabBlock) k_bend_packed(float4* pos, PackedBend pk, float sStiff,
                                                            SolverCtl* __restrict__ ctl, SlabContacts kc)
{
    constexpr int kSpare = KV * kSlabBlock;
    __shared__ float4 sv[kSpare + 1];
    const int tid = threadIdx.x;
    const int done = ctl->done;
    const int nc = pk.nc;
    const int slot = kPair ? tid >> 1 : tid;
    const bool lane = slot < pk.width;
    const int* vg = pk.verts + (size_t)blockIdx.x * kSpare;
    int raw[KV];
#pragma unroll
    for (int u = 0; u < KV; u++) raw[u] = vg[tid + u * kSlabBlock];
    uint2 q[kPackMaxBendColours];
    const size_t qb = (size_t)blockIdx.x * nc * pk.width + slot;
#pragma unroll
    for (int c = 0; c < kPackMaxBendColours; c++)
        if (lane && c < nc) q[c] = pk.ids[qb + (size_t)c * pk.width];
    issue_fence();
    if (done) return;
    float4 P[KV];
#pragma unroll
    for (int u = 0; u < KV; u++) P[u] = load4(pos + (raw[u] & kPackIdMask));
    // the contact counts of the vertices whose last bend slab this is come
    // with the positions; the rest of their contact state, needed after the
    // last colour, is requested during the sweep, one vertex slot per colour
    // (counted from the first, whatever the slab's colour count), so that it
    // holds up neither the staging nor any one colour step
    static_assert(!kContacts || KV <= 4, "contact slots requested over the first colours");
    ContactPre cp[KV];
#pragma unroll
    for (int u = 0; u < KV; u++) {
        cp[u].cnt = 0;
        if (kContacts && ((unsigned)raw[u] & kSlabContacts)) cp[u].cnt = kc.cb.count[raw[u] & kPackIdMask];
    }
    issue_fence();
#pragma unroll
    for (int u = 0; u < KV; u++) sv[tid + u * kSlabBlock] = P[u];
    lds_barrier();
    float err = 0.f;
#pragma unroll
    for (int c = 0; c < kPackMaxBendColours; c++) {
        if (kContacts && c < KV) {
            if (cp[c].cnt > 0) contact_pre(raw[c] & kPackIdMask, cp[c], kc);
            issue_fence();
        }
        if (c >= nc) continue;  // (not break: the loop must unroll, q lives in registers)
        if (lane) {
            const int i1 = q[c].x & 0xffff, i2 = q[c].x >> 16, i3 = q[c].y & 0xffff, i4 = q[c].y >> 16;
            float4 P1 = sv[i1], P2 = sv[i2], P3 = sv[i3], P4 = sv[i4];
            if (kPair) {
                const bool b = tid & 1;
                float4 Pa, Pb;
                bool apply;
                const float e1 = bend_project_pair(P1, P2, P3, P4, sStiff, b, Pa, Pb, apply);
                err = err < e1 ? e1 : err;
                if (apply) {
                    sv[b ? i2 : i1] = Pa;
                    sv[b ? i4 : i3] = Pb;
                }
            } else {
                const float e1 = bend_project(P1, P2, P3, P4, sStiff);
                err = err < e1 ? e1 : err;
                sv[i1] = P1;
                sv[i2] = P2;
                sv[i3] = P3;
                sv[i4] = P4;
            }
        }
        lds_barrier();
    }
    if (kContacts) {
        float cerr = 0.f;
#pragma unroll
        for (int u = 0; u < KV; u++) {
            if (((unsigned)raw[u] & kSlabContacts) && cp[u].cnt > 0) {
                const float e = packed_vertex_contacts(raw[u] & kPackIdMask, cp[u], sv, tid + u * kSlabBlock, kc);
                cerr = cerr < e ? e : cerr;
            }
        }
        if (kc.report & 1) wave_max_atomic(err_part(ctl, 1), cerr);
    }
#pragma unroll
    for (int u = 0; u < KV; u++)
        if (!((unsigned)raw[u] & kPackPad)) pos[raw[u] & kPackIdMask] = sv[tid + u * kSlabBlock];
    if (kc.report & 2) wave_max_atomic(err_part(ctl, 0), err);
}

// Physics/ClothCollisionHandler.cpp:134-202 — vertex contacts with friction.
// One thread per vertex applies that vertex's contacts in list order;
// contacts of different vertices commute, so this equals the reference's
// sequential pass over the contact list.
__global__ void __launch_bounds__(kBlock) k_contacts(float4* __restrict__ pos, const float4* __restrict__ prev,
                                                     float4* __restrict__ vel, ContactBuf cb, float thickness,
                                                     float mu, float invH, SolverCtl* ctl)
{
    if (ctl->done) return;
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    float err = 0.f;
    int n = v < cb.nv ? cb.count[v] : 0;
    if (n > 0) {
        float4 P = pos[v];
        f3 p = xyz(P);
        f3 ve = xyz(vel[v]);
        err = vertex_contacts(v, n, p, xyz(prev[v]), ve, cb, thickness, mu, invH);
        pos[v] = f4(p, P.w);
        vel[v] = f4(ve, 0.f);
    }
    wave_max_atomic(&ctl->cv, err);
}

// The last iteration's vertex contacts followed by the velocity update
// (ClothPBD.cpp:348, 314-321), one thread per vertex.
__global__ void __launch_bounds__(kBlock) k_contacts_velocity(float4* __restrict__ pos, const float4* __restrict__ prev,
                                                              float4* __restrict__ vel, ContactBuf cb, float thickness,
                                                              float mu, float invH, SolverCtl* ctl)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v == 0) ctl->maxTv = 0;  // accumulated afresh by the next prediction
    float err = 0.f;
    if (v < cb.nv) {
        float4 P = pos[v];
        const f3 pp = xyz(prev[v]);
        f3 p = xyz(P);
        f3 ve = xyz(vel[v]);
        const int n = cb.count[v];
        if (n > 0) {
            err = vertex_contacts(v, n, p, pp, ve, cb, thickness, mu, invH);
            pos[v] = f4(p, P.w);
        }
        if (P.w != 0.f) ve = mul(invH, sub(p, pp));
        if (n > 0 || P.w != 0.f) vel[v] = f4(ve, 0.f);
    }
    wave_max_atomic(&ctl->cv, err);  // the last iteration's vertex-contact error
}

// Physics/ClothPBD.cpp:354-361 — iteration counter and the early-exit test.
// The FEM model reports no link error, so that term of the test is 0.
__global__ void k_iter_end(SolverCtl* ctl, int maxIter, int fixedIter, int bendFreq, int dihedral, float selfThr)
{
    // one wave: lane t folds in part t of the bend and contact errors
    const int t = threadIdx.x;
    if (ctl->done) return;
    const unsigned pb = wave_max(ctl->errPart[0][t * kErrStride]), pv = wave_max(ctl->errPart[1][t * kErrStride]);
    const int k = ctl->k + 1;
    float cb = __uint_as_float(max(ctl->cb, pb)), cv = __uint_as_float(max(ctl->cv, pv));
    const float ce = __uint_as_float(ctl->ce), ct = __uint_as_float(ctl->ct);
    const float sct = __uint_as_float(ctl->sct), sce = __uint_as_float(ctl->sce);
    const float cl = __uint_as_float(ctl->cl), eps = 0.01f;
    bool exitNow = !fixedIter && cl < 0.05f && cb < 0.005f && sce < selfThr && sct < selfThr && cv < eps &&
                   ce < eps && ct < eps;
    const bool resetBend = dihedral && k % bendFreq == 0;
    if (!(exitNow || k >= maxIter)) {
        ctl->errPart[1][t * kErrStride] = 0;
        if (resetBend) ctl->errPart[0][t * kErrStride] = 0;
    }
    if (t != 0) return;
    ctl->k = k;
    if (exitNow || k >= maxIter) {
        ctl->done = 1;
        ctl->iterations = k;
        ctl->bendError = cb;
        ctl->contactError = cv;
        return;
    }
    ctl->cv = 0;
    ctl->ce = 0;
    ctl->ct = 0;
    ctl->sct = 0;
    ctl->sce = 0;
    ctl->cl = 0;
    if (resetBend) ctl->cb = 0;
}

// Fixed iteration count: the error statistics are those of the last pass of
// each kind (ClothPBD.cpp:337-361), so the accumulators are cleared before it.
// mask: 1 per-iteration errors (links, contacts, self), 2 the bend error.
__global__ void k_reset_errors(SolverCtl* ctl, int mask)
{
    const int t = threadIdx.x;  // one wave: lane t clears part t
    if (mask & 1) ctl->errPart[1][t * kErrStride] = 0;
    if (mask & 2) ctl->errPart[0][t * kErrStride] = 0;
    if (t != 0) return;
    if (mask & 1) {
        ctl->cl = 0;
        ctl->cv = 0;
        ctl->ce = 0;
        ctl->ct = 0;
        ctl->sct = 0;
        ctl->sce = 0;
    }
    if (mask & 2) ctl->cb = 0;
}

void launch_reset_errors(hipStream_t s, SolverCtl* ctl, int mask)
{
    hipLaunchKernelGGL(k_reset_errors, dim3(1), dim3(kErrParts), 0, s, ctl, mask);
}

// Physics/ClothPBD.cpp:314-321
__global__ void k_velocity(const float4* __restrict__ pos, const float4* __restrict__ prev, float4* __restrict__ vel,
                           int n, float invH, SolverCtl* ctl)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) ctl->maxTv = 0;  // accumulated afresh by the next prediction
    if (i >= n) return;
    float4 P = pos[i];
    if (P.w == 0.f) return;
    vel[i] = f4(mul(invH, sub(xyz(P), xyz(prev[i]))), 0.f);
}

// Plain streaming copy (16-byte loads and stores, grid-stride), used to
// measure the attainable HBM bandwidth and to calibrate the FETCH_SIZE and
// WRITE_SIZE counters against a kernel of known traffic.
__global__ void __launch_bounds__(kBlock) k_copy(const float4* __restrict__ src, float4* __restrict__ dst, long n)
{
    long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// Physics/ClothPBD.cpp:267-282: the barycentric point p of the spring's
// particles is pulled towards the mouse point; each particle moves by
// -h a_k disp and its velocity by -a_k disp, pinned particles included, one
// after another (as the reference's statements, should two indices agree).
__global__ void k_mouse_spring(float4* __restrict__ pos, float4* __restrict__ vel,
                               const MouseSpringDev* __restrict__ msp, float h, SolverCtl* ctl, int resetTravel)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const MouseSpringDev m = *msp;
    const f3 x1 = xyz(pos[m.i1]), x2 = xyz(pos[m.i2]), x3 = xyz(pos[m.i3]);
    const f3 p = add(add(mul(m.a1, x1), mul(m.a2, x2)), mul(m.a3, x3));
    f3 delta = sub(p, f3{m.px, m.py, m.pz});
    const float l = len(delta);
    delta = mul(1.f / l, delta);
    const f3 disp = mul(m.stiffness * (l - m.len), delta);
    const int ids[3] = {m.i1, m.i2, m.i3};
    const float as[3] = {m.a1, m.a2, m.a3};
    for (int k = 0; k < 3; k++) {
        const float4 P = pos[ids[k]];
        pos[ids[k]] = f4(sub(xyz(P), mul(h * as[k], disp)), P.w);
        const float4 V = vel[ids[k]];
        vel[ids[k]] = f4(sub(xyz(V), mul(as[k], disp)), V.w);
    }
    if (resetTravel) ctl->maxTv = 0;
}

__global__ void __launch_bounds__(kBlock) k_max_travel(const float4* __restrict__ pos,
                                                       const float4* __restrict__ prev, int n, SolverCtl* ctl)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float tv = 0.f;
    if (i < n) {
        tv = len(sub(xyz(pos[i]), xyz(prev[i])));
        tv = tv > 0.f ? tv : 0.f;  // a NaN never raises the reference's maximum
    }
    wave_max_atomic(&ctl->maxTv, tv);
}

void launch_mouse_spring(hipStream_t s, float4* pos, float4* vel, const MouseSpringDev* ms, float h, SolverCtl* ctl,
                         int resetTravel)
{
    hipLaunchKernelGGL(k_mouse_spring, dim3(1), dim3(64), 0, s, pos, vel, ms, h, ctl, resetTravel);
}

void launch_max_travel(hipStream_t s, const float4* pos, const float4* prev, int n, SolverCtl* ctl)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_max_travel, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, pos, prev, n, ctl);
}

void launch_copy(hipStream_t s, const float4* src, float4* dst, long n)
{
    if (n <= 0) return;
    long nb = (n + kBlock - 1) / kBlock;
    if (nb > 256 * 64) nb = 256 * 64;
    hipLaunchKernelGGL(k_copy, dim3((unsigned)nb), dim3(kBlock), 0, s, src, dst, n);
}

// ----------------------------------------------------------------- batch ----

// Workgroup-wide maximum of non-negative floats (every thread gets it).
template <int BS>
__device__ __forceinline__ float block_max(float v, float* red)
{
    for (int o = 32; o > 0; o >>= 1) {
        const float w = __shfl_xor(v, o, 64);
        v = v < w ? w : v;
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    float m = red[0];
    for (int k = 1; k < BS / 64; k++) m = m < red[k] ? red[k] : m;
    return m;
}

// The Gauss-Seidel iterations of one cloth on positions P (LDS or HBM):
// the same projections, in the same colour order, as the single-cloth
// colour-schedule kernels (k_fem, k_bend, k_contacts, k_iter_end).
template <int BS, class Pos>
__device__ __forceinline__ void batch_iterations(const ClothDesc& d, Pos P, float* red, int& itersOut, float& cbOut,
                                                 float& cvOut)
{
    float cb = 0.f, cv = 0.f;
    int k = 0;
    while (true) {
        for (int c = 0; c < d.triColours; c++) {
            const int e = d.triCol[c + 1];
            for (int t = d.triCol[c] + threadIdx.x; t < e; t += BS) {
                const int4 id = d.tr.idx[t];
                float4 P1 = P[id.x], P2 = P[id.y], P3 = P[id.z];
                fem_project(P1, P2, P3, d.tr.a[t], d.tr.b[t], d.tr.c[t], d.h2, d.epsUV);
                P[id.x] = P1;
                P[id.y] = P2;
                P[id.z] = P3;
            }
            __syncthreads();
        }
        if (d.dihedral && k % d.bendFreq == 0) {
            float err = 0.f;
            for (int c = 0; c < d.bendColours; c++) {
                const int e = d.bendCol[c + 1];
                for (int t = d.bendCol[c] + threadIdx.x; t < e; t += BS) {
                    const int4 b = d.bidx[t];
                    float4 P1 = P[b.x], P2 = P[b.y], P3 = P[b.z], P4 = P[b.w];
                    const float e1 = bend_project(P1, P2, P3, P4, d.sStiff);
                    err = err < e1 ? e1 : err;
                    P[b.x] = P1;
                    P[b.y] = P2;
                    P[b.z] = P3;
                    P[b.w] = P4;
                }
                __syncthreads();
            }
            cb = block_max<BS>(err, red);
        }
        if (d.anyContacts) {
            float err = 0.f;
            for (int v = threadIdx.x; v < d.nv; v += BS) {
                const int n = d.cb.count[v];
                if (n == 0) continue;
                float4 Q = P[v];
                f3 p = xyz(Q), ve = xyz(d.vel[v]);
                const float e1 = vertex_contacts(v, n, p, xyz(d.prev[v]), ve, d.cb, d.thickness, d.mu, d.invH);
                err = err < e1 ? e1 : err;
                P[v] = f4(p, Q.w);
                d.vel[v] = f4(ve, 0.f);
            }
            cv = block_max<BS>(err, red);
        } else {
            cv = 0.f;
        }
        k++;
        // ClothPBD.cpp:354-361; link, edge, triangle and self terms are 0 here
        const bool exitNow = !d.fixedIterations && cb < 0.005f && cv < 0.01f && 0.f < d.selfThreshold;
        if (exitNow || k >= d.iterations) break;
    }
    itersOut = k;
    cbOut = cb;
    cvOut = cv;
}

template <int BS>
__global__ void __launch_bounds__(BS) k_batch_substep(const ClothDesc* __restrict__ descs, int stageVerts,
                                                      unsigned long long* __restrict__ wgClock)
{
    extern __shared__ float4 sv[];
    __shared__ float red[BS / 64];
    // with wgClock: the workgroup's start and end on the constant-rate clock
    // and the unit it ran on (XCC id << 32 | HW_ID), three words per cloth
    const unsigned long long t0 = wgClock ? wall_clock64() : 0ull;
    const ClothDesc d = descs[blockIdx.x];
    // prediction and detection (ClothPBD.cpp:258-266, ClothCollision.cpp)
    for (int i = threadIdx.x; i < d.nv; i += BS) {
        const float4 P = d.pos[i];
        const f3 pv = xyz(P);
        d.prev[i] = make_float4(P.x, P.y, P.z, 0.f);
        f3 p = pv;
        if (P.w != 0.f) {
            f3 v = add(xyz(d.vel[i]), d.gdv);
            v = f3{v.x * d.drag, v.y * d.drag, v.z * d.drag};
            p = add(p, mul(d.h, v));
            d.vel[i] = f4(v, 0.f);
            d.pos[i] = f4(p, P.w);
        }
        if (d.anyContacts) detect_vertex<false>(i, p, pv, *d.cs, d.cb, d.thickness, d.tolerance, d.collFlags, 0.f);
    }
    __syncthreads();
    int iters;
    float cb, cv;
    const bool stage = d.nv <= stageVerts;
    if (stage) {
        for (int i = threadIdx.x; i < d.nv; i += BS) sv[i] = d.pos[i];
        __syncthreads();
        batch_iterations<BS>(d, sv, red, iters, cb, cv);
        for (int i = threadIdx.x; i < d.nv; i += BS) d.pos[i] = sv[i];
    } else {
        batch_iterations<BS>(d, d.pos, red, iters, cb, cv);
    }
    __syncthreads();
    // ClothPBD.cpp:314-321
    for (int i = threadIdx.x; i < d.nv; i += BS) {
        const float4 P = d.pos[i];
        if (P.w == 0.f) continue;
        d.vel[i] = f4(mul(d.invH, sub(xyz(P), xyz(d.prev[i]))), 0.f);
    }
    if (threadIdx.x == 0) {
        d.ctl->done = 1;
        d.ctl->k = iters;
        d.ctl->iterations = iters;
        d.ctl->cb = __float_as_uint(cb);
        d.ctl->cv = __float_as_uint(cv);
        d.ctl->bendError = cb;
        d.ctl->contactError = cv;
    }
    if (threadIdx.x < kErrParts) {
        d.ctl->errPart[0][threadIdx.x * kErrStride] = 0;
        d.ctl->errPart[1][threadIdx.x * kErrStride] = 0;
    }
    if (wgClock) {
        __syncthreads();
        if (threadIdx.x == 0) {
            const unsigned hw = __builtin_amdgcn_s_getreg((4) | (0 << 6) | (31 << 11));     // HW_ID
            const unsigned xcc = __builtin_amdgcn_s_getreg((20) | (0 << 6) | (15 << 11));   // XCC_ID
            wgClock[3 * (size_t)blockIdx.x] = t0;
            wgClock[3 * (size_t)blockIdx.x + 1] = wall_clock64();
            wgClock[3 * (size_t)blockIdx.x + 2] = ((unsigned long long)xcc << 32) | hw;
        }
    }
}

void batch_allow_lds(int bytes)
{
    HIP_OK(hipFuncSetAttribute((const void*)k_batch_substep<1024>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    HIP_OK(hipFuncSetAttribute((const void*)k_batch_substep<512>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    HIP_OK(hipFuncSetAttribute((const void*)k_batch_substep<256>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
}

void launch_batch_substep(hipStream_t s, const ClothDesc* descs, int numCloths, int blockSize, int stageVerts,
                          unsigned long long* wgClock)
{
    if (numCloths <= 0) return;
    const size_t lds = (size_t)(stageVerts > 0 ? stageVerts : 0) * sizeof(float4);
    if (blockSize == 1024)
        hipLaunchKernelGGL(k_batch_substep<1024>, dim3(numCloths), dim3(1024), lds, s, descs, stageVerts, wgClock);
    else if (blockSize == 512)
        hipLaunchKernelGGL(k_batch_substep<512>, dim3(numCloths), dim3(512), lds, s, descs, stageVerts, wgClock);
    else
        hipLaunchKernelGGL(k_batch_substep<256>, dim3(numCloths), dim3(256), lds, s, descs, stageVerts, wgClock);
}

static inline int blocks(int n) { return (n + kBlock - 1) / kBlock; }

void launch_predict(hipStream_t s, float4* pos, float4* prev, float4* vel, int n, float h, f3 gdv, float drag,
                    SolverCtl* ctl)
{
    int nb = blocks(n > 0 ? n : 1);
    hipLaunchKernelGGL(k_predict, dim3(nb), dim3(kBlock), 0, s, pos, prev, vel, n, h, gdv, drag, ctl);
}

void launch_detect(hipStream_t s, const float4* pos, const float4* prev, int n, const CollSet& cs, ContactBuf cb,
                   float thickness, float tolerance, int collFlags, const SolverCtl* ctl)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_detect, dim3(blocks(n)), dim3(kBlock), 0, s, pos, prev, n, cs, cb, thickness, tolerance,
                       collFlags, ctl);
}

void launch_fem(hipStream_t s, float4* pos, TriStreams tr, int start, int count, float h2, float epsUV,
                const SolverCtl* ctl)
{
    if (count <= 0) return;
    hipLaunchKernelGGL(k_fem, dim3(blocks(count)), dim3(kBlock), 0, s, pos, tr, start, count, h2, epsUV, ctl);
}

void launch_fem_slab(hipStream_t s, float4* pos, TriStreams tr, const ushort4* lidx, SlabSet ss, int numSlabs,
                     int maxVerts, float h2, float epsUV, SolverCtl* ctl, const SlabContacts& kc, bool tiles)
{
    if (numSlabs <= 0) return;
    if (tiles)
        hipLaunchKernelGGL(k_fem_slab<true>, dim3(numSlabs), dim3(kSlabBlock), (maxVerts + 1) * sizeof(float4), s, pos,
                           tr, lidx, ss, h2, epsUV, ctl, kc);
    else
        hipLaunchKernelGGL(k_fem_slab<false>, dim3(numSlabs), dim3(kSlabBlock), (maxVerts + 1) * sizeof(float4), s,
                           pos, tr, lidx, ss, h2, epsUV, ctl, kc);
}

void launch_predict_detect(hipStream_t s, float4* pos, float4* prev, float4* vel, int n, float h, f3 gdv, float drag,
                           SolverCtl* ctl, const CollSet& cs, ContactBuf cb, float thickness, float tolerance,
                           int collFlags)
{
    int nb = blocks(n > 0 ? n : 1);
    hipLaunchKernelGGL(k_predict_detect, dim3(nb), dim3(kBlock), 0, s, pos, prev, vel, n, h, gdv, drag, ctl, cs, cb,
                       thickness, tolerance, collFlags);
}

void launch_contacts_velocity(hipStream_t s, float4* pos, const float4* prev, float4* vel, ContactBuf cb,
                              float thickness, float mu, float invH, SolverCtl* ctl)
{
    hipLaunchKernelGGL(k_contacts_velocity, dim3(blocks(cb.nv > 0 ? cb.nv : 1)), dim3(kBlock), 0, s, pos, prev, vel,
                       cb, thickness, mu, invH, ctl);
}

void launch_bend_slab(hipStream_t s, float4* pos, const ushort4* lidx, SlabSet ss, int numSlabs, int maxVerts,
                      float stiff, SolverCtl* ctl, const SlabContacts& kc)
{
    if (numSlabs <= 0) return;
    hipLaunchKernelGGL(k_bend_slab, dim3(numSlabs), dim3(kSlabBlock), (maxVerts + 1) * sizeof(float4), s, pos, lidx, ss,
                       stiff, ctl, kc);
}

template <int NC, int KV>
static void fem_packed_kv(hipStream_t s, float4* pos, PackedFem pk, int numTiles, float h2, float epsUV,
                          SolverCtl* ctl, const SlabContacts& kc)
{
    if (kc.enabled == 2)
        hipLaunchKernelGGL((k_fem_packed<NC, KV, true>), dim3(numTiles), dim3(kSlabBlock), 0, s, pos, pk, h2, epsUV,
                           ctl, kc);
    else
        hipLaunchKernelGGL((k_fem_packed<NC, KV, false>), dim3(numTiles), dim3(kSlabBlock), 0, s, pos, pk, h2, epsUV,
                           ctl, kc);
}

void launch_fem_packed(hipStream_t s, float4* pos, PackedFem pk, int numTiles, int nc, int kv, float h2, float epsUV,
                       SolverCtl* ctl, const SlabContacts& kc)
{
    if (numTiles <= 0) return;
    if (nc == 8 && kv == 2) fem_packed_kv<8, 2>(s, pos, pk, numTiles, h2, epsUV, ctl, kc);
    else if (nc == 8 && kv == 4) fem_packed_kv<8, 4>(s, pos, pk, numTiles, h2, epsUV, ctl, kc);
    else if (nc == 12 && kv == 2) fem_packed_kv<12, 2>(s, pos, pk, numTiles, h2, epsUV, ctl, kc);
    else if (nc == 12 && kv == 4) fem_packed_kv<12, 4>(s, pos, pk, numTiles, h2, epsUV, ctl, kc);
    else throw std::runtime_error("packed FEM tiles: unsupported colour count or vertex slots");
}

template <int KV, bool kPair>
static void bend_packed_kv(hipStream_t s, float4* pos, PackedBend pk, int numSlabs, float stiff, SolverCtl* ctl,
                           const SlabContacts& kc)
{
    if (kc.enabled == 3)
        hipLaunchKernelGGL((k_bend_packed<KV, true, kPair>), dim3(numSlabs), dim3(kSlabBlock), 0, s, pos, pk, stiff,
                           ctl, kc);
    else
        hipLaunchKernelGGL((k_bend_packed<KV, false, kPair>), dim3(numSlabs), dim3(kSlabBlock), 0, s, pos, pk, stiff,
                           ctl, kc);
}

// Wider slabs (kv 5, 6) without vertex contacts.
template <int KV, bool kPair>
static void bend_packed_nc(hipStream_t s, float4* pos, PackedBend pk, int numSlabs, float stiff, SolverCtl* ctl,
                           const SlabContacts& kc)
{
    hipLaunchKernelGGL((k_bend_packed<KV, false, kPair>), dim3(numSlabs), dim3(kSlabBlock), 0, s, pos, pk, stiff, ctl,
                       kc);
}

void launch_bend_packed(hipStream_t s, float4* pos, PackedBend pk, int numSlabs, int kv, float stiff, SolverCtl* ctl,
                        const SlabContacts& kc)
{
    if (numSlabs <= 0) return;
    if (pk.nc < 1 || pk.nc > kPackMaxBendColours || pk.width < 1 || pk.width > kSlabBlock)
        throw std::runtime_error("packed bend slabs: unsupported colour count or width");
    // a pair of lanes per bend when a colour fits half the workgroup
    const bool pair = pk.width <= kSlabBlock / 2 && pk.width % 32 == 0;
    if (kv == 2) pair ? bend_packed_kv<2, true>(s, pos, pk, numSlabs, stiff, ctl, kc)
                      : bend_packed_kv<2, false>(s, pos, pk, numSlabs, stiff, ctl, kc);
    else if (kv == 3) pair ? bend_packed_kv<3, true>(s, pos, pk, numSlabs, stiff, ctl, kc)
                           : bend_packed_kv<3, false>(s, pos, pk, numSlabs, stiff, ctl, kc);
    else if (kv == 4) pair ? bend_packed_kv<4, true>(s, pos, pk, numSlabs, stiff, ctl, kc)
                           : bend_packed_kv<4, false>(s, pos, pk, numSlabs, stiff, ctl, kc);
    else if (kc.enabled == 3) throw std::runtime_error("packed bend slabs: contacts need at most 4 vertex slots");
    else if (kv == 5) pair ? bend_packed_nc<5, true>(s, pos, pk, numSlabs, stiff, ctl, kc)
                           : bend_packed_nc<5, false>(s, pos, pk, numSlabs, stiff, ctl, kc);
    else if (kv == 6) pair ? bend_packed_nc<6, true>(s, pos, pk, numSlabs, stiff, ctl, kc)
                           : bend_packed_nc<6, false>(s, pos, pk, numSlabs, stiff, ctl, kc);
    else throw std::runtime_error("packed bend slabs: unsupported vertex slots");
}

void launch_bend(hipStream_t s, float4* pos, const int4* bidx, int start, int count, float stiff, SolverCtl* ctl)
{
    if (count <= 0) return;
    hipLaunchKernelGGL(k_bend, dim3(blocks(count)), dim3(kBlock), 0, s, pos, bidx, start, count, stiff, ctl);
}

void launch_contacts(hipStream_t s, float4* pos, const float4* prev, float4* vel, ContactBuf cb, float thickness,
                     float mu, float invH, SolverCtl* ctl)
{
    if (cb.nv <= 0) return;
    hipLaunchKernelGGL(k_contacts, dim3(blocks(cb.nv)), dim3(kBlock), 0, s, pos, prev, vel, cb, thickness, mu, invH,
                       ctl);
}

void launch_iter_end(hipStream_t s, SolverCtl* ctl, int maxIter, int fixedIter, int bendFreq, int dihedral,
                     float selfThreshold)
{
    hipLaunchKernelGGL(k_iter_end, dim3(1), dim3(kErrParts), 0, s, ctl, maxIter, fixedIter, bendFreq, dihedral,
                       selfThreshold);
}

void launch_links(hipStream_t s, float4* pos, const int2* lidx, const float2* lrest, int start, int count, float k,
                  SolverCtl* ctl)
{
    if (count <= 0) return;
    hipLaunchKernelGGL(k_links, dim3(blocks(count)), dim3(kBlock), 0, s, pos, lidx, lrest, start, count, k, ctl);
}

void launch_link_slab(hipStream_t s, float4* pos, const ushort4* lidx, const float2* lrest, SlabSet ss, int numSlabs,
                      int maxVerts, float k, SolverCtl* ctl)
{
    if (numSlabs <= 0) return;
    hipLaunchKernelGGL(k_link_slab, dim3(numSlabs), dim3(kSlabBlock), (maxVerts + 1) * sizeof(float4), s, pos, lidx, lrest,
                       ss, k, ctl);
}

void launch_velocity(hipStream_t s, float4* pos, const float4* prev, float4* vel, int n, float invH, SolverCtl* ctl)
{
    hipLaunchKernelGGL(k_velocity, dim3(blocks(n > 0 ? n : 1)), dim3(kBlock), 0, s, pos, prev, vel, n, invH, ctl);
}

}  // namespace sc
