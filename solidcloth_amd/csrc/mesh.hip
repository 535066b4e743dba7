// gfx950 kernels for triangle-mesh obstacles: prediction with the travel
// reduction the mesh query radius needs, detection of obstacle-edge and
// obstacle-vertex contacts, and their Gauss-Seidel passes. Arithmetic follows
// the reference's operation order (see kernels.hip).
#include "bvh_query.h"
#include "collide_dev.h"
#include "ctl_dev.h"
#include "devutil.h"
#include "kernels.h"

namespace sc {

static constexpr int kBlock = 256;
static constexpr int kMaxSolveBlocks = 2048;

// Physics/ClothPBD.cpp:258-266 plus max_i |pos_i - prev_i|
// (ClothCollision.cpp:95-104). Thread 0 resets the control block, the first
// threads clear the contact groups of the substep.
__global__ void __launch_bounds__(kBlock) k_predict_mesh(float4* __restrict__ pos, float4* __restrict__ prev,
                                                         float4* __restrict__ vel, int n, float h, f3 gdv, float drag,
                                                         SolverCtl* ctl, FeatureContactBuf ec, FeatureContactBuf tc)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        begin_substep_ctl(ctl);
    }
    if (i < kErrParts) begin_substep_parts(ctl, i);
    if (i < ec.groups) ec.groupCount[i * kGroupStride] = 0;
    if (i < tc.groups) tc.groupCount[i * kGroupStride] = 0;
    float tv = 0.f;
    if (i < n) {
        const float4 P = pos[i];
        prev[i] = make_float4(P.x, P.y, P.z, 0.f);
        if (P.w != 0.f) {
            f3 v = add(xyz(vel[i]), gdv);
            v = f3{v.x * drag, v.y * drag, v.z * drag};
            const f3 p = add(xyz(P), mul(h, v));
            vel[i] = f4(v, 0.f);
            pos[i] = f4(p, P.w);
            tv = len(sub(p, xyz(P)));
            tv = tv > 0.f ? tv : 0.f;  // a NaN never raises the reference's maximum
        }
    }
    wave_max_atomic(&ctl->maxTv, tv);
}

// Lists cloth edge / triangle id in its colour group. The calling lanes of a
// wave share one atomic per group among them: a group's lanes take
// consecutive places after the leader's reservation.
__device__ __forceinline__ void group_append(const FeatureContactBuf& b, int id)
{
    const int g = b.group[id];
    const int lane = __lane_id();
    unsigned long long todo = __ballot(true);
    while (todo) {
        const int leader = __ffsll(todo) - 1;
        const int gl = __builtin_amdgcn_readlane(g, leader);
        const unsigned long long same = __ballot(g == gl) & todo;
        int base = 0;
        if (lane == leader) base = atomicAdd(&b.groupCount[gl * kGroupStride], __popcll(same));
        base = __builtin_amdgcn_readlane(base, leader);
        if (g == gl) b.list[b.groupBase[g] + base + __popcll(same & ((1ull << lane) - 1ull))] = id;
        todo &= ~same;
    }
}

__device__ __forceinline__ void put_feature(const FeatureContactBuf& b, int id, int& cnt, f3 point, f3 normal,
                                            float w1, float w2, float w3, int feature, int ci)
{
    const int k = cnt * b.n + id;
    b.P[k] = f4(point, 0.f);
    b.N[k] = f4(normal, 0.f);
    b.W[k] = make_float4(w1, w2, w3, 0.f);
    b.F[k] = make_int2(feature, ci);
    cnt++;
}

// Edge and triangle contacts of the cloth against every collidable, in world
// order, from the previous positions (Physics/ClothCollision.cpp:50-87):
//  * mesh obstacles (ClothCollision.cpp:137-197): per cloth edge the nearest
//    obstacle edge within thickness + max(tolerance, 1.5 max travel), per
//    cloth triangle the nearest obstacle vertex within thickness +
//    max(tolerance, 1.2 x its vertices' travel);
//  * spheres and grid SDFs (ClothCollision.cpp:472-579): the collidable's
//    segment / triangle query within thickness + max(tolerance, 1.1 x
//    travel), border coordinates rejected unless vertex (for edges) or edge
//    (for triangles) contacts are off. The grid SDF's segment query is a
//    stub that the border test always rejects (the host requires vertex
//    contacts with it).
// Blocks [0, edgeBlocks) take edges, the rest triangles.
__global__ void __launch_bounds__(kBlock) k_detect_features(const float4* __restrict__ pos,
                                                            const float4* __restrict__ prev,
                                                            const int2* __restrict__ cEdges,
                                                            const int4* __restrict__ cTris, CollSet cs,
                                                            FeatureContactBuf ec, FeatureContactBuf tc,
                                                            int edgeBlocks, float thickness, float tolerance,
                                                            int collFlags, const SolverCtl* __restrict__ ctl)
{
    const bool borderE = (collFlags & CF_VERTICES) == 0;
    const bool borderT = borderE || (collFlags & CF_EDGES) == 0;
    if ((int)blockIdx.x < edgeBlocks) {
        const int e = blockIdx.x * blockDim.x + threadIdx.x;
        if (e >= ec.n) return;
        const int2 ev = cEdges[e];
        const float4 P1 = prev[ev.x], P2 = prev[ev.y];
        const f3 p = xyz(P1), q = xyz(P2);
        const float meshRad = thickness + smax(tolerance, __uint_as_float(ctl->maxTv) * 1.5f);
        const float tv = smax(len(sub(p, xyz(pos[ev.x]))), len(sub(q, xyz(pos[ev.y]))));
        const float sdfRad = thickness + smax(tolerance, tv * 1.1f);
        const float eps = 0.02f;
        int cnt = 0;
        for (int ci = 0; ci < cs.n; ci++) {
            const CollDev& c = cs.c[ci];
            if (c.type == CT_MESH) {
                NearestResult r;
                if (nearest_edge(*c.mesh, p, q, meshRad, r))
                    put_feature(ec, e, cnt, r.point, r.normal, 1.0f - r.param, r.param, 0.f, r.feature, ci);
            } else if (c.type == CT_SPHERE || c.type == CT_CAPSULE) {
                f3 cp, n;
                float w1, w2;
                const float dist = c.type == CT_SPHERE ? sphere_segment(c, p, q, cp, n, w1, w2)
                                                       : capsule_segment(c, p, q, cp, n, w1, w2);
                if (dist < sdfRad && dist >= 0.f &&
                    ((w1 > eps && w1 < 1.f - eps && w2 > eps && w2 < 1.f - eps) || borderE))
                    put_feature(ec, e, cnt, cp, n, w1, w2, 0.f, -1, ci);
            }
        }
        ec.count[e] = cnt;
        if (cnt) group_append(ec, e);
    } else {
        const int t = (blockIdx.x - edgeBlocks) * blockDim.x + threadIdx.x;
        if (t >= tc.n) return;
        const int4 ix = cTris[t];
        const f3 a = xyz(prev[ix.x]), b = xyz(prev[ix.y]), cc = xyz(prev[ix.z]);
        const float tv1 = len(sub(xyz(pos[ix.x]), a));
        const float tv2 = len(sub(xyz(pos[ix.y]), b));
        const float tv3 = len(sub(xyz(pos[ix.z]), cc));
        const float tv = smax(tv1, smax(tv2, tv3));
        const float meshRad = thickness + smax(tolerance, tv * 1.2f);
        const float sdfRad = thickness + smax(tolerance, tv * 1.1f);
        const float eps = 0.001f;
        int cnt = 0;
        for (int ci = 0; ci < cs.n; ci++) {
            const CollDev& c = cs.c[ci];
            if (c.type == CT_MESH) {
                NearestResult r;
                if (nearest_vertex(*c.mesh, a, b, cc, meshRad, r))
                    put_feature(tc, t, cnt, r.point, r.normal, r.bary.x, r.bary.y, r.bary.z, r.feature, ci);
            } else if (c.type == CT_SPHERE || c.type == CT_SDF || c.type == CT_CAPSULE) {
                f3 cp, n, bar;
                const float dist = c.type == CT_SPHERE    ? sphere_triangle(c, a, b, cc, cp, n, bar)
                                   : c.type == CT_CAPSULE ? capsule_triangle(c, a, b, cc, cp, n, bar)
                                                          : sdf_triangle(c, a, b, cc, cp, n, bar);
                if (dist < sdfRad && dist >= 0.f && ((bar.x > eps && bar.y > eps && bar.z > eps) || borderT))
                    put_feature(tc, t, cnt, cp, n, bar.x, bar.y, bar.z, -1, ci);
            }
        }
        tc.count[t] = cnt;
        if (cnt) group_append(tc, t);
    }
}

// Friction and position update shared by edge and triangle contacts
// (ClothCollisionHandler.cpp:236-286, 318-361). Returns the displacement
// factors; the caller applies them per vertex in the reference's order.
struct FeatureStep {
    f3 pf;      // friction displacement (zero when not applied)
    bool fric;  // friction applied
    f3 disp;
};

__device__ __forceinline__ FeatureStep feature_step(f3 normal, float s, float depth, f3 p, f3 p0, float mu,
                                                    float& lambda, float& lambdaF)
{
    FeatureStep st;
    float dLambda = s * depth;
    lambda += dLambda;
    st.disp = mul(dLambda, normal);
    st.fric = false;
    st.pf = f3{0.f, 0.f, 0.f};
    if (mu != 0.f) {
        const f3 v12 = sub(p, p0);
        const float vnrel = dot(normal, v12);
        f3 vt = sub(v12, mul(vnrel, normal));
        const float vtrel = len(vt);
        const float limit = mu * lambda;
        if (vtrel > 0.001f) {
            const float dLF = vtrel;
            const float l0 = lambdaF;
            lambdaF = l0 + dLF;
            if (lambdaF >= limit) lambdaF = limit;
            dLambda = lambdaF - l0;
            vt = mul(1.f / vtrel, vt);
            st.pf = mul(-s * dLambda, vt);
            st.fric = true;
        }
    }
    return st;
}

// Physics/ClothCollisionHandler.cpp:289-362 — obstacle-edge contacts of one
// colour group (one launch per group, in group order). The cloth edges of a
// group share no vertex, so their contacts are solved in parallel across the
// grid; the group's contact count is read on the device and the grid, sized
// for the group's capacity, strides over it.
__global__ void __launch_bounds__(kBlock) k_edge_contacts(float4* __restrict__ pos, const float4* __restrict__ prev,
                                                          const int2* __restrict__ cEdges, FeatureContactBuf ec,
                                                          const MeshGeo* __restrict__ geo, float len0, float mu, int g,
                                                          SolverCtl* ctl)
{
    if (ctl->done) return;
    const int cnt = ec.groupCount[g * kGroupStride], base = ec.groupBase[g];
    if ((int)(blockIdx.x * blockDim.x) >= cnt) return;
    float err = 0.f;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < cnt; k += gridDim.x * blockDim.x) {
        const int e = ec.list[base + k];
        const int2 ev = cEdges[e];
        const float4 X1 = pos[ev.x], X2 = pos[ev.y];
        f3 x1 = xyz(X1), x2 = xyz(X2);
        const float m1 = X1.w, m2 = X2.w;
        const int ns = ec.count[e];
        for (int sl = 0; sl < ns; sl++) {
            const int q = sl * ec.n + e;
            const float4 CP = ec.P[q], CN = ec.N[q];
            const int2 F = ec.F[q];
            f3 point = xyz(CP), p;
            float w1, w2;
            if (F.x >= 0) {
                // mesh contact: closest points on the same obstacle edge again
                const MeshGeo& M = geo[F.y];
                const int2 me = M.edge[F.x];
                float s1, t;
                closest_pt_segm_segm(xyz(M.v[me.x]), xyz(M.v[me.y]), x1, x2, s1, t, point, p);
                w1 = 1.0f - t;
                w2 = t;
            } else {
                const float4 W = ec.W[q];
                w1 = W.x;
                w2 = W.y;
                p = add(mul(w1, x1), mul(w2, x2));
            }
            const f3 normal = xyz(CN);
            float lambda = CP.w, lambdaF = CN.w;
            const float l = dot(normal, sub(p, point));
            if (!(l > len0)) {
                const float s = 1.f / (w1 * w1 / m1 + w2 * w2 / m2);
                const float depth = len0 - l;
                if (depth > err) err = depth;
                const f3 p0 = add(mul(w1, xyz(prev[ev.x])), mul(w2, xyz(prev[ev.y])));
                const FeatureStep st = feature_step(normal, s, depth, p, p0, mu, lambda, lambdaF);
                if (st.fric) {
                    x1 = add(x1, mul(w1 * m1, st.pf));
                    x2 = add(x2, mul(w2 * m2, st.pf));
                }
                x1 = add(x1, mul(w1 * m1, st.disp));
                x2 = add(x2, mul(w2 * m2, st.disp));
            }
            ec.P[q] = f4(point, lambda);
            ec.N[q] = make_float4(CN.x, CN.y, CN.z, lambdaF);
            ec.W[q] = make_float4(w1, w2, 0.f, 0.f);
        }
        pos[ev.x] = f4(x1, m1);
        pos[ev.y] = f4(x2, m2);
    }
    wave_max_atomic(&ctl->ce, err);
}

// Physics/ClothCollisionHandler.cpp:204-287 — obstacle-vertex contacts on
// the cloth triangles of one colour group, launched per group as above.
__global__ void __launch_bounds__(kBlock) k_tri_contacts(float4* __restrict__ pos, const float4* __restrict__ prev,
                                                         const int4* __restrict__ cTris, FeatureContactBuf tc,
                                                         const MeshGeo* __restrict__ geo, float len0, float mu, int g,
                                                         SolverCtl* ctl)
{
    if (ctl->done) return;
    const int cnt = tc.groupCount[g * kGroupStride], base = tc.groupBase[g];
    if ((int)(blockIdx.x * blockDim.x) >= cnt) return;
    float err = 0.f;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < cnt; k += gridDim.x * blockDim.x) {
        const int t = tc.list[base + k];
        const int4 ix = cTris[t];
        const float4 X1 = pos[ix.x], X2 = pos[ix.y], X3 = pos[ix.z];
        f3 x1 = xyz(X1), x2 = xyz(X2), x3 = xyz(X3);
        const float m1 = X1.w, m2 = X2.w, m3 = X3.w;
        const int ns = tc.count[t];
        for (int sl = 0; sl < ns; sl++) {
            const int q = sl * tc.n + t;
            const float4 CP = tc.P[q], CN = tc.N[q];
            const int2 F = tc.F[q];
            const f3 point = xyz(CP);
            f3 p, co;
            closest_pt_point_triangle(point, x1, x2, x3, p, co);
            const float w1 = co.x, w2 = co.y, w3 = co.z;
            f3 n = normalized(neg(cross(sub(x2, x1), sub(x3, x1))));
            f3 normal = n;
            if (F.x >= 0) {
                const f3 mn = xyz(geo[F.y].vn[F.x]);
                if (dot(n, mn) < 0.f) normal = neg(n);
            }
            float lambda = CP.w, lambdaF = CN.w;
            const float l = dot(normal, sub(p, point));
            if (!(l > len0)) {
                const float s = 1.f / (w1 * w1 / m1 + w2 * w2 / m2 + w3 * w3 / m3);
                const float depth = len0 - l;
                if (depth > err) err = depth;
                const f3 p0 =
                    add(add(mul(w1, xyz(prev[ix.x])), mul(w2, xyz(prev[ix.y]))), mul(w3, xyz(prev[ix.z])));
                const FeatureStep st = feature_step(normal, s, depth, p, p0, mu, lambda, lambdaF);
                if (st.fric) {
                    x1 = add(x1, mul(w1 * m1, st.pf));
                    x2 = add(x2, mul(w2 * m2, st.pf));
                    x3 = add(x3, mul(w3 * m3, st.pf));
                }
                x1 = add(x1, mul(w1 * m1, st.disp));
                x2 = add(x2, mul(w2 * m2, st.disp));
                x3 = add(x3, mul(w3 * m3, st.disp));
            }
            tc.P[q] = make_float4(CP.x, CP.y, CP.z, lambda);
            tc.N[q] = f4(normal, lambdaF);
            tc.W[q] = make_float4(w1, w2, w3, 0.f);
        }
        pos[ix.x] = f4(x1, m1);
        pos[ix.y] = f4(x2, m2);
        pos[ix.z] = f4(x3, m3);
    }
    wave_max_atomic(&ctl->ct, err);
}

// SDF::CreateFromMesh (Geometry/SDF.cpp:58-76), one thread per node: node
// (x, y, z) at (x, y, z) * steps + min (SDF.h:64-70); its value is the
// distance ClosestPointOnMeshToPointAcc returns, the smallest FP32 distance
// to a triangle with distance 0 skipped (FLT_MAX if none), here from the
// obstacle's triangle LBVH. That query's normal is the unit delta, so the
// reference's sign test never flips a value: the field is unsigned. Nodes
// x-fastest.
__global__ void __launch_bounds__(kBlock) k_sdf_from_mesh(MeshGeo g, int vx, int vy, int count, f3 mn, f3 steps,
                                                          float* __restrict__ out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const int x = k % vx, w = k / vx, y = w % vy, z = w / vy;
    const f3 org{(float)x * steps.x + mn.x, (float)y * steps.y + mn.y, (float)z * steps.z + mn.z};
    constexpr float kFltMax = 3.402823466e+38f;
    NearestResult r;
    out[k] = nearest_triangle(g, org, kFltMax, r) ? r.dist : kFltMax;
}

static inline int blocks(int n) { return (n + kBlock - 1) / kBlock; }

void launch_sdf_from_mesh(hipStream_t s, const MeshGeo& g, int nx, int ny, int nz, f3 mn, f3 steps, float* out)
{
    const int count = (nx + 1) * (ny + 1) * (nz + 1);
    if (count <= 0) return;
    hipLaunchKernelGGL(k_sdf_from_mesh, dim3(blocks(count)), dim3(kBlock), 0, s, g, nx + 1, ny + 1, count, mn, steps,
                       out);
}

void launch_predict_mesh(hipStream_t s, float4* pos, float4* prev, float4* vel, int n, float h, f3 gdv, float drag,
                         SolverCtl* ctl, FeatureContactBuf ec, FeatureContactBuf tc)
{
    int m = n > ec.groups ? n : ec.groups;
    m = m > tc.groups ? m : tc.groups;
    hipLaunchKernelGGL(k_predict_mesh, dim3(blocks(m > 0 ? m : 1)), dim3(kBlock), 0, s, pos, prev, vel, n, h, gdv,
                       drag, ctl, ec, tc);
}

void launch_detect_features(hipStream_t s, const float4* pos, const float4* prev, const int2* cEdges,
                            const int4* cTris, const CollSet& cs, FeatureContactBuf ec, FeatureContactBuf tc,
                            float thickness, float tolerance, int collFlags, const SolverCtl* ctl)
{
    const int eb = (collFlags & CF_EDGES) && ec.slots > 0 ? blocks(ec.n) : 0;
    const int tb = (collFlags & CF_TRIANGLES) && tc.slots > 0 ? blocks(tc.n) : 0;
    if (eb + tb == 0) return;
    hipLaunchKernelGGL(k_detect_features, dim3(eb + tb), dim3(kBlock), 0, s, pos, prev, cEdges, cTris, cs, ec, tc,
                       eb, thickness, tolerance, collFlags, ctl);
}

// grid for a group of `cap` cloth edges / triangles: one thread per possible
// contact, up to kMaxSolveBlocks workgroups (the kernels stride beyond)
static inline int group_blocks(int cap)
{
    const int b = blocks(cap);
    return b < kMaxSolveBlocks ? b : kMaxSolveBlocks;
}

void launch_edge_contacts(hipStream_t s, float4* pos, const float4* prev, const int2* cEdges, FeatureContactBuf ec,
                          const int* groupBase, const MeshGeo* geo, float thickness, float mu, SolverCtl* ctl)
{
    for (int g = 0; g < ec.groups; g++) {
        const int cap = groupBase[g + 1] - groupBase[g];
        if (cap > 0)
            hipLaunchKernelGGL(k_edge_contacts, dim3(group_blocks(cap)), dim3(kBlock), 0, s, pos, prev, cEdges, ec,
                               geo, thickness, mu, g, ctl);
    }
}

void launch_tri_contacts(hipStream_t s, float4* pos, const float4* prev, const int4* cTris, FeatureContactBuf tc,
                         const int* groupBase, const MeshGeo* geo, float thickness, float mu, SolverCtl* ctl)
{
    for (int g = 0; g < tc.groups; g++) {
        const int cap = groupBase[g + 1] - groupBase[g];
        if (cap > 0)
            hipLaunchKernelGGL(k_tri_contacts, dim3(group_blocks(cap)), dim3(kBlock), 0, s, pos, prev, cTris, tc, geo,
                               thickness, mu, g, ctl);
    }
}

}  // namespace sc
