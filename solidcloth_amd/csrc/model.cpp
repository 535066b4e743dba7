// Host implementation: cloth construction, constraint colouring, device
// buffers, substep recording and hipGraph replay.
#include "model.h"

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "hiputil.h"

namespace sc {

// ---------------------------------------------------------------- SDF ----

SDF::~SDF() { dev_free(d_nodes); }

// Geometry/SDF.cpp:10-40
void SDF::create(int nx, int ny, int nz, const scAabb3& b)
{
    // the device lookup indexes nodes with 32-bit ints
    if (nx < 1 || ny < 1 || nz < 1 || (double)(nx + 1.0) * (ny + 1.0) * (nz + 1.0) > (double)INT_MAX)
        throw std::runtime_error("SDF grid out of range: cell counts must be >= 1 and the node count at most 2^31 - 1");
    nc[0] = nx; nc[1] = ny; nc[2] = nz;
    nv[0] = nx + 1; nv[1] = ny + 1; nv[2] = nz + 1;
    box = b;
    vals.assign((size_t)nv[0] * nv[1] * nv[2], 0.f);
    f3 ext = sub(f3{b.max.x, b.max.y, b.max.z}, f3{b.min.x, b.min.y, b.min.z});
    f3 inv{1.f / (float)nx, 1.f / (float)ny, 1.f / (float)nz};
    steps = f3{ext.x * inv.x, ext.y * inv.y, ext.z * inv.z};
}

// Geometry/SDF.cpp:210-256: forward differences, the last layer of each axis
// copies its neighbour.
void SDF::compute_gradient()
{
    grads.assign(vals.size(), f3{0.f, 0.f, 0.f});
    auto oob = [&](int x, int y, int z) { return x > nc[0] || y > nc[1] || z > nc[2] || x < 0 || y < 0 || z < 0; };
    auto idx = [&](int x, int y, int z) { return (size_t)x + (size_t)y * nv[0] + (size_t)z * nv[0] * nv[1]; };
    auto val = [&](int x, int y, int z) { return oob(x, y, z) ? 0.f : vals[idx(x, y, z)]; };
    auto grd = [&](int x, int y, int z) { return oob(x, y, z) ? f3{0.f, 0.f, 0.f} : grads[idx(x, y, z)]; };
    auto set = [&](int x, int y, int z, f3 g) {
        if (!oob(x, y, z)) grads[idx(x, y, z)] = g;
    };
    for (int x = 0; x < nc[0]; x++) {
        for (int y = 0; y < nc[1]; y++) {
            for (int z = 0; z < nc[2]; z++) {
                float v = val(x, y, z);
                float dfdx = (val(x + 1, y, z) - v) / steps.x;
                float dfdy = (val(x, y + 1, z) - v) / steps.y;
                float dfdz = (val(x, y, z + 1) - v) / steps.z;
                set(x, y, z, f3{dfdx, dfdy, dfdz});
            }
            set(x, y, nc[2], grd(x, y, nc[2] - 1));
        }
        for (int z = 0; z <= nc[2]; z++) set(x, nc[1], z, grd(x, nc[1] - 1, z));
    }
    for (int y = 0; y <= nc[1]; y++)
        for (int z = 0; z <= nc[2]; z++) set(nc[0], y, z, grd(nc[0] - 1, y, z));
}

void SDF::upload()
{
    std::vector<float4> nodes(vals.size());
    for (size_t i = 0; i < vals.size(); i++) nodes[i] = make_float4(grads[i].x, grads[i].y, grads[i].z, vals[i]);
    dev_alloc(d_nodes, nodes.size());
    HIP_OK(hipMemcpy(d_nodes, nodes.data(), nodes.size() * sizeof(float4), hipMemcpyHostToDevice));
}

// --------------------------------------------------------------- Mesh ----

void Mesh::clear()
{
    vertices.clear(); indices.clear(); normals.clear(); uvs.clear(); colors.clear(); edges.clear();
    triNormals.clear(); edgeNormals.clear();
    dev.reset();
}

// Geometry/Mesh.cpp:65-147. The duplicate list is sorted with std::sort by
// (i1, i2) from index order, exactly as the reference does, so which of an
// interior edge's triangles becomes t1 is the same as in the reference
// compiled with the same standard library.
void Mesh::construct_edges()
{
    edges.clear();
    const int ne = (int)indices.size();
    if (ne == 0) return;
    std::vector<MeshEdge> dup(ne);
    for (int i = 0; i < ne; i++) {
        uint32_t a = indices[i];
        uint32_t b = ((i + 1) % 3 == 0) ? indices[i - 2] : indices[i + 1];
        if (a > b) std::swap(a, b);
        dup[i].i1 = a; dup[i].i2 = b; dup[i].t1 = i / 3; dup[i].t2 = -1;
    }
    std::sort(dup.begin(), dup.end(), [](const MeshEdge& e1, const MeshEdge& e2) {
        return e1.i1 < e2.i1 || (e1.i1 == e2.i1 && e1.i2 < e2.i2);
    });
    MeshEdge cur = dup[0];
    for (size_t i = 1; i < dup.size(); i++) {
        if (dup[i].i1 == dup[i - 1].i1 && dup[i].i2 == dup[i - 1].i2) {
            cur.t2 = dup[i].t1;
            continue;
        }
        edges.push_back(cur);
        cur = dup[i];
    }
    edges.push_back(cur);
}

// Math/Quaternion.cpp:116-122,128-131,142-147: qRotate(q, v) = (q (v q*)).v
// with qMultiply(a, b) = (a.s b.s - a.v.b.v, a.s b.v + b.s a.v + a.v x b.v)
// and the product of a vector v with a quaternion taken as (0, v).
static f3 quat_rotate(const scQuaternion& q, f3 v)
{
    const f3 qv{q.v.x, q.v.y, q.v.z};
    const f3 iv = neg(qv);  // conjugate (q.s, -q.v)
    // t = (0, v) * (q.s, -q.v)
    const float ts = 0.f * q.s - dot(v, iv);
    const f3 tv = add(add(mul(0.f, iv), mul(q.s, v)), cross(v, iv));
    // r = q * t; only its vector part is kept
    return add(add(mul(q.s, tv), mul(ts, qv)), cross(qv, tv));
}

// Physics/ClothCollisionSDF.h:62-68 (CapsuleSDF): the axis ends
// centre +- qRotate(rot, (0, hh, 0)).
void capsule_axis(const Capsule& k, f3& P, f3& Q)
{
    const f3 half = quat_rotate(k.rot, f3{0.f, k.hh, 0.f});
    P = add(k.center, half);
    Q = sub(k.center, half);
}

// Geometry/Mesh.cpp:10-63 (angle-weighted vertex normals), for rendering.
void Mesh::compute_normals()
{
    normals.assign(vertices.size(), scVector3{0.f, 0.f, 0.f});
    triNormals.assign(indices.size() / 3, scVector3{0.f, 0.f, 0.f});
    std::vector<f3> acc(vertices.size(), f3{0.f, 0.f, 0.f});
    auto V = [&](unsigned i) { return f3{vertices[i].x, vertices[i].y, vertices[i].z}; };
    for (size_t i = 0; i + 2 < indices.size(); i += 3) {
        unsigned i1 = indices[i], i2 = indices[i + 1], i3 = indices[i + 2];
        f3 v1 = V(i1), v2 = V(i2), v3 = V(i3);
        f3 nt = normalized(cross(sub(v2, v1), sub(v3, v1)));
        triNormals[i / 3] = scVector3{nt.x, nt.y, nt.z};
        f3 e1 = normalized(sub(v2, v1)), e2 = normalized(sub(v3, v2)), e3 = normalized(sub(v1, v3));
        float c1 = -dot(e1, e3), c2 = -dot(e1, e2), c3 = -dot(e2, e3);
        float t1 = std::fabs(c1) > 1 ? std::fabs(std::atan2(len(cross(e1, neg(e3))), c1)) : std::acos(c1);
        float t2 = std::fabs(c2) > 1 ? std::fabs(std::atan2(len(cross(e2, neg(e1))), c2)) : std::acos(c2);
        float t3 = std::fabs(c3) > 1 ? std::fabs(std::atan2(len(cross(e3, neg(e1))), c3)) : std::acos(c3);
        acc[i1] = add(acc[i1], mul(t1, nt));
        acc[i2] = add(acc[i2], mul(t2, nt));
        acc[i3] = add(acc[i3], mul(t3, nt));
    }
    for (size_t i = 0; i < vertices.size(); i++) {
        if (len(acc[i]) == 0.f) continue;
        f3 n = normalized(acc[i]);
        normals[i] = scVector3{n.x, n.y, n.z};
    }
}

// Geometry/Mesh.cpp:105, 124-126: a border edge takes its triangle's normal,
// an interior edge the normalised sum of both (the last duplicate's
// triangle for a non-manifold edge, as construct_edges records it).
void Mesh::compute_edge_normals()
{
    edgeNormals.assign(edges.size(), scVector3{0.f, 0.f, 0.f});
    if (triNormals.size() != indices.size() / 3) compute_normals();
    auto T = [&](int t) { return f3{triNormals[t].x, triNormals[t].y, triNormals[t].z}; };
    for (size_t e = 0; e < edges.size(); e++) {
        f3 n = T(edges[e].t1);
        if (edges[e].t2 >= 0) n = normalized(add(T(edges[e].t1), T(edges[e].t2)));
        edgeNormals[e] = scVector3{n.x, n.y, n.z};
    }
}

MeshDevice::~MeshDevice()
{
    dev_free(v); dev_free(vn); dev_free(en); dev_free(tri); dev_free(edge);
    dev_free(triLo); dev_free(triHi); dev_free(vtxLo); dev_free(edgeLo); dev_free(edgeHi);
}

const MeshGeo& Mesh::device_geo(hipStream_t s)
{
    if (dev) return dev->geo;
    if (edgeNormals.size() != edges.size()) compute_edge_normals();
    std::unique_ptr<MeshDevice> d(new MeshDevice());
    const int nv = (int)vertices.size(), nt = (int)indices.size() / 3, ne = (int)edges.size();
    auto P = [&](unsigned i) { return make_float4(vertices[i].x, vertices[i].y, vertices[i].z, 0.f); };
    std::vector<float4> hv(nv), hvn(nv), hen(ne), tlo(nt), thi(nt), elo(ne), ehi(ne);
    std::vector<int4> ht(nt);
    std::vector<int2> he(ne);
    float scale = 0.f;
    for (int i = 0; i < nv; i++) {
        hv[i] = P(i);
        hvn[i] = make_float4(normals[i].x, normals[i].y, normals[i].z, 0.f);
        scale = std::max(scale, std::max(std::fabs(vertices[i].x), std::max(std::fabs(vertices[i].y),
                                                                            std::fabs(vertices[i].z))));
    }
    for (int t = 0; t < nt; t++) {
        const unsigned a = indices[3 * t], b = indices[3 * t + 1], c = indices[3 * t + 2];
        ht[t] = make_int4((int)a, (int)b, (int)c, 0);
        const float4 A = P(a), B = P(b), C = P(c);
        tlo[t] = make_float4(std::min(std::min(A.x, B.x), C.x), std::min(std::min(A.y, B.y), C.y),
                             std::min(std::min(A.z, B.z), C.z), 0.f);
        thi[t] = make_float4(std::max(std::max(A.x, B.x), C.x), std::max(std::max(A.y, B.y), C.y),
                             std::max(std::max(A.z, B.z), C.z), 0.f);
    }
    for (int e = 0; e < ne; e++) {
        he[e] = make_int2((int)edges[e].i1, (int)edges[e].i2);
        hen[e] = make_float4(edgeNormals[e].x, edgeNormals[e].y, edgeNormals[e].z, 0.f);
        const float4 A = P(edges[e].i1), B = P(edges[e].i2);
        elo[e] = make_float4(std::min(A.x, B.x), std::min(A.y, B.y), std::min(A.z, B.z), 0.f);
        ehi[e] = make_float4(std::max(A.x, B.x), std::max(A.y, B.y), std::max(A.z, B.z), 0.f);
    }
    auto up = [&](auto*& dst, const auto& src) {
        dev_alloc(dst, src.size());
        if (!src.empty())
            HIP_OK(hipMemcpy(dst, src.data(), src.size() * sizeof(src[0]), hipMemcpyHostToDevice));
    };
    up(d->v, hv); up(d->vn, hvn); up(d->en, hen); up(d->tri, ht); up(d->edge, he);
    up(d->triLo, tlo); up(d->triHi, thi); up(d->vtxLo, hv); up(d->edgeLo, elo); up(d->edgeHi, ehi);
    d->triB.build(s, d->triLo, d->triHi, nt);
    d->vtxB.build(s, d->vtxLo, d->vtxLo, nv);
    d->edgeB.build(s, d->edgeLo, d->edgeHi, ne);
    HIP_OK(hipStreamSynchronize(s));
    MeshGeo& g = d->geo;
    g.v = d->v; g.tri = d->tri; g.vn = d->vn; g.edge = d->edge; g.en = d->en;
    g.triBvh = d->triB.view(); g.vtxBvh = d->vtxB.view(); g.edgeBvh = d->edgeB.view();
    g.scale = scale;
    g.nv = nv; g.nt = nt; g.ne = ne;
    dev = std::move(d);
    return dev->geo;
}

// ------------------------------------------------------------- timing ----

hipEvent_t KernelTimer::get()
{
    if (!pool.empty()) {
        hipEvent_t e = pool.back();
        pool.pop_back();
        return e;
    }
    // timing-only events: no system-scope fence (cache write-back) at each
    // record, which would otherwise add microseconds to every bracketed launch
    hipEvent_t e;
    HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    return e;
}

void KernelTimer::harvest()
{
    for (auto& p : pending) {
        float ms = 0.f;
        HIP_OK(hipEventElapsedTime(&ms, p.second.first, p.second.second));
        launches[p.first]++;
        totalMs[p.first] += ms;
        pool.push_back(p.second.first);
        pool.push_back(p.second.second);
    }
    pending.clear();
}

void KernelTimer::reset()
{
    for (int i = 0; i < 5; i++) { launches[i] = 0; totalMs[i] = 0; }
}

KernelTimer::~KernelTimer()
{
    for (auto e : pool) (void)hipEventDestroy(e);
    for (auto& p : pending) { (void)hipEventDestroy(p.second.first); (void)hipEventDestroy(p.second.second); }
}

// ------------------------------------------------------------- solver ----

GpuSolver::~GpuSolver() { release(); }

void GpuSolver::release_graphs()
{
    if (graphExec) (void)hipGraphExecDestroy(graphExec);
    if (graph) (void)hipGraphDestroy(graph);
    if (graphExec2) (void)hipGraphExecDestroy(graphExec2);
    if (graph2) (void)hipGraphDestroy(graph2);
    graphExec = graphExec2 = nullptr;
    graph = graph2 = nullptr;
    graphKey.clear();
    graphKey2.clear();
}

void GpuSolver::release()
{
    release_graphs();
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    ev0 = ev1 = nullptr;
    dev_free(warm.xhat); dev_free(warm.corr); dev_free(warm.T); dev_free(warm.triStrain); dev_free(warm.X);
    dev_free(warm.adjOff); dev_free(warm.adj); dev_free(warm.vtOff); dev_free(warm.vt);
    dev_free(warm.snapPos); dev_free(warm.snapPrev); dev_free(warm.snapVel); dev_free(warm.snapCorr);
    dev_free(warm.altPos); dev_free(warm.altPrev); dev_free(warm.altVel); dev_free(warm.altCorr);
    warm.altNv = -1;
    dev_free(warm.resPartial); dev_free(warm.resOut);
    warm.nv = -1;
    dev_free(snap.pos); dev_free(snap.prev); dev_free(snap.vel); dev_free(snap.corr);
    snap.nv = -1;
    dev_free(render.idx); dev_free(render.map); dev_free(render.vtOff); dev_free(render.vt); dev_free(render.triA); dev_free(render.triB);
    dev_free(render.outV); dev_free(render.outN); dev_free(render.triN);
    render.nv = render.nt = -1;
    dev_free(d_mouse);
    dev_free(jac.lk); dev_free(jac.lr); dev_free(jac.disp); dev_free(jac.d); dev_free(jac.off); dev_free(jac.inc);
    jac.nl = jac.nv = -1;
    warm.built = warm.saved = false;
    dev_free(d_pos); dev_free(d_prev); dev_free(d_vel);
    dev_free(d_tidx); dev_free(d_tA); dev_free(d_tB); dev_free(d_tC); dev_free(d_bidx);
    dev_free(d_ttA); dev_free(d_ttB); dev_free(d_ttC);
    dev_free(d_triColOff); dev_free(d_bendColOff); dev_free(d_lk); dev_free(d_lr);
    dev_free(d_cCount); dev_free(d_cP); dev_free(d_cN); dev_free(d_cSrc); dev_free(d_cFeat); dev_free(d_ctl);
    dev_free(d_geo); dev_free(d_cEdges); dev_free(d_cTris);
    geoHost.clear();
    for (Features* f : {&ec, &tc}) {
        dev_free(f->count); dev_free(f->P); dev_free(f->N); dev_free(f->W); dev_free(f->F);
        dev_free(f->group); dev_free(f->groupBase); dev_free(f->groupCount); dev_free(f->list);
        f->slotCap = 0;
    }
    dev_free(self.triLo); dev_free(self.triHi); dev_free(self.edgeLo); dev_free(self.edgeHi); dev_free(self.vmax);
    dev_free(self.candVT); dev_free(self.candEE); dev_free(self.vtxLo); dev_free(self.vtxHi);
    dev_free(self.candFillVT); dev_free(self.candFillEE);
    self.candCapVT = self.candCapEE = 0;
    self.candWantVT = self.candWantEE = 0;
    dev_free(self.vloc); dev_free(self.edgeVerts4);
    for (SelfBuf& b : self.buf) {
        dev_free(b.ids); dev_free(b.W); dev_free(b.N); dev_free(b.key); dev_free(b.colour); dev_free(b.order);
        dev_free(b.active); dev_free(b.groupOff); dev_free(b.lid); dev_free(b.vlist);
        b.cap = 0;
    }
    self.cap = 0;
    self.built = false;
    for (DeviceSlabs* d : {&triSlabs, &bendSlabs, &linkSlabs}) {
        dev_free(d->lidx); dev_free(d->info); dev_free(d->colOff); dev_free(d->first); dev_free(d->verts);
    }
    for (DevicePacked* d : {&triPacked, &bendPacked}) {
        dev_free(d->verts); dev_free(d->s0); dev_free(d->s1); dev_free(d->ids);
        d->phases.clear();
    }
    slotCap = 0;
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
}

// -------------------------------------------------------------- model ----

void ClothModel::clear()
{
    pos.clear(); prev.clear(); vel.clear(); invMass.clear(); uv.clear();
    tris.clear(); quads.clear(); links.clear(); edges.clear(); bends.clear();
    contactsHost.clear(); contactsValid = false;
    mark_topology_dirty();
}

// Physics/ClothModel.cpp:54-92 (rest state from material coordinates)
void ClothModel::add_triangle(uint32_t i1, uint32_t i2, uint32_t i3)
{
    add_triangle_uv(i1, i2, i3, uv[i1], uv[i2], uv[i3]);
}

void ClothModel::add_triangle_uv(uint32_t i1, uint32_t i2, uint32_t i3, scVector2 uv1, scVector2 uv2, scVector2 uv3)
{
    Triangle t{};
    t.i1 = i1; t.i2 = i2; t.i3 = i3;
    float d1x = uv2.x - uv1.x, d1y = uv2.y - uv1.y;
    float d2x = uv3.x - uv1.x, d2y = uv3.y - uv1.y;
    t.du1 = d1x; t.du2 = d2x; t.dv1 = d1y; t.dv2 = d2y;
    t.det = t.du1 * t.dv2 - t.du2 * t.dv1;
    t.invDet = (t.det != 0.f) ? 1.f / t.det : 0.f;
    f3 dx1 = sub(pos[i2], pos[i1]);
    f3 dx2 = sub(pos[i3], pos[i1]);
    f3 wu = mul(t.invDet, sub(mul(t.dv2, dx1), mul(t.dv1, dx2)));
    f3 wv = mul(t.invDet, add(mul(-t.du2, dx1), mul(t.du1, dx2)));
    t.lu0 = len(wu);
    t.lv0 = len(wv);
    t.su = t.lu0 != 0.f ? 1.f / t.lu0 : 0.f;
    t.sv = t.lv0 != 0.f ? 1.f / t.lv0 : 0.f;
    t.dotuv = dot(wu, wv);
    t.area = 0.5f * t.det;
    t.lu2 = t.lu0 * t.lu0;
    t.lv2 = t.lv0 * t.lv0;
    tris.push_back(t);
}

// Physics/ClothModel.inl:12-31
void ClothModel::add_link(uint32_t i1, uint32_t i2, float stiffness)
{
    if (invMass[i1] == 0.f && invMass[i2] == 0.f) return;
    if (stiffness == 0.f) return;
    links.push_back(Link{i1, i2, len(sub(pos[i1], pos[i2])), stiffness});
}

// Physics/ClothModel.cpp:94-126 (dihedral branch)
void ClothModel::add_bend(const MeshEdge& e)
{
    if (e.t2 < 0) return;
    auto opp = [](const Triangle& t, uint32_t a, uint32_t b) {
        if (t.i1 != a && t.i1 != b) return t.i1;
        if (t.i2 != a && t.i2 != b) return t.i2;
        return t.i3;
    };
    bends.push_back(Bend{e.i1, e.i2, opp(tris[e.t1], e.i1, e.i2), opp(tris[e.t2], e.i1, e.i2)});
}

void ClothModel::sync_to_host()
{
    if (!mDeviceNewer) return;
    const int n = (int)pos.size();
    std::vector<float4> b(n);
    HIP_OK(hipMemcpy(b.data(), gpu.d_pos, n * sizeof(float4), hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++) { pos[i] = f3{b[i].x, b[i].y, b[i].z}; invMass[i] = b[i].w; }
    HIP_OK(hipMemcpy(b.data(), gpu.d_prev, n * sizeof(float4), hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++) prev[i] = f3{b[i].x, b[i].y, b[i].z};
    HIP_OK(hipMemcpy(b.data(), gpu.d_vel, n * sizeof(float4), hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++) vel[i] = f3{b[i].x, b[i].y, b[i].z};
    mDeviceNewer = false;
}

// Slab structure of a schedule: local vertex indices per constraint position,
// per-slab (first vertex, vertex count, first colour, colour count), colour
// offsets and the slabs' global vertex lists.
// With `owners`, bit 31 of a slab vertex id marks the vertex's first slab in
// phase order; returns the number of vertices below `numVertices` that no
// slab contains.
static int upload_slabs(const Schedule& s, DeviceSlabs& d, bool owners, int numVertices, bool lastFlags = false,
                        std::vector<int>* flaggedOut = nullptr)
{
    dev_free(d.lidx); dev_free(d.info); dev_free(d.colOff); dev_free(d.first); dev_free(d.verts);
    if (!s.slabbed || s.order.empty()) return numVertices;
    // a tile schedule's positions are its instances
    const int n = s.tiled ? (int)s.instCons.size() : (int)s.order.size();
    std::vector<ushort4> li(n);
    for (int k = 0; k < n; k++) {
        const uint16_t* l = &s.localIdx[(size_t)k * s.arity];
        li[k] = make_ushort4(l[0], l[1], s.arity > 2 ? l[2] : 0, s.arity > 3 ? l[3] : 0);
    }
    const int ns = (int)s.slabVertBegin.size() - 1;
    std::vector<int4> info(ns);
    std::vector<int> first(ns);
    for (int i = 0; i < ns; i++) {
        info[i] = make_int4(s.slabVertBegin[i], s.slabVertBegin[i + 1] - s.slabVertBegin[i], s.slabColBegin[i],
                            s.slabColBegin[i + 1] - s.slabColBegin[i]);
        first[i] = s.colOffsets[s.slabColBegin[i]];
    }
    std::vector<int> verts(s.slabVerts.begin(), s.slabVerts.end());
    std::vector<char> seen(numVertices, 0);
    int orphans = numVertices;
    for (size_t i = 0; i < verts.size(); i++) {
        const int v = verts[i];
        const bool own = s.tiled ? s.slabOwned[i] != 0 : (v >= 0 && v < numVertices && !seen[v]);
        if (v >= 0 && v < numVertices && !seen[v]) {
            seen[v] = 1;
            orphans--;
        }
        if (owners && own) verts[i] = (int)((unsigned)v | kSlabOwner);
    }
    if (lastFlags) {
        // each vertex's last slab in phase order (slabs are numbered phase-major)
        std::vector<int> lastAt(numVertices, -1);
        for (size_t i = 0; i < s.slabVerts.size(); i++) {
            const int v = (int)s.slabVerts[i];
            if (v >= 0 && v < numVertices) lastAt[v] = (int)i;
        }
        for (int v = 0; v < numVertices; v++)
            if (lastAt[v] >= 0) verts[lastAt[v]] = (int)((unsigned)verts[lastAt[v]] | kSlabContacts);
    }
    dev_alloc(d.lidx, n);
    dev_alloc(d.info, ns);
    dev_alloc(d.colOff, s.colOffsets.size());
    dev_alloc(d.first, ns);
    dev_alloc(d.verts, verts.size());
    HIP_OK(hipMemcpy(d.lidx, li.data(), n * sizeof(ushort4), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(d.info, info.data(), ns * sizeof(int4), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(d.colOff, s.colOffsets.data(), s.colOffsets.size() * sizeof(int), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(d.first, first.data(), ns * sizeof(int), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(d.verts, verts.data(), verts.size() * sizeof(int), hipMemcpyHostToDevice));
    if (flaggedOut) *flaggedOut = std::move(verts);
    return orphans;
}

// Packs a tile schedule (FEM triangles) or a slab schedule (bends) for the
// preloading sweep kernels: per slab kv * kSlabBlock vertex slots, per colour
// of a phase the same number of constraint slots (see PackedFem and
// PackedBend). `flagged` is the slabs' vertex list with the flags
// upload_slabs sets. Leaves `d` without phases when the schedule exceeds the
// kernels' limits: a tile of kPackMaxKv * kSlabBlock vertices or more, a slab
// of more than kPackMaxBendKv * kSlabBlock, a
// colour of a slab wider than the workgroup, too many colours, a tile owning
// more vertices than the workgroup has threads, or vertex ids of 2^29 or more.
static void pack_schedule(bool enabled, const Schedule& s, const std::vector<int>& flagged, int numVertices,
                          const std::vector<Triangle>* tris, DevicePacked& d)
{
    dev_free(d.verts); dev_free(d.s0); dev_free(d.s1); dev_free(d.ids);
    d.phases.clear();
    const bool fem = tris != nullptr;
    if (!enabled || !s.slabbed || s.order.empty() || fem != s.tiled || numVertices >= (int)kPackIdMask) return;
    const int arity = s.arity;
    if (fem ? arity != 3 : arity != 4) return;
    std::vector<DevicePacked::Phase> phases;
    for (int p = 0; p < s.numPhases(); p++) {
        DevicePacked::Phase ph;
        int maxVerts = 0, maxCol = 0, maxWidth = 0;
        for (int sl = s.phaseSlabBegin[p]; sl < s.phaseSlabBegin[p + 1]; sl++) {
            maxVerts = std::max(maxVerts, s.slabVertBegin[sl + 1] - s.slabVertBegin[sl]);
            maxCol = std::max(maxCol, s.slabColBegin[sl + 1] - s.slabColBegin[sl]);
            for (int c = s.slabColBegin[sl]; c < s.slabColBegin[sl + 1]; c++)
                maxWidth = std::max(maxWidth, s.colOffsets[c + 1] - s.colOffsets[c]);
            if (fem) {
                int owned = 0;
                for (int i = s.slabVertBegin[sl]; i < s.slabVertBegin[sl + 1]; i++) owned += s.slabOwned[i] ? 1 : 0;
                if (owned > kSlabBlock) return;
            }
        }
        if (maxVerts > (fem ? kPackMaxKv : kPackMaxBendKv) * kSlabBlock || maxWidth > kSlabBlock || maxCol < 1) return;
        // FEM tiles keep their last vertex slot free as the spare entry
        if (fem && maxVerts >= kPackMaxKv * kSlabBlock) return;
        ph.slabs = s.phaseSlabBegin[p + 1] - s.phaseSlabBegin[p];
        ph.kv = std::max(2, (maxVerts + kSlabBlock - 1) / kSlabBlock);
        ph.nc = maxCol;
        if (fem) {
            // the kernel's colour count (instantiations for 8 and 12)
            if (maxCol > 12) return;
            ph.ncKernel = maxCol <= 8 ? 8 : 12;
            ph.kv = maxVerts < 2 * kSlabBlock ? 2 : 4;
            ph.width = kSlabBlock;
        } else {
            if (maxCol > kPackMaxBendColours) return;
            ph.ncKernel = maxCol;
            ph.width = std::max(64, (maxWidth + 63) / 64 * 64);
        }
        phases.push_back(ph);
    }
    size_t nVert = 0, nSlot = 0;
    for (DevicePacked::Phase& ph : phases) {
        ph.vertOff = nVert;
        ph.slotOff = nSlot;
        nVert += (size_t)ph.slabs * ph.kv * kSlabBlock;
        nSlot += (size_t)ph.slabs * ph.ncKernel * ph.width;
    }
    std::vector<int> verts(nVert);
    std::vector<uint4> s0;
    std::vector<float4> s1, s2;
    std::vector<uint2> ids;
    if (fem) {
        s0.assign(nSlot, make_uint4(0, 0, 0, 0));
        s1.assign(nSlot, make_float4(0, 0, 0, 0));
    } else {
        ids.assign(nSlot, make_uint2(0, 0));
    }
    std::vector<int> newLocal;
    for (int p = 0; p < (int)phases.size(); p++) {
        const DevicePacked::Phase& ph = phases[p];
        const int slots = ph.kv * kSlabBlock;
        const int spare = fem ? slots - 1 : slots;
        for (int j = 0; j < ph.slabs; j++) {
            const int sl = s.phaseSlabBegin[p] + j;
            const int vb = s.slabVertBegin[sl], nvs = s.slabVertBegin[sl + 1] - vb;
            // vertex slots: a tile's own vertices first (one per thread), then
            // the rest in list order, then padding repeating the first
            int* vs = verts.data() + ph.vertOff + (size_t)j * slots;
            newLocal.assign(nvs, -1);
            int k = 0;
            if (fem)
                for (int i = 0; i < nvs; i++)
                    if (s.slabOwned[vb + i]) newLocal[i] = k++;
            for (int i = 0; i < nvs; i++)
                if (newLocal[i] < 0) newLocal[i] = k++;
            for (int i = 0; i < nvs; i++) vs[newLocal[i]] = flagged[vb + i];
            for (int i = nvs; i < slots; i++) vs[i] = (int)(((unsigned)flagged[vb] & (unsigned)kPackIdMask) | kPackPad);
            // constraint slots, colour by colour
            const int nc = s.slabColBegin[sl + 1] - s.slabColBegin[sl];
            for (int c = 0; c < ph.ncKernel; c++) {
                const size_t base = ph.slotOff + ((size_t)j * ph.ncKernel + c) * ph.width;
                const int lo = c < nc ? s.colOffsets[s.slabColBegin[sl] + c] : 0;
                const int cnt = c < nc ? s.colOffsets[s.slabColBegin[sl] + c + 1] - lo : 0;
                for (int i = 0; i < ph.width; i++) {
                    uint16_t l[4] = {(uint16_t)spare, (uint16_t)spare, (uint16_t)spare, (uint16_t)spare};
                    if (i < cnt)
                        for (int a = 0; a < arity; a++) l[a] = (uint16_t)newLocal[s.localIdx[(size_t)(lo + i) * arity + a]];
                    if (fem) {
                        uint4 w = make_uint4((unsigned)l[0] | ((unsigned)l[1] << 10) | ((unsigned)l[2] << 20), 0, 0, 0);
                        if (i < cnt) {
                            const Triangle& t = (*tris)[s.instCons[lo + i]];
                            const float rest[3] = {t.invDet, t.lu0, t.lv0};
                            std::memcpy(&w.y, &rest[0], 4);
                            std::memcpy(&w.z, &rest[1], 4);
                            std::memcpy(&w.w, &rest[2], 4);
                            s1[base + i] = make_float4(t.du1, t.du2, t.dv1, t.dv2);
                        }
                        s0[base + i] = w;
                    } else {
                        ids[base + i] = make_uint2((unsigned)l[0] | ((unsigned)l[1] << 16),
                                                   (unsigned)l[2] | ((unsigned)l[3] << 16));
                    }
                }
            }
        }
    }
    dev_alloc(d.verts, nVert);
    HIP_OK(hipMemcpy(d.verts, verts.data(), nVert * sizeof(int), hipMemcpyHostToDevice));
    if (fem) {
        dev_alloc(d.s0, nSlot); dev_alloc(d.s1, nSlot);
        HIP_OK(hipMemcpy(d.s0, s0.data(), nSlot * sizeof(uint4), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(d.s1, s1.data(), nSlot * sizeof(float4), hipMemcpyHostToDevice));
    } else {
        dev_alloc(d.ids, nSlot);
        HIP_OK(hipMemcpy(d.ids, ids.data(), nSlot * sizeof(uint2), hipMemcpyHostToDevice));
    }
    d.phases = std::move(phases);
}

void ClothModel::build_device_topology()
{
    GpuSolver& g = gpu;
    if (!g.stream) HIP_OK(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
    g.release_graphs();
    g.warm.built = false;
    g.jac.nl = g.jac.nv = -1;  // rebuilt from the new links when a Jacobi-type solver runs
    g.render.nv = g.render.nt = -1;  // the render topology and weld map are rebuilt on the next sync
    g.nv = (int)pos.size();
    g.nt = (int)tris.size();
    g.nb = (int)bends.size();
    dev_alloc(g.d_pos, g.nv);
    dev_alloc(g.d_prev, g.nv);
    dev_alloc(g.d_vel, g.nv);
    if (!g.d_ctl) dev_alloc(g.d_ctl, 1);
    HIP_OK(hipMemset(g.d_ctl, 0, sizeof(SolverCtl)));

    ensure_colouring();
    // triangles in colour order
    std::vector<int4> ti(g.nt);
    std::vector<float4> ta(g.nt), tb(g.nt);
    std::vector<float2> tc(g.nt);
    for (int k = 0; k < g.nt; k++) {
        const Triangle& t = tris[triCol.order[k]];
        ti[k] = make_int4((int)t.i1, (int)t.i2, (int)t.i3, triCol.order[k]);
        ta[k] = make_float4(t.du1, t.du2, t.dv1, t.dv2);
        tb[k] = make_float4(t.invDet, t.lu2, t.lv2, t.area);
        tc[k] = make_float2(t.lu0, t.lv0);
    }
    dev_free(g.d_ttA); dev_free(g.d_ttB); dev_free(g.d_ttC);
    if (triCol.tiled) {
        const size_t ni = triCol.instCons.size();
        std::vector<float4> ia(ni), ib(ni);
        std::vector<float2> ic(ni);
        for (size_t k = 0; k < ni; k++) {
            const Triangle& t = tris[triCol.instCons[k]];
            ia[k] = make_float4(t.du1, t.du2, t.dv1, t.dv2);
            ib[k] = make_float4(t.invDet, t.lu2, t.lv2, t.area);
            ic[k] = make_float2(t.lu0, t.lv0);
        }
        dev_alloc(g.d_ttA, std::max<size_t>(ni, 1)); dev_alloc(g.d_ttB, std::max<size_t>(ni, 1));
        dev_alloc(g.d_ttC, std::max<size_t>(ni, 1));
        if (ni) {
            HIP_OK(hipMemcpy(g.d_ttA, ia.data(), ni * sizeof(float4), hipMemcpyHostToDevice));
            HIP_OK(hipMemcpy(g.d_ttB, ib.data(), ni * sizeof(float4), hipMemcpyHostToDevice));
            HIP_OK(hipMemcpy(g.d_ttC, ic.data(), ni * sizeof(float2), hipMemcpyHostToDevice));
        }
    }
    dev_alloc(g.d_tidx, g.nt); dev_alloc(g.d_tA, g.nt); dev_alloc(g.d_tB, g.nt); dev_alloc(g.d_tC, g.nt);
    if (g.nt) {
        HIP_OK(hipMemcpy(g.d_tidx, ti.data(), g.nt * sizeof(int4), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(g.d_tA, ta.data(), g.nt * sizeof(float4), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(g.d_tB, tb.data(), g.nt * sizeof(float4), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(g.d_tC, tc.data(), g.nt * sizeof(float2), hipMemcpyHostToDevice));
    }
    // bends in colour order
    std::vector<int4> bi(g.nb);
    for (int k = 0; k < g.nb; k++) {
        const Bend& b = bends[bendCol.order[k]];
        bi[k] = make_int4((int)b.i1, (int)b.i2, (int)b.i3, (int)b.i4);
    }
    dev_alloc(g.d_bidx, g.nb);
    if (g.nb) HIP_OK(hipMemcpy(g.d_bidx, bi.data(), g.nb * sizeof(int4), hipMemcpyHostToDevice));
    dev_alloc(g.d_triColOff, triCol.phaseOffsets.size());
    HIP_OK(hipMemcpy(g.d_triColOff, triCol.phaseOffsets.data(), triCol.phaseOffsets.size() * sizeof(int),
                     hipMemcpyHostToDevice));
    dev_alloc(g.d_bendColOff, bendCol.phaseOffsets.size());
    HIP_OK(hipMemcpy(g.d_bendColOff, bendCol.phaseOffsets.data(), bendCol.phaseOffsets.size() * sizeof(int),
                     hipMemcpyHostToDevice));
    std::vector<int> triFlagged, bendFlagged;
    g.triOrphans = upload_slabs(triCol, g.triSlabs, true, g.nv, false, &triFlagged);
    g.bendOrphans = upload_slabs(bendCol, g.bendSlabs, false, g.nv, true, &bendFlagged);
    const bool packed = mSweepMode == SWEEP_PACKED;
    pack_schedule(packed, triCol, triFlagged, g.nv, &tris, g.triPacked);
    pack_schedule(packed, bendCol, bendFlagged, g.nv, nullptr, g.bendPacked);
    // links in schedule order
    g.nl = (int)links.size();
    std::vector<int2> lk(g.nl);
    std::vector<float2> lr(g.nl);
    for (int k = 0; k < g.nl; k++) {
        const Link& l = links[linkCol.order[k]];
        lk[k] = make_int2((int)l.i1, (int)l.i2);
        lr[k] = make_float2(l.len, l.stiffness);
    }
    dev_alloc(g.d_lk, g.nl);
    dev_alloc(g.d_lr, g.nl);
    if (g.nl) {
        HIP_OK(hipMemcpy(g.d_lk, lk.data(), g.nl * sizeof(int2), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(g.d_lr, lr.data(), g.nl * sizeof(float2), hipMemcpyHostToDevice));
    }
    upload_slabs(linkCol, g.linkSlabs, false, g.nv);
    dev_alloc(g.d_cCount, g.nv);
    if (g.nv) HIP_OK(hipMemset(g.d_cCount, 0, g.nv * sizeof(int)));
    g.slotCap = 0;
    dev_free(g.d_cP);
    dev_free(g.d_cN);
    dev_free(g.d_cSrc);
    dev_free(g.d_cFeat);
    // cloth edges and triangles in index order, for obstacle-feature contacts
    g.ne = (int)edges.size();
    std::vector<int2> he(g.ne);
    for (int e = 0; e < g.ne; e++) he[e] = make_int2((int)edges[e].i1, (int)edges[e].i2);
    std::vector<int4> ht(g.nt);
    for (int t = 0; t < g.nt; t++) ht[t] = make_int4((int)tris[t].i1, (int)tris[t].i2, (int)tris[t].i3, 0);
    dev_alloc(g.d_cEdges, g.ne);
    dev_alloc(g.d_cTris, g.nt);
    if (g.ne) HIP_OK(hipMemcpy(g.d_cEdges, he.data(), g.ne * sizeof(int2), hipMemcpyHostToDevice));
    if (g.nt) HIP_OK(hipMemcpy(g.d_cTris, ht.data(), g.nt * sizeof(int4), hipMemcpyHostToDevice));
    auto groups = [&](GpuSolver::Features& f, const Schedule& c, int n) {
        f.n = n;
        f.groups = c.numPhases() > 0 ? c.numPhases() : 0;
        std::vector<int> grp(n);
        for (int p = 0; p < f.groups; p++)
            for (int k = c.phaseOffsets[p]; k < c.phaseOffsets[p + 1]; k++) grp[c.order[k]] = p;
        dev_alloc(f.group, n);
        dev_alloc(f.groupBase, f.groups + 1);
        dev_alloc(f.groupCount, std::max(f.groups, 1) * (size_t)kGroupStride);
        dev_alloc(f.list, n);
        dev_alloc(f.count, n);
        if (n) HIP_OK(hipMemcpy(f.group, grp.data(), n * sizeof(int), hipMemcpyHostToDevice));
        if (n) HIP_OK(hipMemset(f.count, 0, n * sizeof(int)));
        HIP_OK(hipMemset(f.groupCount, 0, std::max(f.groups, 1) * kGroupStride * sizeof(int)));
        HIP_OK(hipMemcpy(f.groupBase, c.phaseOffsets.data(), (f.groups + 1) * sizeof(int), hipMemcpyHostToDevice));
        f.base = f.groups > 0 ? std::vector<int>(c.phaseOffsets.begin(), c.phaseOffsets.begin() + f.groups + 1)
                              : std::vector<int>(1, 0);
        f.slotCap = 0;
        f.slots = 0;
        dev_free(f.P); dev_free(f.N); dev_free(f.W); dev_free(f.F);
    };
    groups(g.ec, edgeContactCol, g.ne);
    groups(g.tc, triContactCol, g.nt);
    g.self.built = false;
    mTopoDirty = false;
}

// Own vertices per tile of a tile schedule: one per thread of a slab
// workgroup, so that a tile's extended region (own vertices and a halo of as
// many rings as colours, about four times as many on a grid) has about one
// constraint per thread per colour.
constexpr int kTileVerts = kSlabBlock;
static_assert(kMaxTileColours <= kMaxSlabColours, "tile colours must fit the slab kernels");

void ClothModel::ensure_colouring()
{
    if (!mColourDirty) return;
    const int nv = (int)pos.size(), nt = (int)tris.size(), nb = (int)bends.size();
    std::vector<uint32_t> tv(3 * (size_t)nt);
    for (int t = 0; t < nt; t++) { tv[3 * t] = tris[t].i1; tv[3 * t + 1] = tris[t].i2; tv[3 * t + 2] = tris[t].i3; }
    std::vector<uint32_t> bv(4 * (size_t)nb);
    for (int b = 0; b < nb; b++) {
        bv[4 * b] = bends[b].i1; bv[4 * b + 1] = bends[b].i2; bv[4 * b + 2] = bends[b].i3; bv[4 * b + 3] = bends[b].i4;
    }
    const int ne = (int)edges.size();
    std::vector<uint32_t> ev(2 * (size_t)ne);
    for (int e = 0; e < ne; e++) { ev[2 * e] = edges[e].i1; ev[2 * e + 1] = edges[e].i2; }
    edgeContactCol = schedule_colours(ne, 2, ev.data(), nv);
    triContactCol = schedule_colours(nt, 3, tv.data(), nv);
    const int nl = (int)links.size();
    std::vector<uint32_t> lv(2 * (size_t)nl);
    for (int l = 0; l < nl; l++) { lv[2 * l] = links[l].i1; lv[2 * l + 1] = links[l].i2; }
    // LDS budget of a slab workgroup: 4000 float4 vertices (64,000 bytes)
    const int maxVerts = 4000;
    if (mSweepMode == SWEEP_COLOURS) {
        triCol = schedule_colours(nt, 3, tv.data(), nv);
        bendCol = schedule_colours(nb, 4, bv.data(), nv);
        linkCol = schedule_colours(nl, 2, lv.data(), nv);
    } else if (mSlabSize > 0 && mSweepMode != SWEEP_TILES && mSweepMode != SWEEP_PACKED) {
        const bool seamed = mSweepMode == SWEEP_SEAMS;
        triCol = schedule_slabs(nt, 3, tv.data(), nv, pos.data(), mSlabSize, maxVerts, seamed);
        bendCol = schedule_slabs(nb, 4, bv.data(), nv, pos.data(), mSlabSize, maxVerts, seamed);
        linkCol = schedule_slabs(nl, 2, lv.data(), nv, pos.data(), mSlabSize, maxVerts, seamed);
    } else {
        bool tiled = false;
        if ((mSweepMode == SWEEP_TILES || mSweepMode == SWEEP_PACKED) && nt > 0) {
            // tiles of about one owned vertex per thread of the workgroup;
            // a mesh whose halo makes a tile too large keeps the slabs
            try {
                triCol = schedule_tiles(nt, 3, tv.data(), nv, pos.data(), kTileVerts, maxVerts);
                tiled = true;
            } catch (const std::runtime_error&) {
                tiled = false;
            }
        }
        // a slab size given with the tiled modes sets the interiors of the
        // other kinds' interior and seam slabs
        auto slabs = [&](int n, int arity, const uint32_t* v) {
            return mSlabSize > 0 ? schedule_slabs(n, arity, v, nv, pos.data(), mSlabSize, maxVerts, true)
                                 : schedule_slabs_auto(n, arity, v, nv, pos.data(), maxVerts, kSlabBlock);
        };
        if (!tiled) triCol = slabs(nt, 3, tv.data());
        bendCol = slabs(nb, 4, bv.data());
        linkCol = slabs(nl, 2, lv.data());
    }
    mColourDirty = false;
}

void ClothModel::set_sweep(int mode, int slabSize)
{
    if (mode != SWEEP_COLOURS && mode != SWEEP_SLABS && mode != SWEEP_SEAMS && mode != SWEEP_TILES &&
        mode != SWEEP_PACKED)
        throw std::runtime_error("unknown sweep mode");
    if (slabSize < 0 || slabSize > 65536) throw std::runtime_error("slab size out of range");
    if (mode == mSweepMode && slabSize == mSlabSize) return;
    sync_to_host();
    mSweepMode = mode;
    mSlabSize = slabSize;
    mColourDirty = true;
    mTopoDirty = true;
    mHostDirty = true;
}

void ClothModel::sync_to_device()
{
    if (mTopoDirty) build_device_topology();
    if (!mHostDirty) return;
    const int n = (int)pos.size();
    std::vector<float4> b(n);
    for (int i = 0; i < n; i++) b[i] = make_float4(pos[i].x, pos[i].y, pos[i].z, invMass[i]);
    if (n) HIP_OK(hipMemcpy(gpu.d_pos, b.data(), n * sizeof(float4), hipMemcpyHostToDevice));
    for (int i = 0; i < n; i++) b[i] = make_float4(prev[i].x, prev[i].y, prev[i].z, 0.f);
    if (n) HIP_OK(hipMemcpy(gpu.d_prev, b.data(), n * sizeof(float4), hipMemcpyHostToDevice));
    for (int i = 0; i < n; i++) b[i] = make_float4(vel[i].x, vel[i].y, vel[i].z, 0.f);
    if (n) HIP_OK(hipMemcpy(gpu.d_vel, b.data(), n * sizeof(float4), hipMemcpyHostToDevice));
    mHostDirty = false;
}

void ClothModel::ensure_contact_capacity(int slots)
{
    GpuSolver& g = gpu;
    g.slots = slots;
    if (slots <= g.slotCap) return;
    dev_alloc(g.d_cP, (size_t)slots * g.nv);
    dev_alloc(g.d_cN, (size_t)slots * g.nv);
    dev_alloc(g.d_cSrc, (size_t)slots * g.nv);
    dev_alloc(g.d_cFeat, (size_t)slots * g.nv);
    g.slotCap = slots;
    g.release_graphs();
}

void ClothModel::ensure_feature_capacity(int edgeSlots, int triSlots)
{
    GpuSolver& g = gpu;
    bool grown = false;
    for (auto fs : {std::make_pair(&g.ec, edgeSlots), std::make_pair(&g.tc, triSlots)}) {
        GpuSolver::Features& f = *fs.first;
        f.slots = fs.second;
        if (f.slots <= f.slotCap) continue;
        const size_t m = (size_t)f.slots * f.n;
        dev_alloc(f.P, m); dev_alloc(f.N, m); dev_alloc(f.W, m); dev_alloc(f.F, m);
        f.slotCap = f.slots;
        grown = true;
    }
    if (grown) {
        g.release_graphs();
    }
}

SelfParams ClothModel::self_params() const
{
    SelfParams sp;
    sp.tol = mThickness + mTolerance * 0.5f;
    sp.thickness = mThickness;
    sp.shrink = mSelfShrink;
    sp.volume = (double)mCcdVolume;
    sp.bisect = (double)mCcdBisect;
    sp.vtDistance = mCcdVtDistance;
    sp.eeWeights = mSelfEdgeWeights;
    return sp;
}

// Buffers of the self-collision path and a fresh LBVH pair over the current
// swept boxes (called outside graph capture, once per step).
void ClothModel::ensure_self()
{
    GpuSolver& g = gpu;
    GpuSolver::Self& S = g.self;
    const int cap = std::max(mSelfCapacity > 0 ? mSelfCapacity : std::max(4096, 2 * (g.nv + g.ne)),
                             mSelfCapacityGrown);
    if (cap != S.cap) {
        for (SelfBuf& b : S.buf) {
            dev_alloc(b.ids, cap); dev_alloc(b.W, cap); dev_alloc(b.N, cap); dev_alloc(b.key, cap);
            dev_alloc(b.colour, cap); dev_alloc(b.order, cap); dev_alloc(b.active, cap);
            dev_alloc(b.groupOff, (size_t)cap + 1);  // at most one colour per contact
            dev_alloc(b.lid, cap);
            b.cap = cap;
            b.maxColours = cap;
        }
        S.cap = cap;
        g.release_graphs();
    }
    // candidate pools: 80 pairs per vertex and 64 per edge to start with (the
    // 512^2 perturbed cloth with 2 mm thickness has about 65 and 55), grown to
    // 1.5 x a step's peak when it came near the end, in whole chunks
    auto chunks = [](long pairs) {
        const long c = (std::max(pairs, (long)kCandChunk) + kCandChunk - 1) / kCandChunk * kCandChunk;
        return (int)std::min(c, (long)(INT_MAX / kCandChunk) * kCandChunk);
    };
    const int capVT = chunks(std::max(80L * g.nv, S.candWantVT));
    const int capEE = chunks(std::max(64L * g.ne, S.candWantEE));
    if (capVT != S.candCapVT || capEE != S.candCapEE) {
        dev_alloc(S.candVT, (size_t)capVT);
        dev_alloc(S.candEE, (size_t)capEE);
        dev_alloc(S.candFillVT, capVT / kCandChunk);
        dev_alloc(S.candFillEE, capEE / kCandChunk);
        S.candCapVT = capVT;
        S.candCapEE = capEE;
        g.release_graphs();
    }
    if (!S.triLo) {
        dev_alloc(S.triLo, std::max(g.nt, 1)); dev_alloc(S.triHi, std::max(g.nt, 1));
        dev_alloc(S.edgeLo, std::max(g.ne, 1)); dev_alloc(S.edgeHi, std::max(g.ne, 1));
        dev_alloc(S.vtxLo, std::max(g.nv, 1)); dev_alloc(S.vtxHi, std::max(g.nv, 1));
        dev_alloc(S.vmax, 2 * (size_t)std::max(g.nv, 1));
        // the colouring's vertex claims: every byte 0x7f between colourings
        dev_alloc(S.vloc, 2 * (size_t)std::max(g.nv, 1));
        HIP_OK(hipMemset(S.vloc, 0x7f, 2 * sizeof(int) * std::max(g.nv, 1)));
        for (SelfBuf& b : S.buf) dev_alloc(b.vlist, std::max(g.nv, 1));
        std::vector<int4> ev(std::max(g.ne, 1), make_int4(-1, -1, -1, -1));
        for (int e = 0; e < g.ne; e++) ev[e] = make_int4((int)edges[e].i1, (int)edges[e].i2, -1, -1);
        dev_alloc(S.edgeVerts4, ev.size());
        HIP_OK(hipMemcpy(S.edgeVerts4, ev.data(), ev.size() * sizeof(int4), hipMemcpyHostToDevice));
    }
    // hierarchy from the current configuration; substeps refit it (the
    // vertex tree only orders the vertex queries, so it is not refit)
    launch_self_boxes(g.stream, g.d_pos, g.d_prev, g.d_cTris, g.nt, g.d_cEdges, g.ne, mThickness + mTolerance * 0.5f,
                      S.triLo, S.triHi, S.edgeLo, S.edgeHi, g.nv, S.vtxLo, S.vtxHi);
    S.triB.set_payload(g.d_cTris);
    S.edgeB.set_payload(S.edgeVerts4);
    S.triB.build(g.stream, S.triLo, S.triHi, g.nt);
    S.edgeB.build(g.stream, S.edgeLo, S.edgeHi, g.ne);
    S.vtxB.build(g.stream, S.vtxLo, S.vtxHi, g.nv);
    S.built = true;
}

FeatureContactBuf ClothModel::feature_buf(GpuSolver::Features& f) const
{
    return FeatureContactBuf{f.count, f.P, f.N, f.W, f.F, f.group, f.groupBase, f.groupCount, f.list,
                             f.n, f.slots, f.groups};
}

// Device descriptors of the world's collidables, in world order. Mesh
// obstacles are uploaded (with their LBVHs) on first use and their
// descriptors written to the device table, indexed by world position.
CollSet ClothModel::make_collset(int& slots, int& edgeSlots, int& triSlots)
{
    CollSet cs;
    memset(&cs, 0, sizeof(cs));
    slots = 0;
    edgeSlots = 0;
    triSlots = 0;
    int meshes = 0;
    const bool fe = (mCollFlags & CF_EDGES) != 0, ft = (mCollFlags & CF_TRIANGLES) != 0;
    GpuSolver& gs = gpu;
    std::vector<MeshGeo> geo(MAX_COLLIDABLES);
    memset(geo.data(), 0, geo.size() * sizeof(MeshGeo));
    const CollisionWorld* w = mOwner->mWorld;
    if (w->colls.size() > MAX_COLLIDABLES) throw std::runtime_error("too many collidables (max 8)");
    for (const auto& sp : w->colls) {
        const Collidable* c = sp.get();
        CollDev d;
        memset(&d, 0, sizeof(d));
        d.type = c->type;
        if (c->type == CT_WALLS) {
            const Walls* wl = (const Walls*)c;
            d.ax = wl->box.min.x; d.ay = wl->box.min.y; d.az = wl->box.min.z;
            d.bx = wl->box.max.x; d.by = wl->box.max.y; d.bz = wl->box.max.z;
            if (mCollFlags & CF_WALLS) slots += 5;
        } else if (c->type == CT_SPHERE) {
            const Sphere* s = (const Sphere*)c;
            d.ax = s->center.x; d.ay = s->center.y; d.az = s->center.z; d.r = s->radius;
            if (mCollFlags & CF_VERTICES) slots += 1;
            edgeSlots += fe;
            triSlots += ft;
        } else if (c->type == CT_SDF) {
            const SDF* s = ((const CollisionSDF*)c)->sdf;
            d.ax = s->box.min.x; d.ay = s->box.min.y; d.az = s->box.min.z;
            d.bx = s->steps.x; d.by = s->steps.y; d.bz = s->steps.z;
            d.nodes = s->d_nodes;
            d.nc0 = s->nc[0]; d.nc1 = s->nc[1]; d.nc2 = s->nc[2];
            d.nv0 = s->nv[0]; d.nv1 = s->nv[1];
            if (mCollFlags & CF_VERTICES) slots += 1;
            // GridSDF::QuerySegment is a stub (ClothCollisionSDF.h:141-146):
            // with vertex contacts on the border test rejects every edge, and
            // without them the reference produces undefined contacts
            if (fe && !(mCollFlags & CF_VERTICES))
                throw std::runtime_error("edge contacts against a grid SDF need CF_VERTICES "
                                         "(the reference's grid-SDF segment query is a stub)");
            triSlots += ft;
        } else if (c->type == CT_CAPSULE) {
            const Capsule* k = (const Capsule*)c;
            f3 P, Q;
            capsule_axis(*k, P, Q);
            d.ax = P.x; d.ay = P.y; d.az = P.z;
            d.bx = Q.x; d.by = Q.y; d.bz = Q.z;
            d.r = k->r;
            if (mCollFlags & CF_VERTICES) slots += 1;
            edgeSlots += fe;
            triSlots += ft;
        } else if (c->type == CT_MESH) {
            Mesh* me = ((const CollisionMesh*)c)->mesh;
            if (!me) throw std::runtime_error("mesh collidable without a mesh");
            if (!gs.stream) HIP_OK(hipStreamCreateWithFlags(&gs.stream, hipStreamNonBlocking));
            geo[cs.n] = me->device_geo(gs.stream);
            if (!gs.d_geo) dev_alloc(gs.d_geo, MAX_COLLIDABLES);
            d.mesh = gs.d_geo + cs.n;
            if (mCollFlags & CF_VERTICES) slots += 1;
            edgeSlots += fe;
            triSlots += ft;
            meshes++;
        } else {
            throw std::runtime_error("collidable type " + std::to_string(c->type) +
                                     " is not supported by the GPU substep");
        }
        cs.c[cs.n++] = d;
    }
    if (meshes) {
        const bool same = gs.geoHost.size() == geo.size() &&
                          memcmp(gs.geoHost.data(), geo.data(), geo.size() * sizeof(MeshGeo)) == 0;
        if (!same) {
            HIP_OK(hipMemcpy(gs.d_geo, geo.data(), geo.size() * sizeof(MeshGeo), hipMemcpyHostToDevice));
            gs.geoHost = geo;
        }
    }
    cs.maxSlots = slots;
    return cs;
}

std::string ClothModel::graph_key(float h, const CollSet& cs) const
{
    std::string k;
    auto put = [&](const void* p, size_t n) { k.append((const char*)p, n); };
    put(&h, 4); put(&mUnit, 4); put(&mThickness, 4); put(&mTolerance, 4); put(&mShearStiff, 4);
    put(&mBendStiff, 4); put(&mPoisson, 4); put(&mFriction, 4); put(&mDragCoeff, 4);
    put(&mNumIterations, 4); put(&mCollFlags, 4); put(&mBendFreq, 4);
    put(&mArea, 4); put(&mDihedral, 1); put(&mFEM, 1); put(&mFixedIterations, 1); put(&mSweepMode, 4); put(&mSlabSize, 4);
    put(&msActive, 1); put(&gpu.d_mouse, sizeof(void*)); put(&mSolver, 4); put(&gpu.jac.lk, sizeof(void*));
    put(&cs, sizeof(cs));
    put(&gpu.d_cP, sizeof(void*));
    put(&gpu.d_cFeat, sizeof(void*));
    put(&gpu.ec.P, sizeof(void*)); put(&gpu.ec.slots, 4);
    put(&gpu.tc.P, sizeof(void*)); put(&gpu.tc.slots, 4);
    put(&mSelfShrink, 4); put(&mSelfCollThreshold, 4); put(&mCcdVolume, 4); put(&mCcdBisect, 4);
    put(&mCcdVtDistance, 4); put(&gpu.self.cap, 4); put(&mSelfEdgeWeights, 4);
    const BvhView tv = gpu.self.triB.view(), ev = gpu.self.edgeB.view();
    put(&tv, sizeof(tv)); put(&ev, sizeof(ev));
    put(&gpu.self.buf[0].ids, sizeof(void*)); put(&gpu.self.buf[1].ids, sizeof(void*));
    put(&gpu.self.candVT, sizeof(void*)); put(&gpu.self.candEE, sizeof(void*)); put(&gpu.self.candCapVT, 4);
    put(&gpu.self.candCapEE, 4); put(&gpu.self.candFillVT, sizeof(void*)); put(&gpu.self.candFillEE, sizeof(void*));
    const BvhView vv = gpu.self.vtxB.view();
    put(&vv, sizeof(vv));
    const int wmode = mWarmPlainNow ? 0 : mWarmMode;
    put(&mWarm, sizeof(void*)); put(&wmode, 4); put(&mWarmIters, 4);
    put(&gpu.warm.corr, sizeof(void*)); put(&gpu.warm.X, sizeof(void*)); put(&gpu.warm.triStrain, sizeof(void*));
    put(&gpu.warm.adj, sizeof(void*)); put(&gpu.warm.vt, sizeof(void*)); put(&gpu.warm.nFree, 4);
    return k;
}

static inline void timed_begin(KernelTimer& t, bool timed, int which, hipStream_t s, hipEvent_t& a)
{
    if (!timed) return;
    a = t.get();
    HIP_OK(hipEventRecord(a, s));
    (void)which;
}

static inline void timed_end(KernelTimer& t, bool timed, int which, hipStream_t s, hipEvent_t a)
{
    if (!timed) return;
    hipEvent_t b = t.get();
    HIP_OK(hipEventRecord(b, s));
    t.pending.push_back({which, {a, b}});
}

// One substep (Physics/ClothPBD.cpp:255-322) as a fixed kernel sequence. In
// early-exit mode every iteration is recorded and the kernels of iterations
// after the exit return at once (SolverCtl::done).
void ClothModel::record_substep(float h, const CollSet& cs, bool timed)
{
    GpuSolver& g = gpu;
    hipStream_t s = g.stream;
    hipEvent_t ev = nullptr;
    // h * gravity * unit, evaluated in the reference's order (ClothPBD.cpp:263)
    const f3 gravity{0.f, -9.8f, 0.f};
    const f3 hg = mul(h, gravity);
    const f3 gdv = f3{mUnit * hg.x, mUnit * hg.y, mUnit * hg.z};
    bool anyMesh = false;
    for (int i = 0; i < cs.n; i++) anyMesh = anyMesh || cs.c[i].type == CT_MESH;
    ContactBuf cb{g.d_cCount, g.d_cP, g.d_cN, g.d_cSrc, g.d_cFeat, anyMesh ? g.d_geo : nullptr, g.nv, g.slots};
    const bool anyContacts = cs.maxSlots > 0;
    const bool edgeContacts = (mCollFlags & CF_EDGES) && g.ec.slots > 0 && g.ne > 0;
    const bool triContacts = (mCollFlags & CF_TRIANGLES) && g.tc.slots > 0 && g.nt > 0;
    const bool selfColl = (mCollFlags & CF_SELF) && g.self.built;
    const SelfParams sp = self_params();
    const FeatureContactBuf ecb = feature_buf(g.ec), tcb = feature_buf(g.tc);
    const float invH = 1.f / h;
    // With a fixed iteration count and every vertex in some triangle slab, the
    // vertex contacts of iteration k are applied by the FEM slab kernels of
    // iteration k+1 (by each vertex's owner slab, before its constraints) and
    // the last iteration's by the velocity kernel: one launch per iteration
    // fewer, same sequential order. The early-exit test needs the contact
    // error at the end of each iteration, so that mode keeps separate launches.
    // Edge and triangle contacts follow the vertex contacts inside an
    // iteration, so they also keep the launches separate.
    const bool fuse = mSolver == SOLVER_GAUSS_SEIDEL && anyContacts && mFEM && mFixedIterations && triCol.slabbed &&
                      !triCol.tiled && triCol.numPhases() > 0 &&
                      g.triOrphans == 0 && !edgeContacts && !triContacts && !selfColl;
    // A tile schedule stages every tile's halo in the same launch, so a
    // vertex's contacts cannot be applied while staging (another tile may
    // stage it before or after its owner updates the contact state).
    // Instead, in an iteration without a bend pass the FEM launch applies the
    // iteration's vertex contacts at its end, each tile to its own vertices;
    // iterations with a bend pass keep the contact launch after the bends.
    const bool fuseTiles = mSolver == SOLVER_GAUSS_SEIDEL && anyContacts && mFEM && mFixedIterations &&
                           triCol.tiled && !edgeContacts && !triContacts && !selfColl;
    if (anyMesh || edgeContacts || triContacts || selfColl) {
        // the mesh query radius needs max |pos - prev| over all vertices
        // before any detection, and the contact groups are cleared there, so
        // prediction and detection are separate launches
        timed_begin(g.timer, timed, 4, s, ev);
        launch_predict_mesh(s, g.d_pos, g.d_prev, g.d_vel, g.nv, h, gdv, mDragCoeff, g.d_ctl, ecb, tcb);
        if (msActive) {
            // the spring moves its particles after the prediction and before
            // the travel maximum is taken (ClothPBD.cpp:267-285)
            launch_mouse_spring(s, g.d_pos, g.d_vel, g.d_mouse, h, g.d_ctl, 1);
            launch_max_travel(s, g.d_pos, g.d_prev, g.nv, g.d_ctl);
        }
        timed_end(g.timer, timed, 4, s, ev);
        timed_begin(g.timer, timed, 3, s, ev);
        if (anyContacts)
            launch_detect(s, g.d_pos, g.d_prev, g.nv, cs, cb, mThickness, mTolerance, mCollFlags, g.d_ctl);
        if (edgeContacts || triContacts)
            launch_detect_features(s, g.d_pos, g.d_prev, g.d_cEdges, g.d_cTris, cs, ecb, tcb, mThickness,
                                   mTolerance, mCollFlags, g.d_ctl);
        if (selfColl) {
            GpuSolver::Self& S = g.self;
            launch_self_boxes(s, g.d_pos, g.d_prev, g.d_cTris, g.nt, g.d_cEdges, g.ne, sp.tol, S.triLo, S.triHi,
                              S.edgeLo, S.edgeHi, g.nv, nullptr, nullptr);
            S.triB.refit(s, S.triLo, S.triHi);
            S.edgeB.refit(s, S.edgeLo, S.edgeHi);
            launch_self_detect(s, g.d_pos, g.d_prev, g.d_cTris, g.d_cEdges, g.nv, g.ne, S.triB.view(),
                               S.edgeB.view(), S.vtxB.view().leafPrim, sp,
                               SelfCandidates{S.candVT, S.candEE, S.candFillVT, S.candFillEE, S.candCapVT,
                                              S.candCapEE},
                               S.buf[0], S.buf[1], g.d_ctl);
            launch_self_colour(s, S.buf[0], S.buf[1], S.vmax, S.vloc, std::max(g.nv, 1), g.d_ctl);
        }
        timed_end(g.timer, timed, 3, s, ev);
    } else if (anyContacts && msActive) {
        // detection must see the spring's moves: three launches
        timed_begin(g.timer, timed, 3, s, ev);
        launch_predict(s, g.d_pos, g.d_prev, g.d_vel, g.nv, h, gdv, mDragCoeff, g.d_ctl);
        launch_mouse_spring(s, g.d_pos, g.d_vel, g.d_mouse, h, g.d_ctl, 0);
        launch_detect(s, g.d_pos, g.d_prev, g.nv, cs, cb, mThickness, mTolerance, mCollFlags, g.d_ctl);
        timed_end(g.timer, timed, 3, s, ev);
    } else if (anyContacts) {
        timed_begin(g.timer, timed, 3, s, ev);
        launch_predict_detect(s, g.d_pos, g.d_prev, g.d_vel, g.nv, h, gdv, mDragCoeff, g.d_ctl, cs, cb, mThickness,
                              mTolerance, mCollFlags);
        timed_end(g.timer, timed, 3, s, ev);
    } else {
        timed_begin(g.timer, timed, 4, s, ev);
        launch_predict(s, g.d_pos, g.d_prev, g.d_vel, g.nv, h, gdv, mDragCoeff, g.d_ctl);
        if (msActive) launch_mouse_spring(s, g.d_pos, g.d_vel, g.d_mouse, h, g.d_ctl, 0);
        timed_end(g.timer, timed, 4, s, ev);
    }

    const float h2 = h * h;
    const float epsUV = 2 * (1 + mPoisson) / (smax(mShearStiff, 0.1f) * mUnit * mUnit);
    const float sStiff = 0.0005f * mBendStiff;
    TriStreams tr{g.d_tidx, g.d_tA, g.d_tB, g.d_tC};
    // learned warm start: features at the inertial positions, then (solve
    // mode) the predicted correction added before the sweeps
    const bool track = mWarm && mWarmMode != 0 && g.warm.built;
    const int wmode = mWarmPlainNow ? 0 : mWarmMode;
    const WarmBufs wb = warm_bufs();
    if (track) {
        launch_ws_tri_strain(s, g.d_pos, tr, g.nt, g.warm.triStrain);
        launch_ws_features(s, g.d_pos, g.d_prev, wb, g.nv, 1.f / g.warm.ell);
        if (wmode == 2) launch_mlp_apply(s, mWarm->weights(), g.warm.X, g.nv, g.d_pos, g.warm.ell / kWsScaleOut);
    }
    const int iterations = wmode == 2 ? mWarmIters : mNumIterations;
    // ClothPBD.cpp:9-50: k = h^2 area unit, evaluated in the reference's order
    const float linkK = h2 * mArea * mUnit;
    // with a fixed count, the reported errors are the last passes' (the exit
    // test's kernel clears them per iteration otherwise)
    const int lastBend = mBendFreq > 0 ? ((iterations - 1) / mBendFreq) * mBendFreq : 0;
    // accumulators that launches of this substep may have written since the
    // substep cleared them (bit 0: the per-iteration errors, bit 1: the bend
    // error); a fixed count clears only what earlier passes wrote
    int dirty = 0;
    bool velocityDone = false;  // set by FEM tiles of the last iteration
    if (mSolver != SOLVER_GAUSS_SEIDEL) {
        // ClothPBD.cpp:290-298: Jacobi or conjugate-residual passes over the
        // links only (no triangles, bends or contacts), a fixed count
        const JacobiLinks J{g.jac.lk, g.jac.lr, g.jac.disp, g.jac.d, g.jac.off, g.jac.inc, g.jac.nl, g.jac.nv};
        timed_begin(g.timer, timed, 0, s, ev);
        if (mSolver == SOLVER_JACOBI) {
            for (int k = 0; k < mNumIterations; k++) launch_jacobi_pass(s, g.d_pos, J, linkK, 0.5f);
        } else {
            launch_cr_reset(s, J);
            for (int k = 0; k < mNumIterations; k++) {
                // ClothPBD.cpp:442-443 (SKIP_BETA): pow on the host, as the reference
                float beta = (float)k / (float)(mNumIterations - 1);
                beta = std::pow(beta, 0.6f);
                launch_cr_pass(s, g.d_pos, J, mUnit, beta, 0.1f);
            }
        }
        timed_end(g.timer, timed, 0, s, ev);
    }
    for (int k = 0; k < (mSolver == SOLVER_GAUSS_SEIDEL ? iterations : 0); k++) {
        if (mFixedIterations) {
            const int mask =
                ((k == iterations - 1 && k > 0 ? 1 : 0) | (mDihedral && k == lastBend && k > 0 ? 2 : 0)) & dirty;
            if (mask) launch_reset_errors(s, g.d_ctl, mask);
            dirty &= ~mask;
        }
        if (!mFEM) {
            for (int p = 0; p < linkCol.numPhases(); p++) {
                timed_begin(g.timer, timed, 0, s, ev);
                if (linkCol.slabbed) {
                    const int b = linkCol.phaseSlabBegin[p], ns = linkCol.phaseSlabBegin[p + 1] - b;
                    SlabSet ss{g.linkSlabs.info, g.linkSlabs.colOff, g.linkSlabs.first, g.linkSlabs.verts, b,
                               (int)linkCol.order.size()};
                    launch_link_slab(s, g.d_pos, g.linkSlabs.lidx, g.d_lr, ss, ns, linkCol.maxSlabVerts, linkK,
                                     g.d_ctl);
                    dirty |= 1;
                } else {
                    const int st = linkCol.phaseOffsets[p], cnt = linkCol.phaseOffsets[p + 1] - st;
                    launch_links(s, g.d_pos, g.d_lk, g.d_lr, st, cnt, linkK, g.d_ctl);
                    dirty |= 1;
                }
                timed_end(g.timer, timed, 0, s, ev);
            }
        }
        const bool bendPass = mDihedral && k % mBendFreq == 0;
        // the error maxima that survive to the substep's report (a fixed
        // count clears them before the last pass of each kind)
        const int report = !mFixedIterations ? 3 : (k == iterations - 1 ? 1 : 0) | (k >= lastBend ? 2 : 0);
        const bool contactsInFem = fuseTiles && !bendPass;
        // with a bend pass, each vertex's last bend slab applies its contacts
        const bool contactsInBends = fuseTiles && bendPass && bendCol.slabbed && g.bendOrphans == 0;
        if (mFEM) {
            for (int p = 0; p < triCol.numPhases(); p++) {
                timed_begin(g.timer, timed, 0, s, ev);
                if (triCol.slabbed) {
                    const int b = triCol.phaseSlabBegin[p], ns = triCol.phaseSlabBegin[p + 1] - b;
                    const int positions = triCol.tiled ? (int)triCol.instCons.size() : (int)triCol.order.size();
                    SlabSet ss{g.triSlabs.info, g.triSlabs.colOff, g.triSlabs.first, g.triSlabs.verts, b, positions};
                    SlabContacts kc{cb, g.d_prev, g.d_vel, mThickness, mFriction, invH,
                                    (fuse && k > 0) ? 1 : contactsInFem ? 2 : 0, report};
                    const TriStreams fs = triCol.tiled ? TriStreams{g.d_tidx, g.d_ttA, g.d_ttB, g.d_ttC} : tr;
                    if ((int)g.triPacked.phases.size() == triCol.numPhases() && kc.enabled != 1) {
                        const DevicePacked::Phase& ph = g.triPacked.phases[p];
                        const PackedFem pf{g.triPacked.verts + ph.vertOff, g.triPacked.s0 + ph.slotOff,
                                           g.triPacked.s1 + ph.slotOff, ph.nc};
                        // the last iteration's tiles also set the velocities when
                        // every vertex is some tile's own
                        if (kc.enabled == 2 && k == iterations - 1 && triCol.numPhases() == 1 && g.triOrphans == 0) {
                            kc.velocity = 1;
                            velocityDone = true;
                        }
                        launch_fem_packed(s, g.d_pos, pf, ph.slabs, ph.ncKernel, ph.kv, h2, epsUV, g.d_ctl, kc);
                        if (kc.enabled == 2) dirty |= report & 1;
                    } else {
                        launch_fem_slab(s, g.d_pos, fs, g.triSlabs.lidx, ss, ns, triCol.maxSlabVerts, h2, epsUV,
                                        g.d_ctl, kc, triCol.tiled);
                        if (kc.enabled) dirty |= 1;
                    }
                } else {
                    int st = triCol.phaseOffsets[p], cnt = triCol.phaseOffsets[p + 1] - st;
                    launch_fem(s, g.d_pos, tr, st, cnt, h2, epsUV, g.d_ctl);
                }
                timed_end(g.timer, timed, 0, s, ev);
            }
        }
        if (bendPass) {
            for (int p = 0; p < bendCol.numPhases(); p++) {
                timed_begin(g.timer, timed, 1, s, ev);
                if (bendCol.slabbed) {
                    const int b = bendCol.phaseSlabBegin[p], ns = bendCol.phaseSlabBegin[p + 1] - b;
                    SlabSet ss{g.bendSlabs.info, g.bendSlabs.colOff, g.bendSlabs.first, g.bendSlabs.verts, b,
                               (int)bendCol.order.size()};
                    SlabContacts kc{cb, g.d_prev, g.d_vel, mThickness, mFriction, invH, contactsInBends ? 3 : 0,
                                    report};
                    // (slabs wider than kPackMaxKv slots apply no contacts)
                    bool packed = (int)g.bendPacked.phases.size() == bendCol.numPhases();
                    for (const DevicePacked::Phase& ph : g.bendPacked.phases)
                        packed = packed && !(contactsInBends && ph.kv > kPackMaxKv);
                    if (packed) {
                        const DevicePacked::Phase& ph = g.bendPacked.phases[p];
                        const PackedBend pb{g.bendPacked.verts + ph.vertOff, g.bendPacked.ids + ph.slotOff, ph.nc,
                                            ph.width};
                        launch_bend_packed(s, g.d_pos, pb, ph.slabs, ph.kv, sStiff, g.d_ctl, kc);
                        dirty |= (report & 2) | (kc.enabled == 3 ? report & 1 : 0);
                    } else {
                        launch_bend_slab(s, g.d_pos, g.bendSlabs.lidx, ss, ns, bendCol.maxSlabVerts, sStiff, g.d_ctl,
                                         kc);
                        dirty |= kc.enabled == 3 ? 3 : 2;
                    }
                } else {
                    int st = bendCol.phaseOffsets[p], cnt = bendCol.phaseOffsets[p + 1] - st;
                    launch_bend(s, g.d_pos, g.d_bidx, st, cnt, sStiff, g.d_ctl);
                    dirty |= 2;
                }
                timed_end(g.timer, timed, 1, s, ev);
            }
        }
        if (anyContacts && !fuse && !contactsInFem && !contactsInBends) {
            timed_begin(g.timer, timed, 2, s, ev);
            launch_contacts(s, g.d_pos, g.d_prev, g.d_vel, cb, mThickness, mFriction, invH, g.d_ctl);
            dirty |= 1;
            timed_end(g.timer, timed, 2, s, ev);
        }
        if (edgeContacts) {
            timed_begin(g.timer, timed, 2, s, ev);
            launch_edge_contacts(s, g.d_pos, g.d_prev, g.d_cEdges, ecb, g.ec.base.data(), g.d_geo, mThickness,
                                 mFriction, g.d_ctl);
            dirty |= 1;
            timed_end(g.timer, timed, 2, s, ev);
        }
        if (triContacts) {
            timed_begin(g.timer, timed, 2, s, ev);
            launch_tri_contacts(s, g.d_pos, g.d_prev, g.d_cTris, tcb, g.tc.base.data(), g.d_geo, mThickness,
                                mFriction, g.d_ctl);
            dirty |= 1;
            timed_end(g.timer, timed, 2, s, ev);
        }
        if (selfColl) {
            timed_begin(g.timer, timed, 2, s, ev);
            launch_self_solve(s, g.d_pos, g.self.buf[0], 0, sp, h, g.d_ctl);
            launch_self_solve(s, g.d_pos, g.self.buf[1], 1, sp, h, g.d_ctl);
            dirty |= 1;
            timed_end(g.timer, timed, 2, s, ev);
        }
        if (!mFixedIterations)
            launch_iter_end(s, g.d_ctl, iterations, 0, mBendFreq, mDihedral ? 1 : 0, mSelfCollThreshold);
    }
    timed_begin(g.timer, timed, 4, s, ev);
    if (fuse)
        launch_contacts_velocity(s, g.d_pos, g.d_prev, g.d_vel, cb, mThickness, mFriction, invH, g.d_ctl);
    else if (!velocityDone)
        launch_velocity(s, g.d_pos, g.d_prev, g.d_vel, g.nv, invH, g.d_ctl);
    timed_end(g.timer, timed, 4, s, ev);
    if (track) {
        // the correction of this substep; in training mode, a step of Adam on it
        launch_ws_after(s, g.d_pos, wb, g.nv, wmode == 1, kWsScaleOut / g.warm.ell);
        if (wmode == 1 && g.warm.nFree > 0)
            mWarm->enqueue_train(s, g.warm.X, g.warm.T, g.nv, 1.f / (float)g.warm.nFree);
    }
}

// n substeps through one of two cached graphs (keyed by everything a
// recorded substep depends on), or by direct launches when timing.
void ClothModel::run_substeps(float h, int n, const CollSet& cs, bool timed)
{
    GpuSolver& g = gpu;
    if (mUseGraph && !timed) {
        const std::string key = graph_key(h, cs);
        hipGraphExec_t exec = nullptr;
        if (g.graphExec && key == g.graphKey) {
            exec = g.graphExec;
        } else if (g.graphExec2 && key == g.graphKey2) {
            exec = g.graphExec2;
        } else {
            // the newer graph moves to the second slot; the new one takes the first
            if (g.graphExec2) (void)hipGraphExecDestroy(g.graphExec2);
            if (g.graph2) (void)hipGraphDestroy(g.graph2);
            g.graphExec2 = g.graphExec;
            g.graph2 = g.graph;
            g.graphKey2 = g.graphKey;
            g.graphExec = nullptr;
            g.graph = nullptr;
            HIP_OK(hipStreamBeginCapture(g.stream, hipStreamCaptureModeThreadLocal));
            record_substep(h, cs, false);
            HIP_OK(hipStreamEndCapture(g.stream, &g.graph));
            HIP_OK(hipGraphInstantiate(&g.graphExec, g.graph, nullptr, nullptr, 0));
            g.graphKey = key;
            exec = g.graphExec;
        }
        if (!g.ev0) {
            HIP_OK(hipEventCreate(&g.ev0));
            HIP_OK(hipEventCreate(&g.ev1));
        }
        HIP_OK(hipEventRecord(g.ev0, g.stream));
        for (int i = 0; i < n; i++) HIP_OK(hipGraphLaunch(exec, g.stream));
        HIP_OK(hipEventRecord(g.ev1, g.stream));
        HIP_OK(hipStreamSynchronize(g.stream));
        float ms = 0.f;
        HIP_OK(hipEventElapsedTime(&ms, g.ev0, g.ev1));
        mLastStepMs += ms;
    } else {
        for (int i = 0; i < n; i++) record_substep(h, cs, timed);
    }
    HIP_OK(hipStreamSynchronize(g.stream));
    HIP_OK(hipGetLastError());
}

void ClothModel::step_substeps(float h, int n)
{
    if (mNumIterations < 1) throw std::runtime_error("NumIterations must be >= 1");
    if (mBendFreq < 1) throw std::runtime_error("bend frequency must be >= 1");
    const CollisionWorld* w = mOwner->mWorld;
    if (w->colls.size() > MAX_COLLIDABLES) throw std::runtime_error("too many collidables (max 8)");
    for (const auto& c : w->colls) {
        if (c->type != CT_WALLS && c->type != CT_SPHERE && c->type != CT_SDF && c->type != CT_MESH &&
            c->type != CT_CAPSULE)
            throw std::runtime_error("collidable type " + std::to_string(c->type) +
                                     " is not supported by the GPU substep");
        if (c->type == CT_SDF && (mCollFlags & CF_EDGES) && !(mCollFlags & CF_VERTICES))
            throw std::runtime_error("edge contacts against a grid SDF need CF_VERTICES "
                                     "(the reference's grid-SDF segment query is a stub)");
    }
    sync_to_device();
    if (pos.empty()) return;
    for (const auto& c : mOwner->mWorld->colls)
        if (c->type == CT_SDF && !((CollisionSDF*)c.get())->sdf->d_nodes) ((CollisionSDF*)c.get())->sdf->upload();
    int slots = 0, edgeSlots = 0, triSlots = 0;
    CollSet cs = make_collset(slots, edgeSlots, triSlots);
    ensure_contact_capacity(std::max(slots, 1));
    gpu.activeSlots = slots;
    ensure_feature_capacity(edgeSlots, triSlots);
    if (mCollFlags & CF_SELF) {
        ensure_self();
        clear_self_counters();
    }
    GpuSolver& g = gpu;
    const bool timed = g.timer.enabled;
    if (mWarm && mWarmMode != 0) ensure_warm();
    if (mSolver != SOLVER_GAUSS_SEIDEL) {
        if (mWarm && mWarmMode != 0)
            throw std::runtime_error("the learned warm start works with the Gauss-Seidel solver only");
        ensure_jacobi();
    }
    if (msActive) {
        const int nv = (int)pos.size();
        for (int k = 0; k < 3; k++)
            if (msI[k] < 0 || msI[k] >= nv) throw std::runtime_error("mouse spring particle index out of range");
        if (!g.d_mouse) {
            dev_alloc(g.d_mouse, 1);
            msDirty = true;
        }
        if (msDirty) {
            const MouseSpringDev d{msI[0], msI[1], msI[2], msA[0], msA[1], msA[2], msPoint.x, msPoint.y, msPoint.z,
                                   msLen, msStiffness};
            HIP_OK(hipMemcpy(g.d_mouse, &d, sizeof(d), hipMemcpyHostToDevice));
            msDirty = false;
        }
    }
    // the plain solver judges a warm-started step: a step whose residual
    // exceeds the last plain step's (or the set threshold) is rerun plain
    // from the saved state, and the first step calibrates
    // a threshold of +infinity: no judge (the warm-started step is kept as is)
    const bool judge = mWarm && mWarmMode == 2 && !std::isinf(mWarmThreshold);
    // continuous self collision snapshots the step's start: a step whose
    // contacts or candidate pairs outgrew their buffers is rerun from the
    // snapshot with the buffers grown to the step's peaks, so that no step
    // advances with an incomplete contact set
    const bool selfStep = (mCollFlags & CF_SELF) != 0;
    if (selfStep) snapshot_step();
    mLastStepMs = 0.0;  // a rerun's attempts are all charged
    for (int attempt = 0;; attempt++) {
        if (selfStep && attempt > 0) {
            restore_step();
            ensure_self();
            clear_self_counters();
        }
        if (judge) save_state();
        mWarmPlainNow = false;
        if (judge && mWarmThreshold < 0.f) {
            // the plain step from the same start judges the warm-started one:
            // its residual is the bound and its result is kept when the
            // warm-started step leaves a larger residual
            mWarmPlainNow = true;
            run_substeps(h, n, cs, timed);
            mWarmPlainNow = false;
            const double rp = residual();
            keep_plain_result();
            restore_state();
            run_substeps(h, n, cs, timed);
            const double rw = residual();
            mWarmSteps++;
            mWarmPlainResidual = rp;
            mWarmWarmResidual = rw;
            if (!(rw <= rp)) {
                restore_plain_result();
                mWarmFallbacks++;
                mWarmLastPlain = true;
                mWarmResidual = rp;
            } else {
                mWarmLastPlain = false;
                mWarmResidual = rw;
            }
        } else {
            run_substeps(h, n, cs, timed);
            mWarmLastPlain = !judge;
            if (judge) {
                // a fixed bound: a step above it is rerun plain
                const double r = residual();
                mWarmSteps++;
                mWarmResidual = r;
                mWarmWarmResidual = r;
                mWarmPlainResidual = -1.0;
                if (!(r <= (double)mWarmThreshold)) {
                    mWarmFallbacks++;
                    restore_state();
                    mWarmPlainNow = true;
                    run_substeps(h, n, cs, timed);
                    mWarmPlainNow = false;
                    mWarmLastPlain = true;
                    mWarmResidual = mWarmPlainResidual = residual();
                }
            }
        }
        if (timed) g.timer.harvest();
        if (!selfStep || !self_step_overflowed(attempt)) break;
    }
    mDeviceNewer = true;
    mStatsPending = true;
    contactsValid = false;
}

// Copies of the state a step starts from (positions, previous positions,
// velocities and the warm start's last correction).
void ClothModel::snapshot_step()
{
    GpuSolver& g = gpu;
    GpuSolver::StepSnap& S = g.snap;
    if (S.nv != g.nv) {
        dev_alloc(S.pos, std::max(g.nv, 1));
        dev_alloc(S.prev, std::max(g.nv, 1));
        dev_alloc(S.vel, std::max(g.nv, 1));
        dev_alloc(S.corr, std::max(g.nv, 1));
        S.nv = g.nv;
    }
    launch_copy(g.stream, g.d_pos, S.pos, g.nv);
    launch_copy(g.stream, g.d_prev, S.prev, g.nv);
    launch_copy(g.stream, g.d_vel, S.vel, g.nv);
    S.hasCorr = g.warm.built && g.warm.corr && g.warm.nv == g.nv;
    if (S.hasCorr) launch_copy(g.stream, g.warm.corr, S.corr, g.nv);
}

void ClothModel::restore_step()
{
    GpuSolver& g = gpu;
    GpuSolver::StepSnap& S = g.snap;
    launch_copy(g.stream, S.pos, g.d_pos, g.nv);
    launch_copy(g.stream, S.prev, g.d_prev, g.nv);
    launch_copy(g.stream, S.vel, g.d_vel, g.nv);
    if (S.hasCorr) launch_copy(g.stream, S.corr, g.warm.corr, g.nv);
    HIP_OK(hipStreamSynchronize(g.stream));
}

void ClothModel::clear_self_counters()
{
    GpuSolver& g = gpu;
    HIP_OK(hipMemsetAsync(&g.d_ctl->selfOverflow, 0, sizeof(int), g.stream));
    HIP_OK(hipMemsetAsync(g.d_ctl->selfPeak, 0, sizeof(g.d_ctl->selfPeak), g.stream));
}

// After a self-colliding step: grow the candidate pools and contact buffers
// from the step's peaks. Returns true when the step overflowed and must be
// rerun; throws (with the state of the step's start restored) when the rerun
// cannot help.
bool ClothModel::self_step_overflowed(int attempt)
{
    GpuSolver& g = gpu;
    GpuSolver::Self& S = g.self;
    SolverCtl c;
    HIP_OK(hipMemcpy(&c, g.d_ctl, sizeof(c), hipMemcpyDeviceToHost));
    const long peak[4] = {std::max(c.selfPeak[0], c.selfCount[0]), std::max(c.selfPeak[1], c.selfCount[1]),
                          std::max(c.selfPeak[2], c.candSlots[0]), std::max(c.selfPeak[3], c.candSlots[1])};
    lastSelfPeak[0] = (int)peak[0];
    lastSelfPeak[1] = (int)peak[1];
    lastSelfColours[0] = c.selfColours[0];
    lastSelfColours[1] = c.selfColours[1];
    lastSelfUniq[0] = c.selfUniq[0];
    lastSelfUniq[1] = c.selfUniq[1];
    const long limit = 1L << 30;
    if (peak[2] > S.candCapVT / 5 * 4) S.candWantVT = std::min(limit, peak[2] + peak[2] / 2);
    if (peak[3] > S.candCapEE / 5 * 4) S.candWantEE = std::min(limit, peak[3] + peak[3] / 2);
    if (!c.selfOverflow) return false;
    selfReruns++;
    const long need = std::max(peak[0], peak[1]);
    const bool contactsFull = need > S.cap;
    const bool poolsFull = peak[2] > S.candCapVT || peak[3] > S.candCapEE;
    if (contactsFull) mSelfCapacityGrown = (int)std::min(limit, std::max(2L * S.cap, need + need / 2));
    if ((!contactsFull && !poolsFull) || attempt >= 8 || need >= limit) {
        restore_step();
        after_device_step();
        throw std::runtime_error(
            "self collision overflowed and the step was not taken: " + std::to_string(peak[0]) +
            " vertex-triangle and " + std::to_string(peak[1]) + " edge-edge contacts (capacity " +
            std::to_string(S.cap) + "), " + std::to_string(peak[2]) + " / " + std::to_string(peak[3]) +
            " candidate slots (pools " + std::to_string(S.candCapVT) + " / " + std::to_string(S.candCapEE) + "), " +
            std::to_string(c.selfColours[0]) + " / " + std::to_string(c.selfColours[1]) + " colour groups after " +
            std::to_string(attempt + 1) + " attempts");
    }
    return true;
}

// The last substep's broad-phase candidate pairs of one kind (0: (triangle,
// vertex), 1: (edge1, edge2)): the filled slots of the pool's chunks,
// sorted. Empty before the first self-colliding step.
std::vector<std::pair<int, int>> ClothModel::self_candidates(int kind)
{
    std::vector<std::pair<int, int>> out;
    GpuSolver& g = gpu;
    GpuSolver::Self& S = g.self;
    if (!(mCollFlags & CF_SELF) || !S.candVT || !g.d_ctl) return out;
    HIP_OK(hipStreamSynchronize(g.stream));
    SolverCtl c;
    HIP_OK(hipMemcpy(&c, g.d_ctl, sizeof(c), hipMemcpyDeviceToHost));
    const int cap = kind == 0 ? S.candCapVT : S.candCapEE;
    const int slots = std::min(c.candSlots[kind], cap);
    if (slots <= 0) return out;
    std::vector<int> fill(slots / kCandChunk);
    std::vector<int2> pool(slots);
    HIP_OK(hipMemcpy(fill.data(), kind == 0 ? S.candFillVT : S.candFillEE, fill.size() * sizeof(int),
                     hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(pool.data(), kind == 0 ? S.candVT : S.candEE, pool.size() * sizeof(int2), hipMemcpyDeviceToHost));
    for (size_t ch = 0; ch < fill.size(); ch++)
        for (int k = 0; k < fill[ch]; k++) out.push_back({pool[ch * kCandChunk + k].x, pool[ch * kCandChunk + k].y});
    std::sort(out.begin(), out.end());
    return out;
}

void ClothModel::fetch_stats()
{
    if (!mStatsPending || !gpu.d_ctl) return;
    SolverCtl c;
    HIP_OK(hipMemcpy(&c, gpu.d_ctl, sizeof(c), hipMemcpyDeviceToHost));
    if (mFixedIterations) {
        statIterations = (mWarm && mWarmMode == 2 && !mWarmLastPlain) ? mWarmIters : mNumIterations;
        unsigned cb = c.cb, cv = c.cv;
        for (int t = 0; t < kErrParts; t++) {
            cb = std::max(cb, c.errPart[0][t * kErrStride]);
            cv = std::max(cv, c.errPart[1][t * kErrStride]);
        }
        statBendError = __builtin_bit_cast(float, cb);
        statContactError = __builtin_bit_cast(float, cv);
    } else {
        statIterations = c.iterations;
        statBendError = c.bendError;
        statContactError = c.contactError;
    }
    statSelfVTError = __builtin_bit_cast(float, c.sct);
    statSelfEEError = __builtin_bit_cast(float, c.sce);
    statLinkError = mFEM ? 0.f : __builtin_bit_cast(float, c.cl);
    mStatsPending = false;
}

// Contacts of the last substep in the reference's list order: collidable
// order, then vertex, then the wall planes' order (ClothCollision.cpp:199-220).
void ClothModel::download_contacts()
{
    if (contactsValid) return;
    contactsHost.clear();
    edgeContactsHost.clear();
    triContactsHost.clear();
    GpuSolver& g = gpu;
    if (mTopoDirty) { contactsValid = true; return; }
    if (g.d_cP && g.nv > 0 && g.slots > 0 && g.activeSlots > 0) {
        std::vector<int> cnt(g.nv), src((size_t)g.slots * g.nv), feat((size_t)g.slots * g.nv);
        std::vector<float4> P((size_t)g.slots * g.nv), N((size_t)g.slots * g.nv);
        HIP_OK(hipMemcpy(cnt.data(), g.d_cCount, g.nv * sizeof(int), hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(P.data(), g.d_cP, P.size() * sizeof(float4), hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(N.data(), g.d_cN, N.size() * sizeof(float4), hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(src.data(), g.d_cSrc, src.size() * sizeof(int), hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(feat.data(), g.d_cFeat, feat.size() * sizeof(int), hipMemcpyDeviceToHost));
        for (int v = 0; v < g.nv; v++) {
            for (int s = 0; s < cnt[v] && s < g.slots; s++) {
                const size_t k = (size_t)s * g.nv + v;
                ContactOut c;
                c.vertex = v;
                c.slot = s;
                c.coll = src[k];
                c.feature = feat[k];
                c.point = f3{P[k].x, P[k].y, P[k].z};
                c.normal = f3{N[k].x, N[k].y, N[k].z};
                contactsHost.push_back(c);
            }
        }
        std::stable_sort(contactsHost.begin(), contactsHost.end(), [](const ContactOut& a, const ContactOut& b) {
            if (a.coll != b.coll) return a.coll < b.coll;
            if (a.vertex != b.vertex) return a.vertex < b.vertex;
            return a.slot < b.slot;
        });
    }
    // obstacle-edge and obstacle-vertex contacts in the reference's list
    // order: collidable, then cloth edge / triangle
    auto features = [&](GpuSolver::Features& f, std::vector<FeatureContactOut>& out) {
        if (!f.P || f.n == 0 || f.slots == 0) return;
        const size_t m = (size_t)f.slots * f.n;
        std::vector<int> cnt(f.n);
        std::vector<float4> P(m), N(m), W(m);
        std::vector<int2> F(m);
        HIP_OK(hipMemcpy(cnt.data(), f.count, f.n * sizeof(int), hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(P.data(), f.P, m * sizeof(float4), hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(N.data(), f.N, m * sizeof(float4), hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(W.data(), f.W, m * sizeof(float4), hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(F.data(), f.F, m * sizeof(int2), hipMemcpyDeviceToHost));
        for (int i = 0; i < f.n; i++) {
            for (int s = 0; s < cnt[i] && s < f.slots; s++) {
                const size_t k = (size_t)s * f.n + i;
                FeatureContactOut c;
                c.id = i;
                c.feature = F[k].x;
                c.coll = F[k].y;
                c.point = f3{P[k].x, P[k].y, P[k].z};
                c.normal = f3{N[k].x, N[k].y, N[k].z};
                c.w[0] = W[k].x; c.w[1] = W[k].y; c.w[2] = W[k].z;
                c.lambda = P[k].w;
                c.lambdaF = N[k].w;
                out.push_back(c);
            }
        }
        std::stable_sort(out.begin(), out.end(), [](const FeatureContactOut& a, const FeatureContactOut& b) {
            if (a.coll != b.coll) return a.coll < b.coll;
            return a.id < b.id;
        });
    };
    if (mCollFlags & CF_EDGES) features(g.ec, edgeContactsHost);
    if (mCollFlags & CF_TRIANGLES) features(g.tc, triContactsHost);
    selfContactsHost[0].clear();
    selfContactsHost[1].clear();
    if ((mCollFlags & CF_SELF) && g.self.cap > 0 && g.d_ctl) {
        SolverCtl c;
        HIP_OK(hipMemcpy(&c, g.d_ctl, sizeof(c), hipMemcpyDeviceToHost));
        for (int kind = 0; kind < 2; kind++) {
            const SelfBuf& b = g.self.buf[kind];
            const int n = std::min(c.selfCount[kind], b.cap);
            const int nc = c.selfColours[kind];
            if (n <= 0 || nc <= 0) continue;
            std::vector<int4> ids(n);
            std::vector<float4> W(n), N(n);
            std::vector<unsigned long long> key(n);
            std::vector<int> order(n), off(nc + 1);
            HIP_OK(hipMemcpy(ids.data(), b.ids, n * sizeof(int4), hipMemcpyDeviceToHost));
            HIP_OK(hipMemcpy(W.data(), b.W, n * sizeof(float4), hipMemcpyDeviceToHost));
            HIP_OK(hipMemcpy(N.data(), b.N, n * sizeof(float4), hipMemcpyDeviceToHost));
            HIP_OK(hipMemcpy(key.data(), b.key, n * sizeof(unsigned long long), hipMemcpyDeviceToHost));
            HIP_OK(hipMemcpy(off.data(), b.groupOff, (nc + 1) * sizeof(int), hipMemcpyDeviceToHost));
            HIP_OK(hipMemcpy(order.data(), b.order, off[nc] * sizeof(int), hipMemcpyDeviceToHost));
            // colour groups in order; inside a group by key (members commute)
            for (int grp = 0; grp < nc; grp++) {
                std::vector<int> m(order.begin() + off[grp], order.begin() + off[grp + 1]);
                std::sort(m.begin(), m.end(), [&](int a, int bb) { return key[a] < key[bb]; });
                for (int i : m) {
                    SelfContactOut o;
                    o.key = key[i];
                    o.ids[0] = ids[i].x; o.ids[1] = ids[i].y; o.ids[2] = ids[i].z; o.ids[3] = ids[i].w;
                    o.w[0] = W[i].x; o.w[1] = W[i].y; o.w[2] = W[i].z;
                    o.side = W[i].w;
                    o.normal[0] = N[i].x; o.normal[1] = N[i].y; o.normal[2] = N[i].z;
                    o.lambda = N[i].w;
                    selfContactsHost[kind].push_back(o);
                }
            }
        }
    }
    contactsValid = true;
}

// -------------------------------------------------------------- patch ----

// Physics/ClothPatch.cpp:135-246 (grid branch)
// Physics/ClothPatch.cpp:43-133. Vertices at the same position are welded
// into one particle (the reference's quadratic search, kept for its exact
// test: squared distance 0); each triangle corner takes the mesh's material
// coordinates; links, edges and bends come from the sorted corner edges, in
// which the closing edge of the last triangle is missing as in the reference
// (its loop over corners stops one short).
void ClothPatch::init_mesh(const Mesh& mesh, f3 offset, bool attached)
{
    if (mesh.vertices.empty() || mesh.indices.empty()) return;
    const size_t n = mesh.indices.size();
    if (n % 3 != 0) throw std::runtime_error("mesh index count is not a multiple of 3");
    for (unsigned v : mesh.indices)
        if (v >= mesh.vertices.size()) throw std::runtime_error("mesh index out of range");
    if (!mesh.uvs.empty() && mesh.uvs.size() != mesh.vertices.size())
        throw std::runtime_error("mesh uv count differs from its vertex count");
    if (!mesh.colors.empty() && mesh.colors.size() != mesh.vertices.size())
        throw std::runtime_error("mesh colour count differs from its vertex count");
    ClothModel& m = *mModel;
    mMesh.clear();
    mMesh.vertices = mesh.vertices;
    mMesh.indices = mesh.indices;
    mMesh.uvs = mesh.uvs;
    mMesh.colors = mesh.colors;
    m.clear();
    const size_t nv = mesh.vertices.size();
    mMap.assign(nv, -1);
    auto V = [&](size_t i) { return f3{mesh.vertices[i].x, mesh.vertices[i].y, mesh.vertices[i].z}; };
    for (size_t i = 0; i < nv; i++) {
        if (mMap[i] >= 0) continue;
        const int np = (int)m.pos.size();
        for (size_t j = i + 1; j < nv; j++) {
            const f3 d = sub(V(i), V(j));
            if (dot(d, d) == 0.f) mMap[j] = np;
        }
        mMap[i] = np;
        const f3 p = add(V(i), offset);
        m.pos.push_back(p);
        m.prev.push_back(p);
        m.vel.push_back(f3{0.f, 0.f, 0.f});
        m.invMass.push_back((attached && !mesh.colors.empty() && mesh.colors[i].x != 0.f) ? 0.f : 1.f);
        m.uv.push_back(mesh.uvs.empty() ? scVector2{0.f, 0.f} : mesh.uvs[i]);
    }
    const bool uvs = !mesh.uvs.empty();
    auto tri = [&](size_t a, size_t b, size_t c, uint32_t i1, uint32_t i2, uint32_t i3) {
        if (uvs) m.add_triangle_uv(i1, i2, i3, mesh.uvs[a], mesh.uvs[b], mesh.uvs[c]);
        else m.add_triangle(i1, i2, i3);
    };
    const std::vector<unsigned>& idx = mesh.indices;
    std::vector<MeshEdge> dup(n - 1);
    for (size_t i = 0; i + 1 < n; i++) {
        uint32_t i1 = (uint32_t)mMap[idx[i]], i2 = (uint32_t)mMap[idx[i + 1]];
        if ((i + 1) % 3 == 0) {
            i2 = (uint32_t)mMap[idx[i - 2]];
            tri(idx[i - 2], idx[i - 1], idx[i], i2, (uint32_t)mMap[idx[i - 1]], i1);
        }
        if (i1 > i2) std::swap(i1, i2);
        MeshEdge e;
        e.i1 = i1;
        e.i2 = i2;
        e.t1 = (int)(i / 3);
        dup[i] = e;
    }
    tri(idx[n - 3], idx[n - 2], idx[n - 1], (uint32_t)mMap[idx[n - 3]], (uint32_t)mMap[idx[n - 2]],
        (uint32_t)mMap[idx[n - 1]]);
    std::sort(dup.begin(), dup.end(), [](const MeshEdge& a, const MeshEdge& b) {
        return a.i1 < b.i1 || (a.i1 == b.i1 && a.i2 < b.i2);
    });
    const float stretch = m.mStretchStiff;
    if (!dup.empty()) {
        m.add_link(dup[0].i1, dup[0].i2, stretch);
        MeshEdge cur = dup[0];
        for (size_t i = 1; i < dup.size(); i++) {
            if (dup[i].i1 == dup[i - 1].i1 && dup[i].i2 == dup[i - 1].i2) {
                cur.t2 = dup[i].t1;
                continue;
            }
            m.add_link(dup[i].i1, dup[i].i2, stretch);
            mMesh.edges.push_back(cur);
            m.edges.push_back(cur);
            m.add_bend(cur);
            cur = dup[i];
        }
        mMesh.edges.push_back(cur);
        m.edges.push_back(cur);
        m.add_bend(cur);
    }
    m.mark_topology_dirty();
}

void ClothPatch::init_grid(int divX, int divY, float inc, f3 offset, bool horizontal, bool attached)
{
    mMap.clear();
    ClothModel& m = *mModel;
    mMesh.clear();
    m.clear();
    const int n = divX * divY;
    m.pos.assign(n, f3{0.f, 0.f, 0.f});
    m.prev = m.pos;
    m.vel = m.pos;
    m.invMass.assign(n, 1.f);
    m.uv.assign(n, scVector2{0.f, 0.f});
    const float scale = 1.f;
    float x = -0.5f * divX * inc;
    const float startY = x;
    const float incUVX = 1.f / divX;
    const float incUVY = 1.f / divY;
    const float stretch = m.mStretchStiff, shear = m.mShearStiff, bend = m.mBendStiff;
    const float incline = mIncline;
    mMesh.vertices.resize(n);
    mMesh.uvs.resize(n);
    for (int j = 0, base = 0; j < divX; j++, base += divY) {
        f3 p0 = horizontal ? f3{x, 0.f, startY} : f3{x, startY, 0.f};
        m.pos[base] = mul(scale, add(p0, offset));
        m.prev[base] = m.pos[base];
        m.vel[base] = f3{0.f, 0.f, 0.f};
        m.invMass[base] = 1.f;
        m.uv[base] = scVector2{x, 0.f};
        mMesh.vertices[base] = scVector3{m.pos[base].x, m.pos[base].y, m.pos[base].z};
        mMesh.uvs[base] = scVector2{j * incUVX, 0.f};
        if (base >= divY) {
            m.add_link(base - divY, base, stretch);
            m.add_link(base - divY + 1, base, shear);
            if (base >= 2 * divY && !m.mDihedral) m.add_link(base - 2 * divY, base, bend);
        }
        for (int i = 1; i < divY; i++) {
            const int idx = base + i;
            f3 pi = horizontal ? f3{x, 0.f, startY + inc * i} : f3{x, startY + inc * i, i * incline};
            m.pos[idx] = mul(scale, add(pi, offset));
            m.prev[idx] = m.pos[idx];
            m.vel[idx] = f3{0.f, 0.f, 0.f};
            m.invMass[idx] = 1.f;
            m.uv[idx] = scVector2{x, inc * i};
            mMesh.vertices[idx] = scVector3{m.pos[idx].x, m.pos[idx].y, m.pos[idx].z};
            mMesh.uvs[idx] = scVector2{j * incUVX, i * incUVY};
            m.add_link(idx - 1, idx, stretch);
            if (idx > base + 1 && !m.mDihedral) m.add_link(idx - 2, idx, bend);
            if (idx >= divY) {
                m.add_link(idx - divY, idx, stretch);
                if (i < divY - 1) m.add_link(idx - divY + 1, idx, shear);
                if (j > 1 && !m.mDihedral) m.add_link(idx - 2 * divY, idx, bend);
            }
            if (idx > divY) {
                m.add_link(idx - divY - 1, idx, shear);
                m.quads.push_back(Quad{(uint32_t)(idx - divY - 1), (uint32_t)(idx - divY), (uint32_t)idx,
                                       (uint32_t)(idx - 1)});
                uint32_t t[6];
                if ((i & 1) == (j & 1)) {
                    t[0] = idx - divY - 1; t[1] = idx; t[2] = idx - divY;
                    t[3] = idx - divY - 1; t[4] = idx - 1; t[5] = idx;
                } else {
                    t[0] = idx - divY; t[1] = idx - 1; t[2] = idx;
                    t[3] = idx - 1; t[4] = idx - divY; t[5] = idx - divY - 1;
                }
                m.add_triangle(t[0], t[1], t[2]);
                m.add_triangle(t[3], t[4], t[5]);
                for (int q = 0; q < 6; q++) mMesh.indices.push_back(t[q]);
            }
        }
        if (attached && (j == 0 || j == divX - 1)) m.invMass[base + divY - 1] = 0.f;
        x += inc;
    }
    mMesh.compute_normals();
    mMesh.construct_edges();
    for (const MeshEdge& e : mMesh.edges) {
        m.edges.push_back(e);
        m.add_bend(e);
    }
    m.mark_topology_dirty();
}

// Physics/ClothPatch.cpp:323-335
void ClothPatch::step(float dt)
{
    const int n = mNumSteps;
    const float h = dt / n;
    mModel->step_substeps(h, n);
}

// Physics/ClothPatch.cpp:266-273 with ClothModel::UpdateMesh (ClothModel.inl:123-130)
// Physics/ClothPatch.cpp:266-273: the render mesh from the current state,
// computed on the device (launch_render_mesh) and copied into the mesh's
// vertex, normal and face-normal buffers.
void ClothPatch::update_mesh()
{
    if (mMesh.vertices.empty()) return;
    mModel->render_mesh(mMesh, mPosition, mMap);
}

// The links in the reference's order with each particle's incident links,
// for SolveJacobi / SolveCR
void ClothModel::ensure_jacobi()
{
    GpuSolver::Jacobi& J = gpu.jac;
    const int nl = (int)links.size(), nv = (int)pos.size();
    if (J.nl == nl && J.nv == nv) return;
    dev_free(J.lk); dev_free(J.lr); dev_free(J.disp); dev_free(J.d); dev_free(J.off); dev_free(J.inc);
    std::vector<int2> lk(std::max(1, nl));
    std::vector<float2> lr(std::max(1, nl));
    std::vector<int> off(nv + 1, 0), inc(std::max(1, 2 * nl));
    for (int l = 0; l < nl; l++) {
        lk[l] = make_int2((int)links[l].i1, (int)links[l].i2);
        lr[l] = make_float2(links[l].len, links[l].stiffness);
        off[links[l].i1 + 1]++;
        off[links[l].i2 + 1]++;
    }
    for (int v = 0; v < nv; v++) off[v + 1] += off[v];
    std::vector<int> fill(off.begin(), off.end() - 1);
    for (int l = 0; l < nl; l++) {
        inc[fill[links[l].i1]++] = 2 * l;
        inc[fill[links[l].i2]++] = 2 * l + 1;
    }
    dev_alloc(J.lk, lk.size());
    dev_alloc(J.lr, lr.size());
    dev_alloc(J.disp, lk.size());
    dev_alloc(J.d, lk.size());
    dev_alloc(J.off, off.size());
    dev_alloc(J.inc, inc.size());
    HIP_OK(hipMemcpy(J.lk, lk.data(), lk.size() * sizeof(int2), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(J.lr, lr.data(), lr.size() * sizeof(float2), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(J.off, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(J.inc, inc.data(), inc.size() * sizeof(int), hipMemcpyHostToDevice));
    J.nl = nl;
    J.nv = nv;
}

void ClothModel::render_mesh(Mesh& mesh, f3 offset, const std::vector<int>& map)
{
    const int nv = (int)mesh.vertices.size(), nt = (int)(mesh.indices.size() / 3);
    if (map.empty() ? nv != (int)pos.size() : (int)map.size() != nv)
        throw std::runtime_error("the render mesh does not match the particles");
    sync_to_device();
    GpuSolver& g = gpu;
    GpuSolver::Render& r = g.render;
    if (r.nv != nv || r.nt != nt) {
        dev_free(r.idx); dev_free(r.map); dev_free(r.vtOff); dev_free(r.vt); dev_free(r.triA); dev_free(r.triB);
        dev_free(r.outV); dev_free(r.outN); dev_free(r.triN);
        r.nv = r.nt = -1;  // stays invalid until every buffer below is rebuilt
        if (!map.empty()) {
            for (int v : map)
                if (v < 0 || v >= (int)pos.size()) throw std::runtime_error("render mesh map out of range");
            dev_alloc(r.map, nv);
            HIP_OK(hipMemcpy(r.map, map.data(), nv * sizeof(int), hipMemcpyHostToDevice));
        }
        // each vertex's (triangle, corner) entries in ascending triangle order
        std::vector<int> off(nv + 1, 0), list(3 * (size_t)nt);
        for (int k = 0; k < 3 * nt; k++) {
            const unsigned v = mesh.indices[k];
            if (v >= (unsigned)nv) throw std::runtime_error("render mesh index out of range");
            off[v + 1]++;
        }
        for (int v = 0; v < nv; v++) off[v + 1] += off[v];
        std::vector<int> fill(off.begin(), off.end() - 1);
        for (int k = 0; k < 3 * nt; k++) list[fill[mesh.indices[k]]++] = (k / 3) * 4 + k % 3;
        dev_alloc(r.idx, std::max(1, 3 * nt));
        dev_alloc(r.vtOff, nv + 1);
        dev_alloc(r.vt, std::max(1, 3 * nt));
        dev_alloc(r.triA, std::max(1, nt));
        dev_alloc(r.triB, std::max(1, nt));
        dev_alloc(r.outV, 3 * (size_t)nv);
        dev_alloc(r.outN, 3 * (size_t)nv);
        dev_alloc(r.triN, std::max<size_t>(1, 3 * (size_t)nt));
        if (nt > 0) {
            HIP_OK(hipMemcpy(r.idx, mesh.indices.data(), 3 * (size_t)nt * sizeof(unsigned), hipMemcpyHostToDevice));
            HIP_OK(hipMemcpy(r.vt, list.data(), list.size() * sizeof(int), hipMemcpyHostToDevice));
        }
        HIP_OK(hipMemcpy(r.vtOff, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice));
        r.nv = nv;
        r.nt = nt;
    }
    launch_render_mesh(g.stream, g.d_pos, r.map, nv, r.idx, nt, offset, r.vtOff, r.vt, r.triA, r.triB, r.outV, r.outN,
                       r.triN);
    HIP_OK(hipGetLastError());
    mesh.normals.resize(nv);
    mesh.triNormals.resize(nt);
    static_assert(sizeof(scVector3) == 3 * sizeof(float), "scVector3 is three packed floats");
    HIP_OK(hipMemcpyAsync(mesh.vertices.data(), r.outV, 3 * (size_t)nv * sizeof(float), hipMemcpyDeviceToHost, g.stream));
    HIP_OK(hipMemcpyAsync(mesh.normals.data(), r.outN, 3 * (size_t)nv * sizeof(float), hipMemcpyDeviceToHost, g.stream));
    if (nt > 0)
        HIP_OK(hipMemcpyAsync(mesh.triNormals.data(), r.triN, 3 * (size_t)nt * sizeof(float), hipMemcpyDeviceToHost,
                              g.stream));
    HIP_OK(hipStreamSynchronize(g.stream));
}

// Physics/ClothPatch.cpp:337-381
void ClothPatch::compute_strain_map(scVector3* colors, int count)
{
    ClothModel& m = *mModel;
    m.sync_to_host();
    const int n = std::min(count, (int)m.pos.size());
    for (int i = 0; i < n; i++) colors[i] = scVector3{0.f, 0.f, 0.f};
    const float PI = 3.1415926535897932384626433832795f;
    for (const Triangle& tri : m.tris) {
        f3 dx1 = sub(m.pos[tri.i2], m.pos[tri.i1]);
        f3 dx2 = sub(m.pos[tri.i3], m.pos[tri.i1]);
        f3 wu = mul(tri.invDet, sub(mul(tri.dv2, dx1), mul(tri.dv1, dx2)));
        f3 wv = mul(tri.invDet, add(mul(-tri.du2, dx1), mul(tri.du1, dx2)));
        float euu = 0.5f * (dot(wu, wu) - 1), evv = 0.5f * (dot(wv, wv) - 1), euv = dot(wu, wv);
        float s = sqrtf(euu * euu + evv * evv + euv * euv) * 10.f;
        f3 e1 = normalized(dx1), e2 = normalized(sub(m.pos[tri.i3], m.pos[tri.i2])), e3 = normalized(neg(dx2));
        float t1 = std::acos(-dot(e1, e3)), t2 = std::acos(-dot(e1, e2)), t3 = std::acos(-dot(e2, e3));
        const uint32_t id[3] = {tri.i1, tri.i2, tri.i3};
        const float tt[3] = {t1, t2, t3};
        for (int k = 0; k < 3; k++) {
            if ((int)id[k] >= n) continue;
            float w = tt[k] / PI;
            colors[id[k]].x += w * s;
            colors[id[k]].z += w * s;
        }
    }
}

}  // namespace sc

namespace sc {

// ---------------------------------------------------------------- batch ----

void ClothModel::validate_batch() const
{
    if (pos.empty()) throw std::runtime_error("a batched patch has no cloth (scInitCloth first)");
    if (!mFEM) throw std::runtime_error("batched cloths use the FEM stretch model (scSetUseFEM(true))");
    if (mNumIterations < 1) throw std::runtime_error("NumIterations must be >= 1");
    if (mBendFreq < 1) throw std::runtime_error("bend frequency must be >= 1");
    if (msActive) throw std::runtime_error("a batched patch cannot carry a mouse spring (step it with scStepCloth)");
    if (mSolver != SOLVER_GAUSS_SEIDEL) throw std::runtime_error("batched cloths use the Gauss-Seidel solver");
    if (mCollFlags & (CF_EDGES | CF_TRIANGLES | CF_SELF))
        throw std::runtime_error("batched cloths support vertex and wall contacts only");
    for (const auto& c : mOwner->mWorld->colls)
        if (c->type != CT_WALLS && c->type != CT_SPHERE && c->type != CT_SDF)
            throw std::runtime_error("batched cloths support walls, spheres and grid SDFs only");
    if (mOwner->mWorld->colls.size() > MAX_COLLIDABLES) throw std::runtime_error("too many collidables (max 8)");
}

CollSet ClothModel::prepare_batch(int& slots)
{
    if (mSweepMode != SWEEP_COLOURS) set_sweep(SWEEP_COLOURS, 0);
    sync_to_device();
    for (const auto& c : mOwner->mWorld->colls)
        if (c->type == CT_SDF && !((CollisionSDF*)c.get())->sdf->d_nodes) ((CollisionSDF*)c.get())->sdf->upload();
    int edgeSlots = 0, triSlots = 0;
    CollSet cs = make_collset(slots, edgeSlots, triSlots);
    ensure_contact_capacity(std::max(slots, 1));
    gpu.activeSlots = slots;
    return cs;
}

ClothDesc ClothModel::batch_desc(float h, const CollSet* d_cs, bool anyContacts) const
{
    const GpuSolver& g = gpu;
    ClothDesc d;
    memset(&d, 0, sizeof(d));
    d.pos = g.d_pos; d.prev = g.d_prev; d.vel = g.d_vel;
    d.tr = TriStreams{g.d_tidx, g.d_tA, g.d_tB, g.d_tC};
    d.triCol = g.d_triColOff;
    d.bidx = g.d_bidx;
    d.bendCol = g.d_bendColOff;
    d.cb = ContactBuf{g.d_cCount, g.d_cP, g.d_cN, g.d_cSrc, g.d_cFeat, nullptr, g.nv, g.slots};
    d.cs = d_cs;
    d.ctl = g.d_ctl;
    d.nv = g.nv; d.nt = g.nt; d.nb = g.nb;
    d.triColours = triCol.numPhases();
    d.bendColours = bendCol.numPhases();
    // the same parameter expressions as record_substep
    const f3 gravity{0.f, -9.8f, 0.f};
    const f3 hg = mul(h, gravity);
    d.gdv = f3{mUnit * hg.x, mUnit * hg.y, mUnit * hg.z};
    d.h = h;
    d.h2 = h * h;
    d.epsUV = 2 * (1 + mPoisson) / (smax(mShearStiff, 0.1f) * mUnit * mUnit);
    d.sStiff = 0.0005f * mBendStiff;
    d.drag = mDragCoeff;
    d.thickness = mThickness;
    d.tolerance = mTolerance;
    d.mu = mFriction;
    d.invH = 1.f / h;
    d.selfThreshold = mSelfCollThreshold;
    d.iterations = mNumIterations;
    d.fixedIterations = mFixedIterations ? 1 : 0;
    d.bendFreq = mBendFreq;
    d.dihedral = mDihedral ? 1 : 0;
    d.collFlags = mCollFlags;
    d.anyContacts = anyContacts ? 1 : 0;
    return d;
}

ClothBatch::~ClothBatch()
{
    release_graph();
    dev_free(d_descs);
    dev_free(d_cs);
    dev_free(d_clock);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
}

void ClothBatch::release_graph()
{
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    graphKey.clear();
}

void ClothBatch::add(ClothPatch* p)
{
    for (ClothPatch* q : patches)
        if (q == p) throw std::runtime_error("patch already in the batch");
    patches.push_back(p);
    release_graph();
}

void ClothBatch::set_launch(int bs, int st)
{
    blockSize = bs;
    stage = st;
    tuned = bs != 0 && st >= 0;
    trials.clear();
    trialConfigs.clear();
    trialMs.clear();
}

// One step of every cloth: `NumSteps` substeps, each one launch of the
// batch kernel. With blockSize 0 the first steps time the candidate
// configurations (every configuration computes identical results: each
// colour's constraints touch disjoint vertices) and keep the fastest.
void ClothBatch::step(float dt)
{
    if (patches.empty()) return;
    const int n = patches[0]->mNumSteps;
    for (ClothPatch* p : patches)
        if (p->mNumSteps != n) throw std::runtime_error("all patches of a batch need the same number of substeps");
    if (n < 1) throw std::runtime_error("NumSteps must be >= 1");
    for (ClothPatch* p : patches) p->mModel->validate_batch();
    const float h = dt / n;
    if (!stream) HIP_OK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    if (stageLimit == 0) {
        int dev = 0, bytes = 0;
        HIP_OK(hipGetDevice(&dev));
        HIP_OK(hipDeviceGetAttribute(&bytes, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        stageLimit = std::max(1, (bytes - 1024) / (int)sizeof(float4));
        batch_allow_lds(stageLimit * (int)sizeof(float4));
    }
    const int m = (int)patches.size();
    // largest cloths first: workgroups start in launch order
    std::vector<int> order(m);
    for (int i = 0; i < m; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        const ClothModel& A = *patches[a]->mModel;
        const ClothModel& B = *patches[b]->mModel;
        return (long)A.pos.size() * A.mNumIterations > (long)B.pos.size() * B.mNumIterations;
    });
    std::vector<CollSet> cs(m);
    std::vector<int> any(m);
    for (int k = 0; k < m; k++) {
        int slots = 0;
        cs[k] = patches[order[k]]->mModel->prepare_batch(slots);
        any[k] = slots > 0;
    }
    if (m > cap) {
        dev_alloc(d_descs, m);
        dev_alloc(d_cs, m);
        cap = m;
        release_graph();
    }
    // the preparation's copies run on the null stream and the models' streams
    HIP_OK(hipDeviceSynchronize());
    std::vector<ClothDesc> descs(m);
    for (int k = 0; k < m; k++) descs[k] = patches[order[k]]->mModel->batch_desc(h, d_cs + k, any[k] != 0);
    if (hostCs.size() != cs.size() || memcmp(hostCs.data(), cs.data(), m * sizeof(CollSet)) != 0) {
        HIP_OK(hipMemcpyAsync(d_cs, cs.data(), m * sizeof(CollSet), hipMemcpyHostToDevice, stream));
        hostCs = cs;
    }
    if (hostDescs.size() != descs.size() || memcmp(hostDescs.data(), descs.data(), m * sizeof(ClothDesc)) != 0) {
        HIP_OK(hipMemcpyAsync(d_descs, descs.data(), m * sizeof(ClothDesc), hipMemcpyHostToDevice, stream));
        hostDescs = descs;
        release_graph();
    }
    HIP_OK(hipStreamSynchronize(stream));

    int bs = blockSize, st = stage;
    bool timing = false;
    if (!tuned && (blockSize == 0 || stage < 0)) {
        if (trials.empty() && trialMs.empty()) {
            for (int b : {256, 512, 1024})
                for (int g : {0, 1})
                    if ((blockSize == 0 || b == blockSize) && (stage < 0 || g == stage)) trials.push_back({b, g});
        }
        const size_t t = trialMs.size();
        bs = trials[t].first;
        st = trials[t].second;
        timing = true;
    } else if (tuned) {
        bs = blockSize;
        st = stage;
    }
    // LDS sized for the largest cloth that fits (every workgroup allocates it)
    int stageVerts = 0;
    if (st)
        for (ClothPatch* p : patches) {
            const int nv = (int)p->mModel->pos.size();
            if (nv <= stageLimit) stageVerts = std::max(stageVerts, nv);
        }
    if (recordClock && m > clockCap) {
        release_graph();
        dev_alloc(d_clock, 3 * (size_t)m);
        clockCap = m;
    }
    unsigned long long* clock = recordClock ? d_clock : nullptr;
    std::string key = std::to_string(bs) + "/" + std::to_string(stageVerts) + "/" + std::to_string(n) + "/" +
                      std::to_string(m) + "/" + std::to_string((uintptr_t)d_descs) + "/" +
                      std::to_string((uintptr_t)clock);
    if (useGraph && (!exec || key != graphKey)) {
        release_graph();
        HIP_OK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        for (int i = 0; i < n; i++) launch_batch_substep(stream, d_descs, m, bs, stageVerts, clock);
        HIP_OK(hipStreamEndCapture(stream, &graph));
        HIP_OK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        graphKey = key;
    }
    if (!ev0) {
        HIP_OK(hipEventCreate(&ev0));
        HIP_OK(hipEventCreate(&ev1));
    }
    HIP_OK(hipEventRecord(ev0, stream));
    if (useGraph)
        HIP_OK(hipGraphLaunch(exec, stream));
    else
        for (int i = 0; i < n; i++) launch_batch_substep(stream, d_descs, m, bs, stageVerts, clock);
    HIP_OK(hipEventRecord(ev1, stream));
    HIP_OK(hipStreamSynchronize(stream));
    HIP_OK(hipGetLastError());
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, ev0, ev1));
    lastStepMs = ms;
    lastSubsteps = n;
    if (clock) {
        std::vector<unsigned long long> c(3 * (size_t)m);
        HIP_OK(hipMemcpy(c.data(), clock, c.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
        clockStart.assign(m, 0);
        clockEnd.assign(m, 0);
        clockUnit.assign(m, 0);
        for (int k = 0; k < m; k++) {
            clockStart[order[k]] = c[3 * (size_t)k];
            clockEnd[order[k]] = c[3 * (size_t)k + 1];
            clockUnit[order[k]] = c[3 * (size_t)k + 2];
        }
    }
    if (timing) {
        trialConfigs.push_back({bs, st});
        trialMs.push_back(ms);
        if (trialMs.size() == trials.size()) {
            size_t best = 0;
            for (size_t t = 1; t < trialMs.size(); t++)
                if (trialMs[t] < trialMs[best]) best = t;
            blockSize = trials[best].first;
            stage = trials[best].second;
            tuned = true;
        }
    }
    for (ClothPatch* p : patches) p->mModel->after_device_step();
}

}  // namespace sc
