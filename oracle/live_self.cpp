// TEST INFRASTRUCTURE ONLY (see oracle_vec.h): the reference's live
// self-collision detection, restated as a CPU baseline for the self-colliding
// cloth. Only bench.py's cpu_baseline leg times it and tests/ check it; no
// product code loads it.
//
// ClothModel::MeshCollisions on the cloth's own previous-position mesh
// (Physics/ClothCollision.cpp:41-46,89-105) calls
// MeshClosestPoints::ClosestPoints(mesh, tree, radTol)
// (Geometry/ClosestPointsBetweenMeshes.cpp:538-592), which ignores the tree
// and is brute force:
//   * per vertex, ClosestPointsOnMeshToPoint over every triangle
//     (Geometry/ClosestPointOnMesh.cpp:142-246) with the Voronoi-region tests
//     of Geometry/VoronoiRegions.cpp:11-113, then CheckAdjacency
//     (ClosestPointsBetweenMeshes.cpp:441-460);
//   * per edge, ClosestPointOnMeshToSegment over every later edge
//     (ClosestPointOnMesh.cpp:499-542), then CheckEdgeAdjacency
//     (ClosestPointsBetweenMeshes.cpp:473-536).
// The mesh is prepared as ClothPatch::Init prepares it
// (Physics/ClothPatch.cpp:232-235): normals, edges, vertex and edge one-rings.
//
// The reference appends under `omp critical` inside `omp parallel for`, so its
// list order follows the threads. Here each query's results are kept apart
// and concatenated in query order: any thread count gives the serial order,
// which is the order the compiled reference gives without OpenMP
// (tests/golden/ref_live_self.npz pins it bit for bit).
#include <omp.h>

#include <vector>

#include "oracle_geom.h"

namespace oracle {
namespace {

struct LiveMesh {
    std::vector<V3<float>> v, vn, tn;
    std::vector<uint32_t> idx;
    std::vector<MeshEdge> edges;
    std::vector<int> sideEdge;                          // 3 t + k -> edge
    std::vector<std::vector<int>> vtxRing, edgeRing;   // Mesh::vtxOneRings, edgeOneRings

    int nt() const { return (int)(idx.size() / 3); }

    void build()
    {
        mesh_face_vertex_normals(v, idx, vn, tn);
        edges = mesh_edges(idx, &sideEdge);
        vtxRing.assign(v.size(), {});
        edgeRing.assign(v.size(), {});
        for (size_t e = 0; e < edges.size(); e++) {
            vtxRing[edges[e].i1].push_back((int)edges[e].i2);
            vtxRing[edges[e].i2].push_back((int)edges[e].i1);
            edgeRing[edges[e].i1].push_back((int)e);
            edgeRing[edges[e].i2].push_back((int)e);
        }
    }

    bool adjacent(int a, int b) const
    {
        for (int x : vtxRing[a])
            if (x == b) return true;
        return false;
    }
    bool border_edge(int e) const { return edges[e].t2 < 0; }
    bool border_vertex(int i) const
    {
        if (i < 0) return false;
        for (int e : edgeRing[i])
            if (border_edge(e)) return true;
        return false;
    }

    // VoronoiRegions.cpp:11-35. The sign test is the reference's chained
    // comparison (signbit1 == signbit2) == signbit3, as written.
    bool in_triangle_region(V3<float> p, int t) const
    {
        const V3<float> v1 = v[idx[3 * t]], v2 = v[idx[3 * t + 1]], v3 = v[idx[3 * t + 2]];
        const V3<float> e1 = v2 - v1, e2 = v3 - v2, e3 = v1 - v3;
        const V3<float> n = tn[t];
        const V3<float> h1 = cross(e1, n), h2 = cross(e2, n), h3 = cross(e3, n);
        const float d1 = dot(h1, p - v1), d2 = dot(h2, p - v2), d3 = dot(h3, p - v3);
        return (std::signbit(d1) == std::signbit(d2)) == std::signbit(d3);
    }

    // VoronoiRegions.cpp:37-83 with its default eps 0.002
    // (VoronoiRegions.h:11)
    bool in_edge_region(V3<float> p, int e, float eps = 0.002f) const
    {
        const MeshEdge& E = edges[e];
        const V3<float> a = v[E.i1], b = v[E.i2];
        V3<float> d = b - a;
        normalize(d);
        const float d3 = dot(d, p - a), d4 = dot(d, p - b);
        if (d3 < 0 || d4 > 0) return false;
        if (E.swapped) d = -d;
        if (E.t2 >= 0) {
            V3<float> h1 = cross(d, tn[E.t1]);
            V3<float> h2 = -cross(d, tn[E.t2]);
            normalize(h1);
            normalize(h2);
            if (dot(h1, p - a) < -eps || dot(h2, p - a) < -eps) return false;
        } else {
            V3<float> h1 = cross(d, tn[E.t1]);
            normalize(h1);
            if (dot(h1, p - a) < -eps) return false;
        }
        return true;
    }

    // VoronoiRegions.cpp:85-111
    bool in_vertex_region(V3<float> p, int i) const
    {
        const V3<float> x = v[i];
        const float side = dot(vn[i], p - x) < 0 ? -1.f : 1.f;
        const float eps = 0.1f;
        for (int e : edgeRing[i]) {
            const MeshEdge& E = edges[e];
            V3<float> d = v[E.i1] - v[E.i2];
            normalize(d);
            if ((int)E.i2 != i) d = -d;
            const float t = dot(d, p - x);
            if (side > 0 && t < -eps) return false;
            if (side < 0 && t > eps) return false;
        }
        return true;
    }

    // ClosestPointsBetweenMeshes.cpp:441-460
    bool check_adjacency(int vtx, int t, int region) const
    {
        const int i1 = (int)idx[3 * t], i2 = (int)idx[3 * t + 1], i3 = (int)idx[3 * t + 2];
        if (i1 == vtx || i2 == vtx || i3 == vtx) return true;
        if (region == TR_VERTEX_A && adjacent(vtx, i1)) return true;
        if (region == TR_VERTEX_B && adjacent(vtx, i2)) return true;
        if (region == TR_VERTEX_C && adjacent(vtx, i3)) return true;
        if (region == TR_EDGE_AB && adjacent(vtx, i1) && adjacent(vtx, i2)) return true;
        if (region == TR_EDGE_BC && adjacent(vtx, i2) && adjacent(vtx, i3)) return true;
        if (region == TR_EDGE_AC && adjacent(vtx, i1) && adjacent(vtx, i3)) return true;
        return false;
    }

    bool same_triangle(int vtx, int t) const
    {
        if (t < 0) return false;
        return vtx == (int)idx[3 * t] || vtx == (int)idx[3 * t + 1] || vtx == (int)idx[3 * t + 2];
    }
};

// ClosestTriangleToPoint (ClosestPointOnMesh.h:21-33) as written out
struct VTInfo {
    int vtx = -1, tri = -1, region = -1, regionType = -1, feature = -1;
    float distance = FLT_MAX, cp[3] = {0, 0, 0}, bar[3] = {0, 0, 0}, n[3] = {0, 0, 0}, side = FLT_MAX;
};

// ClosestEdgeToSegment (ClosestPointOnMesh.h:44-56)
struct EEInfo {
    int edge1 = -1, edge = -1, region = -1;
    float distance = -FLT_MAX, cs[2] = {0, 0}, cm[2] = {0, 0}, ps[3] = {0, 0, 0}, pm[3] = {0, 0, 0},
          n[3] = {0, 0, 0}, side = FLT_MAX;
};

// ClosestPointsOnMeshToPoint(p = vertex i, maxDist = radTol, idx = i), then
// the adjacency filter of ClosestPoints (self)
void vertex_query(const LiveMesh& M, int i, float maxDist, std::vector<VTInfo>& out)
{
    const V3<float> p = M.v[i];
    const bool isBorder = M.border_vertex(i);
    std::vector<VTInfo> results;
    for (int t = 0; t < M.nt(); t++) {
        const int i0 = (int)M.idx[3 * t], i1 = (int)M.idx[3 * t + 1], i2 = (int)M.idx[3 * t + 2];
        const V3<float> a = M.v[i0], b = M.v[i1], c = M.v[i2];
        V3<float> pt, coords;
        const int region = closest_pt_point_triangle(p, a, b, c, pt, coords);
        const bool inVR = M.in_vertex_region(pt, i);
        const V3<float> delta = p - pt;
        const float dist = length(delta);
        if (dist > maxDist) continue;
        if (dist == 0) continue;
        VTInfo r;
        auto edge_case = [&](int k) {
            const int e = M.sideEdge[3 * t + k];
            if (!M.border_edge(e) && !(inVR || M.in_edge_region(p, e))) return false;
            r.regionType = 1;
            r.feature = e;
            return true;
        };
        auto vertex_case = [&](int j) {
            if (!isBorder && !(inVR || M.in_vertex_region(p, j))) return false;
            r.regionType = 2;
            r.feature = j;
            return true;
        };
        bool keep = true;
        switch (region) {
        case TR_FACE_INTERIOR:
            keep = inVR || M.in_triangle_region(p, t);
            r.regionType = 0;
            break;
        case TR_EDGE_AB: keep = edge_case(0); break;
        case TR_EDGE_BC: keep = edge_case(1); break;
        case TR_EDGE_AC: keep = edge_case(2); break;
        case TR_VERTEX_A: keep = vertex_case(i0); break;
        case TR_VERTEX_B: keep = vertex_case(i1); break;
        case TR_VERTEX_C: keep = vertex_case(i2); break;
        default: break;
        }
        if (!keep) continue;
        if (r.feature >= 0) {
            bool found = false;
            for (const VTInfo& q : results)
                if (q.regionType == r.regionType && q.feature == r.feature) {
                    found = true;
                    break;
                }
            if (found) continue;
        }
        r.distance = dist;
        r.tri = t;
        r.cp[0] = pt.x; r.cp[1] = pt.y; r.cp[2] = pt.z;
        r.bar[0] = coords.x; r.bar[1] = coords.y; r.bar[2] = coords.z;
        const V3<float> n = (1.f / dist) * delta;
        r.n[0] = n.x; r.n[1] = n.y; r.n[2] = n.z;
        r.region = region;
        r.side = signed_volume(p, a, b, c);
        results.push_back(r);
    }
    for (VTInfo& r : results) {
        r.vtx = i;
        if (r.tri >= 0 && !M.check_adjacency(i, r.tri, r.region)) out.push_back(r);
    }
}

// ClosestPointOnMeshToSegment(p, q, mesh, info, e2 = e) and CheckEdgeAdjacency
void edge_query(const LiveMesh& M, int e2, std::vector<EEInfo>& out)
{
    const MeshEdge& E2 = M.edges[e2];
    const V3<float> p = M.v[E2.i1], q = M.v[E2.i2];
    EEInfo r;
    r.edge1 = e2;
    float distMin = FLT_MAX;
    for (int e1 = e2 + 1; e1 < (int)M.edges.size(); e1++) {
        const uint32_t i1 = M.edges[e1].i1, i2 = M.edges[e1].i2;
        if (i1 == E2.i1 || i1 == E2.i2 || i2 == E2.i1 || i2 == E2.i2) continue;
        const V3<float> a = M.v[i1], b = M.v[i2];
        float ps, pm;
        V3<float> cs, cm;
        const int region = closest_pt_segm_segm(p, q, a, b, ps, pm, cs, cm);
        const float dist = length(cs - cm);
        if (dist != 0 && dist < distMin) {
            distMin = dist;
            r.edge = e1;
            r.distance = dist;
            r.pm[0] = cm.x; r.pm[1] = cm.y; r.pm[2] = cm.z;
            r.ps[0] = cs.x; r.ps[1] = cs.y; r.ps[2] = cs.z;
            r.cs[0] = 1.0f - ps; r.cs[1] = ps;
            r.cm[0] = 1.0f - pm; r.cm[1] = pm;
            V3<float> n = cm - cs;
            normalize(n);
            r.n[0] = n.x; r.n[1] = n.y; r.n[2] = n.z;
            r.side = signed_volume(p, q, a, b);
            r.region = region;
        }
    }
    // CheckEdgeAdjacency: true rejects the pair
    const float eps = 1e-5f;
    if (r.edge < 0) return;
    const MeshEdge& F = M.edges[r.edge];
    const int i1 = (int)E2.i1, i2 = (int)E2.i2, j1 = (int)F.i1, j2 = (int)F.i2;
    if (i1 == j1 || i1 == j2 || i2 == j1 || i2 == j2) return;
    if (((r.region & ER_VERTEX_P1) != 0 || r.cs[0] < eps) && (M.same_triangle(i1, F.t1) || M.same_triangle(i1, F.t2)))
        return;
    if (((r.region & ER_VERTEX_Q1) != 0 || r.cs[1] < eps) && (M.same_triangle(i2, F.t1) || M.same_triangle(i2, F.t2)))
        return;
    if (((r.region & ER_VERTEX_P2) != 0 || r.cm[0] < eps) &&
        (M.same_triangle(j1, E2.t1) || M.same_triangle(j1, E2.t2)))
        return;
    if (((r.region & ER_VERTEX_Q2) != 0 || r.cm[1] < eps) &&
        (M.same_triangle(j2, E2.t1) || M.same_triangle(j2, E2.t2)))
        return;
    out.push_back(r);
}

// Runs query(k, out) for the listed items on OpenMP threads and concatenates
// the per-item outputs in list order
template <class T, class F>
std::vector<T> ordered_queries(const std::vector<int>& items, F query)
{
    const long n = (long)items.size();
    std::vector<std::vector<T>> part(n);
#pragma omp parallel for schedule(dynamic, 1)
    for (long k = 0; k < n; k++) query(items[k], part[k]);
    std::vector<T> out;
    for (auto& p : part) out.insert(out.end(), p.begin(), p.end());
    return out;
}

std::vector<int> item_list(const int* q, int nq, int count)
{
    std::vector<int> items;
    if (q) {
        items.assign(q, q + nq);
    } else {
        items.resize(count);
        for (int i = 0; i < count; i++) items[i] = i;
    }
    return items;
}

}  // namespace
}  // namespace oracle

using namespace oracle;

extern "C" {

// The cloth mesh at the given (previous) positions
void* ol_create(int nv, const float* verts, int ni, const unsigned* idx)
{
    LiveMesh* M = new LiveMesh;
    M->v.resize(nv);
    for (int i = 0; i < nv; i++) M->v[i] = V3<float>{verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]};
    M->idx.assign(idx, idx + ni);
    M->build();
    return M;
}

void ol_free(void* h) { delete static_cast<LiveMesh*>(h); }

// (vertices, triangles, edges)
void ol_counts(void* h, int* out)
{
    const LiveMesh* M = static_cast<LiveMesh*>(h);
    out[0] = (int)M->v.size();
    out[1] = M->nt();
    out[2] = (int)M->edges.size();
}

// The VT infos of the listed vertices (all, in order, if q is null); ids 5 and
// data 11 per info as ref_live_self writes them. Returns the info count and
// writes at most cap.
long ol_vertex_queries(void* h, float radTol, const int* q, int nq, int* ids, float* data, long cap)
{
    const LiveMesh& M = *static_cast<LiveMesh*>(h);
    const std::vector<VTInfo> out = ordered_queries<VTInfo>(
        item_list(q, nq, (int)M.v.size()), [&](int i, std::vector<VTInfo>& o) { vertex_query(M, i, radTol, o); });
    for (long k = 0; k < (long)out.size() && k < cap; k++) {
        const VTInfo& r = out[k];
        const int iv[5] = {r.vtx, r.tri, r.region, r.regionType, r.feature};
        const float dv[11] = {r.distance, r.cp[0], r.cp[1], r.cp[2], r.bar[0], r.bar[1],
                              r.bar[2],   r.n[0],  r.n[1],  r.n[2],  r.side};
        std::copy(iv, iv + 5, ids + 5 * k);
        std::copy(dv, dv + 11, data + 11 * k);
    }
    return (long)out.size();
}

// The EE infos of the listed edges (all if q is null); ids 3 and data 15 per
// info as ref_live_self writes them
long ol_edge_queries(void* h, const int* q, int nq, int* ids, float* data, long cap)
{
    const LiveMesh& M = *static_cast<LiveMesh*>(h);
    const std::vector<EEInfo> out = ordered_queries<EEInfo>(
        item_list(q, nq, (int)M.edges.size()), [&](int e, std::vector<EEInfo>& o) { edge_query(M, e, o); });
    for (long k = 0; k < (long)out.size() && k < cap; k++) {
        const EEInfo& r = out[k];
        const int iv[3] = {r.edge1, r.edge, r.region};
        const float dv[15] = {r.distance, r.cs[0], r.cs[1], r.cm[0], r.cm[1], r.ps[0], r.ps[1], r.ps[2],
                              r.pm[0],    r.pm[1], r.pm[2], r.n[0],  r.n[1],  r.n[2],  r.side};
        std::copy(iv, iv + 3, ids + 3 * k);
        std::copy(dv, dv + 15, data + 15 * k);
    }
    return (long)out.size();
}

}  // extern "C"
