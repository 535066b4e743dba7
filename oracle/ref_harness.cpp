// TEST INFRASTRUCTURE ONLY. A C-callable harness around the reference's own
// Geometry translation units, compiled from /root/reference by
// oracle/build_ref.sh into oracle/_ref/libref_geometry.so. It is run only by
// tests/golden/make_golden.py, in a container that has the reference, to
// write golden vectors; nothing on a GPU host loads it.
//
// Only Geometry/{Mesh,Collision3D,AabbTree,ClosestPointOnMesh,VoronoiRegions,
// ClosestPointsBetweenMeshes}.cpp are compiled: they build with the image's
// own headers. The Physics translation units do not (they need Win32
// types from Engine/Engine.h and std::floorf, which libstdc++ 11 lacks), so
// the solver itself is pinned by the oracle's mathematical tests instead.
#include "Geometry/AabbTree.h"
#include "Geometry/ClosestPointOnMesh.h"
#include "Geometry/ClosestPointsBetweenMeshes.h"
#include "Geometry/Collision3D.h"
#include "Geometry/Mesh.h"

#include <omp.h>

#include <chrono>
#include <cstring>

using Math::Vector3;

// Defined in Geometry/ClosestPointsBetweenMeshes.cpp:441,473 (external
// linkage, no header declares them)
namespace Geometry {
bool CheckAdjacency(const Mesh& mesh, int vtx, int tri, int region);
bool CheckEdgeAdjacency(ClosestEdgeToSegment& info, const Mesh& mesh1, int i1, int i2, const Mesh::Edge& edge1);
}  // namespace Geometry

extern "C" {

// Mesh::ConstructEdges on an indexed triangle list (Geometry/Mesh.cpp:65-147).
// Returns the number of unique edges; out = (i1, i2, t1, t2) per edge.
int ref_construct_edges(const float* verts, int nv, const unsigned* idx, int ni, int* out, int cap)
{
    Geometry::Mesh m;
    m.vertices.resize(nv);
    for (int i = 0; i < nv; i++) m.vertices[i] = Vector3(verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]);
    m.indices.assign(idx, idx + ni);
    m.ComputeNormals();
    m.ConstructEdges();
    int n = (int)m.edges.size();
    for (int e = 0; e < n && e < cap; e++) {
        out[4 * e] = (int)m.edges[e].i1;
        out[4 * e + 1] = (int)m.edges[e].i2;
        out[4 * e + 2] = m.edges[e].t1;
        out[4 * e + 3] = m.edges[e].t2;
    }
    return n;
}

// Mesh::ComputeNormals (angle weighted, Geometry/Mesh.cpp:10-63)
void ref_compute_normals(const float* verts, int nv, const unsigned* idx, int ni, float* out)
{
    Geometry::Mesh m;
    m.vertices.resize(nv);
    for (int i = 0; i < nv; i++) m.vertices[i] = Vector3(verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]);
    m.indices.assign(idx, idx + ni);
    m.ComputeNormals();
    for (int i = 0; i < nv; i++) {
        out[3 * i] = m.normals[i].x;
        out[3 * i + 1] = m.normals[i].y;
        out[3 * i + 2] = m.normals[i].z;
    }
}

// Geometry/Collision3D.cpp ClosestPtPointTriangle; out = cp(3), bar(3)
int ref_closest_pt_point_triangle(const float* p, const float* a, const float* b, const float* c, float* out)
{
    Vector3 cp, bar;
    int r = Geometry::ClosestPtPointTriangle(Vector3(p[0], p[1], p[2]), Vector3(a[0], a[1], a[2]),
                                             Vector3(b[0], b[1], b[2]), Vector3(c[0], c[1], c[2]), cp, bar);
    float o[6] = {cp.x, cp.y, cp.z, bar.x, bar.y, bar.z};
    memcpy(out, o, sizeof(o));
    return r;
}

// Geometry/Collision3D.cpp:10-58 ClosestPtSegmSegm; out = s, t, c1(3), c2(3)
int ref_closest_pt_segm_segm(const float* p1, const float* q1, const float* p2, const float* q2, float* out)
{
    float s, t;
    Vector3 c1, c2;
    int r = Geometry::ClosestPtSegmSegm(Vector3(p1[0], p1[1], p1[2]), Vector3(q1[0], q1[1], q1[2]),
                                        Vector3(p2[0], p2[1], p2[2]), Vector3(q2[0], q2[1], q2[2]), s, t, c1, c2);
    float o[8] = {s, t, c1.x, c1.y, c1.z, c2.x, c2.y, c2.z};
    memcpy(out, o, sizeof(o));
    return r;
}

// Geometry/Collision3D.cpp:143-210 ClosestPtSegmTriangle; out = dist, param,
// bar (3), closest on the segment (3), closest on the triangle (3)
void ref_closest_pt_segm_triangle(const float* p, const float* q, const float* a, const float* b, const float* c,
                                  float* out)
{
    float param = 0.f;
    Math::BarycentricCoords bar;
    Vector3 ce, ct;
    float d = Geometry::ClosestPtSegmTriangle(Vector3(p[0], p[1], p[2]), Vector3(q[0], q[1], q[2]),
                                              Vector3(a[0], a[1], a[2]), Vector3(b[0], b[1], b[2]),
                                              Vector3(c[0], c[1], c[2]), param, bar, ce, ct);
    float o[11] = {d, param, bar.x, bar.y, bar.z, ce.x, ce.y, ce.z, ct.x, ct.y, ct.z};
    memcpy(out, o, sizeof(o));
}

// Geometry/Collision3D.h:125-141 ClosestPtSegm; out = t, closest point (3)
void ref_closest_pt_segm(const float* p, const float* a, const float* b, float* out)
{
    float t = 0.f;
    Vector3 cp = Geometry::ClosestPtSegm(Vector3(p[0], p[1], p[2]), Vector3(a[0], a[1], a[2]),
                                         Vector3(b[0], b[1], b[2]), t);
    float o[4] = {t, cp.x, cp.y, cp.z};
    memcpy(out, o, sizeof(o));
}

// Geometry/Collision3D.cpp:298-301
float ref_signed_volume(const float* a0, const float* a1, const float* a2, const float* a3)
{
    return Geometry::SignedVolume(Vector3(a0[0], a0[1], a0[2]), Vector3(a1[0], a1[1], a1[2]),
                                  Vector3(a2[0], a2[1], a2[2]), Vector3(a3[0], a3[1], a3[2]));
}

// Geometry/Collision3D.h:42-58
void ref_barycentric(const float* a, const float* b, const float* c, const float* p, float* out)
{
    Vector3 r = Geometry::Barycentric(Vector3(a[0], a[1], a[2]), Vector3(b[0], b[1], b[2]), Vector3(c[0], c[1], c[2]),
                                      Vector3(p[0], p[1], p[2]));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

// An obstacle mesh as the reference's demo prepares one
// (Demo/ClothDemo.cpp:100-103).
static void make_mesh(Geometry::Mesh& m, const float* verts, int nv, const unsigned* idx, int ni)
{
    m.vertices.resize(nv);
    for (int i = 0; i < nv; i++) m.vertices[i] = Vector3(verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]);
    m.indices.assign(idx, idx + ni);
    m.ComputeNormals();
    m.ConstructEdges();
    m.ConstructEdgeOneRings();
}

// ClosestPointOnMeshToPointAcc (Geometry/ClosestPointOnMesh.cpp:304-319)
// over the tree CollisionMesh::Update builds (Physics/Common.h:124-137:
// triangles, vertices and edges, 20 levels, tolerance tol). out per query:
// triangle, closest point (3), normal (3), distance, region.
void ref_closest_point_on_mesh_acc(const float* verts, int nv, const unsigned* idx, int ni, float tol,
                                   const float* q, int nq, float* out)
{
    Geometry::Mesh m;
    make_mesh(m, verts, nv, idx, ni);
    Geometry::AabbTree* tree = Geometry::ComputeMeshTree(
        m, Geometry::ATF_TRIANGLES | Geometry::ATF_VERTICES | Geometry::ATF_EDGES, 20, tol);
    for (int k = 0; k < nq; k++) {
        Geometry::ClosestTriangleToPoint r =
            Geometry::ClosestPointOnMeshToPointAcc(Vector3(q[3 * k], q[3 * k + 1], q[3 * k + 2]), m, tree);
        float o[9] = {(float)r.tri, r.closestPtMesh.x, r.closestPtMesh.y, r.closestPtMesh.z, r.normal.x,
                      r.normal.y, r.normal.z, r.distance, (float)r.region};
        memcpy(out + 9 * k, o, sizeof(o));
    }
    delete tree;
}

// ComputeTriangleNormal (ClosestPointOnMesh.cpp:37-97) at (triangle, coords)
void ref_triangle_normals(const float* verts, int nv, const unsigned* idx, int ni, const int* tris,
                          const float* coords, int n, int usePseudoNormal, int useMeshNormal, float* out)
{
    Geometry::Mesh m;
    make_mesh(m, verts, nv, idx, ni);
    for (int k = 0; k < n; k++) {
        Vector3 nn = Geometry::ComputeTriangleNormal(m, tris[k], Vector3(coords[3 * k], coords[3 * k + 1],
                                                                          coords[3 * k + 2]),
                                                     usePseudoNormal != 0, useMeshNormal != 0);
        out[3 * k] = nn.x; out[3 * k + 1] = nn.y; out[3 * k + 2] = nn.z;
    }
}

// ComputeEdgeNormal (ClosestPointOnMesh.cpp:483-497) of edge e at parameter
// params[k]; edges are the mesh's after ConstructEdges. Returns the edge count.
int ref_edge_normals(const float* verts, int nv, const unsigned* idx, int ni, const int* edges, const float* params,
                     int n, float* out)
{
    Geometry::Mesh m;
    make_mesh(m, verts, nv, idx, ni);
    for (int k = 0; k < n; k++) {
        Vector3 nn = Geometry::ComputeEdgeNormal(m, edges[k], params[k]);
        out[3 * k] = nn.x; out[3 * k + 1] = nn.y; out[3 * k + 2] = nn.z;
    }
    return (int)m.edges.size();
}

// MeshClosestPoints::ClosestPoints on the cloth's own mesh
// (Geometry/ClosestPointsBetweenMeshes.cpp:538-592), the live self path of
// ClothModel::MeshCollisions (Physics/ClothCollision.cpp:89-105), over a mesh
// prepared as ClothPatch::Init prepares it (Physics/ClothPatch.cpp:232-235).
// Call with one OpenMP thread (ref_set_threads) for vertex and edge order.
// Per VT info: ids (vtx, tri, region, regionType, feature) and data (distance,
// closest point 3, barycentric 3, normal 3, side). Per EE info: ids (edge1,
// edge, region) and data (distance, coordsSegm 2, coordsMesh 2, closest point
// on the segment 3, on the mesh 3, normal 3, side). counts = (VT, EE).
void ref_live_self(const float* verts, int nv, const unsigned* idx, int ni, float radTol, int flags, int* vtIds,
                   float* vtData, int vtCap, int* eeIds, float* eeData, int eeCap, int* counts)
{
    Geometry::Mesh m;
    m.vertices.resize(nv);
    for (int i = 0; i < nv; i++) m.vertices[i] = Vector3(verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]);
    m.indices.assign(idx, idx + ni);
    m.ComputeNormals();
    m.ConstructEdges();
    m.ConstructVertexOneRings();
    m.ConstructEdgeOneRings();
    Geometry::MeshClosestPoints mcp;
    mcp.SetParams(flags);
    mcp.ClosestPoints(m, nullptr, radTol);
    counts[0] = (int)mcp.mVertexInfos.size();
    counts[1] = (int)mcp.mEdgeInfos.size();
    for (int k = 0; k < counts[0] && k < vtCap; k++) {
        const Geometry::ClosestTriangleToPoint& r = mcp.mVertexInfos[k];
        const int ids[5] = {r.vtx, r.tri, r.region, r.regionType, r.feature};
        const float d[11] = {r.distance, r.closestPtMesh.x, r.closestPtMesh.y, r.closestPtMesh.z, r.baryOnMesh.x,
                             r.baryOnMesh.y, r.baryOnMesh.z, r.normal.x, r.normal.y, r.normal.z, r.side};
        memcpy(vtIds + 5 * k, ids, sizeof(ids));
        memcpy(vtData + 11 * k, d, sizeof(d));
    }
    for (int k = 0; k < counts[1] && k < eeCap; k++) {
        const Geometry::ClosestEdgeToSegment& r = mcp.mEdgeInfos[k];
        const int ids[3] = {r.edge1, r.edge, r.region};
        const float d[15] = {r.distance, r.coordsSegm.x, r.coordsSegm.y, r.coordsMesh.x, r.coordsMesh.y,
                             r.closestPtSegment.x, r.closestPtSegment.y, r.closestPtSegment.z, r.closestPtMesh.x,
                             r.closestPtMesh.y, r.closestPtMesh.z, r.normal.x, r.normal.y, r.normal.z, r.side};
        memcpy(eeIds + 3 * k, ids, sizeof(ids));
        memcpy(eeData + 15 * k, d, sizeof(d));
    }
}

// OpenMP threads for the reference's parallel loops
void ref_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

// The loop bodies of MeshClosestPoints::ClosestPoints (self)
// (ClosestPointsBetweenMeshes.cpp:546-562 and 573-590) for the listed vertices
// and edges only, in the same OpenMP structure (parallel for, appends under a
// critical section), on a mesh prepared as ref_live_self prepares it. Times a
// bounded sample of the live path: seconds = (VT loop, EE loop), the mesh
// preparation excluded; counts = infos kept (VT, EE).
void ref_live_self_sample(const float* verts, int nv, const unsigned* idx, int ni, float radTol, const int* vq,
                          int nvq, const int* eq, int neq, int* counts, double* seconds)
{
    Geometry::Mesh m;
    m.vertices.resize(nv);
    for (int i = 0; i < nv; i++) m.vertices[i] = Vector3(verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]);
    m.indices.assign(idx, idx + ni);
    m.ComputeNormals();
    m.ConstructEdges();
    m.ConstructVertexOneRings();
    m.ConstructEdgeOneRings();
    std::vector<Geometry::ClosestTriangleToPoint> vertexInfos;
    std::vector<Geometry::ClosestEdgeToSegment> edgeInfos;
    auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel for
    for (int k = 0; k < nvq; k++) {
        const int i = vq[k];
        std::vector<Geometry::ClosestTriangleToPoint> results;
        Geometry::ClosestPointsOnMeshToPoint(m.vertices[i], m, results, radTol, i);
        for (size_t j = 0; j < results.size(); j++) {
            results[j].vtx = i;
            if (results[j].tri >= 0 && !Geometry::CheckAdjacency(m, i, results[j].tri, results[j].region)) {
#pragma omp critical
                vertexInfos.push_back(results[j]);
            }
        }
    }
    auto t1 = std::chrono::steady_clock::now();
#pragma omp parallel for
    for (int k = 0; k < neq; k++) {
        const int i = eq[k];
        const Geometry::Mesh::Edge& edge1 = m.edges[i];
        Geometry::ClosestEdgeToSegment info;
        info.edge1 = i;
        Geometry::ClosestPointOnMeshToSegment(m.vertices[edge1.i1], m.vertices[edge1.i2], m, info, i);
        if (!Geometry::CheckEdgeAdjacency(info, m, edge1.i1, edge1.i2, edge1)) {
#pragma omp critical
            edgeInfos.push_back(info);
        }
    }
    auto t2 = std::chrono::steady_clock::now();
    counts[0] = (int)vertexInfos.size();
    counts[1] = (int)edgeInfos.size();
    seconds[0] = std::chrono::duration<double>(t1 - t0).count();
    seconds[1] = std::chrono::duration<double>(t2 - t1).count();
}

// SDF::CreateFromMesh (Geometry/SDF.cpp:58-76; SDF.cpp itself needs Win32
// functions and does not build here): the reference's own tree
// (ComputeMeshTree with triangles, depth 10, tolerance 0.1) and query
// (ClosestPointOnMeshToPointAcc) at every node, with the node position of
// SDF::GetPosAt (SDF.h:64-70) and the steps of SDF::CacheData
// (SDF.cpp:33-38), and the loop's sign rule. Nodes x-fastest.
void ref_sdf_from_mesh(const float* verts, int nv, const unsigned* idx, int ni, int nx, int ny, int nz,
                       const float* box6, float* out)
{
    Geometry::Mesh m;
    make_mesh(m, verts, nv, idx, ni);
    Geometry::AabbTree* tree = Geometry::ComputeMeshTree(m, Geometry::ATF_TRIANGLES, 10, 0.1f);
    const Vector3 mn(box6[0], box6[1], box6[2]), mx(box6[3], box6[4], box6[5]);
    Vector3 steps = mx - mn;
    const Vector3 invRes(1.f / nx, 1.f / ny, 1.f / nz);
    steps.Scale(invRes);
    const int vx = nx + 1, vy = ny + 1, count = vx * vy * (nz + 1);
    for (int k = 0; k < count; k++) {
        const int x = k % vx, w = k / vx, y = w % vy, z = w / vy;
        Vector3 org((float)x, (float)y, (float)z);
        org.Scale(steps);
        org += mn;
        Geometry::ClosestTriangleToPoint info = Geometry::ClosestPointOnMeshToPointAcc(org, m, tree);
        float val = info.distance;
        Vector3 delta = org - info.closestPtMesh;
        if (delta.Dot(info.normal) < 0) val = -val;
        out[k] = val;
    }
    delete tree;
}

}  // extern "C"
