// TEST INFRASTRUCTURE ONLY (see oracle_vec.h).
//
// Closest-point primitives and the static obstacle mesh, restated from the
// reference's geometry code with its operation order:
//   Geometry/Collision3D.cpp:62-141  ClosestPtPointTriangle (Ericson 5.1.5)
//   Geometry/Collision3D.cpp:10-58   ClosestPtSegmSegm (Ericson 5.1.9)
//   Geometry/Collision3D.cpp:298-301 SignedVolume
//   Geometry/Mesh.cpp:10-63          angle-weighted vertex normals
//   Geometry/Mesh.cpp:65-147         unique edges with winged triangles
// These are pinned bit for bit against the reference's compiled Geometry code
// (tests/golden/ref_collision3d.npz, ref_meshes.npz).
#pragma once
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <vector>

#include "oracle_vec.h"

namespace oracle {

enum { TR_VERTEX_A, TR_VERTEX_B, TR_VERTEX_C, TR_EDGE_AB, TR_EDGE_AC, TR_EDGE_BC, TR_FACE_INTERIOR };
enum { ER_EDGE_INTERIOR = 0, ER_VERTEX_P1 = 1, ER_VERTEX_Q1 = 2, ER_VERTEX_P2 = 4, ER_VERTEX_Q2 = 8 };

template <class R>
int closest_pt_point_triangle(V3<R> p, V3<R> a, V3<R> b, V3<R> c, V3<R>& cp, V3<R>& bar)
{
    bar = V3<R>{0, 0, 0};
    V3<R> ab = b - a, ac = c - a, ap = p - a;
    R d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= R(0) && d2 <= R(0)) {
        bar.x = 1;
        cp = a;
        return TR_VERTEX_A;
    }
    V3<R> bp = p - b;
    R d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= R(0) && d4 <= d3) {
        bar.y = 1;
        cp = b;
        return TR_VERTEX_B;
    }
    R vc = d1 * d4 - d3 * d2;
    if (vc <= R(0) && d1 >= R(0) && d3 <= R(0)) {
        bar.y = d1 / (d1 - d3);
        bar.x = R(1) - bar.y;
        cp = a + bar.y * ab;
        return TR_EDGE_AB;
    }
    V3<R> cq = p - c;
    R d5 = dot(ab, cq), d6 = dot(ac, cq);
    if (d6 >= R(0) && d5 <= d6) {
        bar.z = 1;
        cp = c;
        return TR_VERTEX_C;
    }
    R vb = d5 * d2 - d1 * d6;
    if (vb <= R(0) && d2 >= R(0) && d6 <= R(0)) {
        bar.z = d2 / (d2 - d6);
        bar.x = R(1) - bar.z;
        cp = a + bar.z * ac;
        return TR_EDGE_AC;
    }
    R va = d3 * d6 - d5 * d4;
    R d43 = d4 - d3, d56 = d5 - d6;
    if (va <= R(0) && d43 >= R(0) && d56 >= R(0)) {
        bar.z = d43 / (d43 + d56);
        bar.y = R(1) - bar.z;
        cp = b + bar.z * (c - b);
        return TR_EDGE_BC;
    }
    R denom = R(1.f) / (va + vb + vc);
    bar.y = vb * denom;
    bar.z = vc * denom;
    bar.x = R(1.f) - bar.y - bar.z;
    cp = a + ab * bar.y + ac * bar.z;
    return TR_FACE_INTERIOR;
}

// Math::clamp = std::max(lo, std::min(hi, v)) (Math/Utils.h:24-28)
template <class R> inline R clampf(R v, R lo, R hi) { return smax(lo, smin(hi, v)); }

// Geometry/Collision3D.h:125-141: closest point to p on segment [a, b]
template <class R>
V3<R> closest_pt_segm(V3<R> p, V3<R> a, V3<R> b, R& t)
{
    const V3<R> u = b - a, v = p - a;
    t = dot(u, v) / dot(u, u);
    if (t <= R(0)) {
        t = R(0);
        return a;
    }
    if (t >= R(1)) {
        t = R(1);
        return b;
    }
    return a + t * u;
}

// Geometry/Collision3D.h:42-58: barycentric coordinates of p in (a, b, c)
template <class R>
V3<R> barycentric(V3<R> a, V3<R> b, V3<R> c, V3<R> p)
{
    const V3<R> v0 = b - a, v1 = c - a, v2 = p - a;
    const R d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1);
    const R d20 = dot(v2, v0), d21 = dot(v2, v1);
    const R denom = d00 * d11 - d01 * d01;
    V3<R> r;
    r.y = (d11 * d20 - d01 * d21) / denom;
    r.z = (d00 * d21 - d01 * d20) / denom;
    r.x = R(1.f) - r.y - r.z;
    return r;
}

template <class R>
int closest_pt_segm_segm(V3<R> p1, V3<R> q1, V3<R> p2, V3<R> q2, R& s, R& t, V3<R>& c1, V3<R>& c2)
{
    V3<R> d1 = q1 - p1, d2 = q2 - p2, r = p1 - p2;
    R a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
    R c = dot(d1, r), b = dot(d1, d2);
    R denom = a * e - b * b;
    if (denom != R(0)) s = clampf((b * f - c * e) / denom, R(0), R(1));
    else s = 0;
    int region = ER_EDGE_INTERIOR;
    t = (b * s + f) / e;
    if (t < R(0)) {
        region |= ER_VERTEX_P2;
        t = 0;
        s = clampf(-c / a, R(0), R(1));
    } else if (t > R(1)) {
        region |= ER_VERTEX_Q2;
        t = 1;
        s = clampf((b - c) / a, R(0), R(1));
    }
    if (s == R(0)) region |= ER_VERTEX_P1;
    if (s == R(1)) region |= ER_VERTEX_Q1;
    c1 = p1 + s * d1;
    c2 = p2 + t * d2;
    return region;
}

// Geometry/Collision3D.cpp:143-210 (ClosestPtSegmTriangle): nearest of the
// segment against each edge and each end point against the triangle, first
// strictly nearer kept, with the reference's coordinates — including that the
// query of end point q overwrites them whatever its outcome and, when q is
// nearest, hands them p's.
template <class R>
R closest_pt_segm_triangle(V3<R> p, V3<R> q, V3<R> a, V3<R> b, V3<R> c, R& param, V3<R>& bar, V3<R>& closestEdge,
                           V3<R>& closestTri)
{
    R minDist = R(FLT_MAX);
    R s, t;
    V3<R> c1, c2;
    closest_pt_segm_segm(p, q, a, b, s, t, c1, c2);
    R dist1 = length(c1 - c2);
    if (dist1 < minDist) {
        minDist = dist1; closestEdge = c1; closestTri = c2; param = s;
        bar = V3<R>{R(1) - t, t, R(0)};
    }
    closest_pt_segm_segm(p, q, b, c, s, t, c1, c2);
    R dist2 = length(c1 - c2);
    if (dist2 < minDist) {
        minDist = dist2; closestEdge = c1; closestTri = c2; param = s;
        bar = V3<R>{R(0), R(1) - t, t};
    }
    closest_pt_segm_segm(p, q, c, a, s, t, c1, c2);
    R dist3 = length(c1 - c2);
    if (dist3 < minDist) {
        minDist = dist3; closestEdge = c1; closestTri = c2; param = s;
        bar = V3<R>{t, R(0), R(1) - t};
    }
    V3<R> coords, cp4;
    closest_pt_point_triangle(p, a, b, c, cp4, coords);
    R dist4 = length(p - cp4);
    if (dist4 < minDist) {
        minDist = dist4; closestEdge = p; closestTri = cp4; param = 0;
        bar = coords;
    }
    V3<R> cp5;
    closest_pt_point_triangle(q, a, b, c, cp5, bar);
    R dist5 = length(q - cp5);
    if (dist5 < minDist) {
        minDist = dist5; closestEdge = q; closestTri = cp5; param = 1;
        bar = coords;
    }
    return minDist;
}

template <class R>
inline R signed_volume(V3<R> a0, V3<R> a1, V3<R> a2, V3<R> a3)
{
    // triple(u, v, w) = w . (u x v), Math/Vector3.h:289-292
    V3<R> u = a1 - a0, v = a2 - a0, w = a3 - a0;
    return dot(w, cross(u, v));
}

struct MeshEdge {
    uint32_t i1, i2;
    int t1 = -1, t2 = -1;
    bool swapped = false;  // the first triangle runs the edge i2 -> i1
    int orig = -1;         // index of the first triangle side, 3 t + k
};

// Geometry/Mesh.cpp:65-147 on an index list (std::sort as the reference, so
// the first of two duplicate sides, which sets t1 and swapped, is the
// reference's). map, if given, receives per triangle side 3 t + k the unique
// edge: k = 0 is (a, b), 1 is (b, c), 2 is (c, a) (Mesh::Triangle::e).
inline std::vector<MeshEdge> mesh_edges(const std::vector<uint32_t>& idx, std::vector<int>* map = nullptr)
{
    std::vector<MeshEdge> edges;
    const int ne = (int)idx.size();
    if (ne == 0) return edges;
    std::vector<MeshEdge> dup(ne);
    for (int i = 0; i < ne; i++) {
        uint32_t a = idx[i];
        uint32_t b = ((i + 1) % 3 == 0) ? idx[i - 2] : idx[i + 1];
        dup[i].swapped = a > b;
        if (a > b) std::swap(a, b);
        dup[i].i1 = a; dup[i].i2 = b; dup[i].t1 = i / 3; dup[i].orig = i;
    }
    std::sort(dup.begin(), dup.end(), [](const MeshEdge& e1, const MeshEdge& e2) {
        return e1.i1 < e2.i1 || (e1.i1 == e2.i1 && e1.i2 < e2.i2);
    });
    if (map) map->assign(ne, -1);
    MeshEdge cur = dup[0];
    if (map) (*map)[cur.orig] = 0;
    for (size_t i = 1; i < dup.size(); i++) {
        if (dup[i].i1 == dup[i - 1].i1 && dup[i].i2 == dup[i - 1].i2) {
            cur.t2 = dup[i].t1;
            if (map) (*map)[dup[i].orig] = (int)edges.size();
            continue;
        }
        edges.push_back(cur);
        cur = dup[i];
        if (map) (*map)[cur.orig] = (int)edges.size();
    }
    edges.push_back(cur);
    return edges;
}

// Geometry/Mesh.cpp:10-63: unit face normals (tn) and angle-weighted unit
// vertex normals (vn; zero for a vertex in no triangle)
inline void mesh_face_vertex_normals(const std::vector<V3<float>>& v, const std::vector<uint32_t>& idx,
                                     std::vector<V3<float>>& vn, std::vector<V3<float>>& tn)
{
    const size_t nt = idx.size() / 3;
    vn.assign(v.size(), V3<float>{0, 0, 0});
    tn.assign(nt, V3<float>{0, 0, 0});
    for (size_t t = 0; t < nt; t++) {
        const uint32_t i1 = idx[3 * t], i2 = idx[3 * t + 1], i3 = idx[3 * t + 2];
        const V3<float> v1 = v[i1], v2 = v[i2], v3 = v[i3];
        V3<float> n = cross(v2 - v1, v3 - v1);
        normalize(n);
        tn[t] = n;
        V3<float> e1 = v2 - v1, e2 = v3 - v2, e3 = v1 - v3;
        normalize(e1);
        normalize(e2);
        normalize(e3);
        float c1 = -dot(e1, e3), c2 = -dot(e1, e2), c3 = -dot(e2, e3);
        float s1 = length(cross(e1, -e3)), s2 = length(cross(e2, -e1)), s3 = length(cross(e3, -e1));
        float t1 = std::fabs(c1) > 1 ? std::fabs((float)std::atan2(s1, c1)) : std::acos(c1);
        float t2 = std::fabs(c2) > 1 ? std::fabs((float)std::atan2(s2, c2)) : std::acos(c2);
        float t3 = std::fabs(c3) > 1 ? std::fabs((float)std::atan2(s3, c3)) : std::acos(c3);
        vn[i1] += t1 * n;
        vn[i2] += t2 * n;
        vn[i3] += t3 * n;
    }
    for (auto& n : vn)
        if (length(n) != 0.f) normalize(n);
}

// Uniform grid over boxes for exact closed-interval overlap queries (a speed-up
// of the oracle's exhaustive searches, used to time it at large sizes; it
// returns exactly the boxes an exhaustive scan would). Each box is stored in
// the cell of its minimum corner and the cell size is at least the largest box
// extent, so a box overlapping the query box has its minimum corner in the
// query's cells widened by one cell on each side. Cells are a dense CSR array
// over the boxes' bounds.
struct BoxGrid {
    double cell = 1.0, ox = 0.0, oy = 0.0, oz = 0.0;
    long nx = 0, ny = 0, nz = 0;
    std::vector<V3<float>> lo, hi;
    std::vector<int> start, items;

    long ix(double v, double o) const { return (long)std::floor((v - o) / cell); }

    // boxes with skip[i] set are left out
    void build(const std::vector<V3<float>>& blo, const std::vector<V3<float>>& bhi,
               const std::vector<char>* skip = nullptr)
    {
        lo = blo;
        hi = bhi;
        double ext = 0.0;
        double mx = 1e300, my = 1e300, mz = 1e300, Mx = -1e300, My = -1e300, Mz = -1e300;
        size_t count = 0;
        for (size_t i = 0; i < lo.size(); i++) {
            if (skip && (*skip)[i]) continue;
            ext = std::max({ext, double(hi[i].x) - lo[i].x, double(hi[i].y) - lo[i].y, double(hi[i].z) - lo[i].z});
            mx = std::min(mx, double(lo[i].x));
            my = std::min(my, double(lo[i].y));
            mz = std::min(mz, double(lo[i].z));
            Mx = std::max(Mx, double(lo[i].x));
            My = std::max(My, double(lo[i].y));
            Mz = std::max(Mz, double(lo[i].z));
            count++;
        }
        nx = ny = nz = 0;
        start.assign(1, 0);
        items.clear();
        if (!count) return;
        // no finer than the spacing of points spread over a surface of the
        // bounds' size (boxes of zero extent), and at most 16M cells
        const double span = std::max({Mx - mx, My - my, Mz - mz});
        cell = std::max(ext, span / std::sqrt(double(count)));
        if (!(cell > 0.0)) cell = 1.0;
        ox = mx;
        oy = my;
        oz = mz;
        for (;;) {
            nx = ix(Mx, ox) + 1;
            ny = ix(My, oy) + 1;
            nz = ix(Mz, oz) + 1;
            if (double(nx) * double(ny) * double(nz) <= double(1 << 24)) break;
            cell *= 1.25;
        }
        start.assign(size_t(nx * ny * nz) + 1, 0);
        auto at = [&](size_t i) {
            return size_t((ix(lo[i].x, ox) * ny + ix(lo[i].y, oy)) * nz + ix(lo[i].z, oz));
        };
        for (size_t i = 0; i < lo.size(); i++)
            if (!(skip && (*skip)[i])) start[at(i) + 1]++;
        for (size_t c = 1; c < start.size(); c++) start[c] += start[c - 1];
        items.resize(count);
        std::vector<int> fill(start.begin(), start.end() - 1);
        for (size_t i = 0; i < lo.size(); i++)
            if (!(skip && (*skip)[i])) items[fill[at(i)]++] = (int)i;
    }

    // the indices of the boxes overlapping [qlo, qhi], ascending
    void query(V3<float> qlo, V3<float> qhi, std::vector<int>& out) const
    {
        out.clear();
        if (!nx) return;
        const long x0 = std::max(0L, ix(qlo.x, ox) - 1), x1 = std::min(nx - 1, ix(qhi.x, ox));
        const long y0 = std::max(0L, ix(qlo.y, oy) - 1), y1 = std::min(ny - 1, ix(qhi.y, oy));
        const long z0 = std::max(0L, ix(qlo.z, oz) - 1), z1 = std::min(nz - 1, ix(qhi.z, oz));
        for (long x = x0; x <= x1; x++)
            for (long y = y0; y <= y1; y++) {
                const size_t row = size_t((x * ny + y) * nz);
                for (int k = start[row + z0], e = z1 >= z0 ? start[row + z1 + 1] : k; k < e; k++) {
                    const int i = items[k];
                    if (!(qhi.x < lo[i].x || qhi.y < lo[i].y || qhi.z < lo[i].z || hi[i].x < qlo.x ||
                          hi[i].y < qlo.y || hi[i].z < qlo.z))
                        out.push_back(i);
                }
            }
        std::sort(out.begin(), out.end());
    }
};

// A static triangle mesh obstacle in float (the reference's Mesh stores
// Vector3 = FP32): vertices, indices, angle-weighted vertex normals, face
// normals, unique edges and their unweighted-average normals.
struct ObstacleMesh {
    std::vector<V3<float>> v;
    std::vector<uint32_t> idx;
    std::vector<V3<float>> vn, tn, en;
    std::vector<MeshEdge> edges;
    BoxGrid triGrid, vtxGrid, edgeGrid;  // feature boxes, for the accelerated queries

    void build()
    {
        const size_t nt = idx.size() / 3;
        mesh_face_vertex_normals(v, idx, vn, tn);
        edges = mesh_edges(idx);
        en.resize(edges.size());
        for (size_t e = 0; e < edges.size(); e++) {
            V3<float> n = tn[edges[e].t1];
            if (edges[e].t2 >= 0) {
                n = tn[edges[e].t1] + tn[edges[e].t2];
                normalize(n);
            }
            en[e] = n;
        }
        std::vector<V3<float>> lo, hi;
        auto box = [&](std::initializer_list<uint32_t> ids) {
            V3<float> a{1e30f, 1e30f, 1e30f}, b{-1e30f, -1e30f, -1e30f};
            for (uint32_t i : ids) {
                a = V3<float>{std::min(a.x, v[i].x), std::min(a.y, v[i].y), std::min(a.z, v[i].z)};
                b = V3<float>{std::max(b.x, v[i].x), std::max(b.y, v[i].y), std::max(b.z, v[i].z)};
            }
            lo.push_back(a);
            hi.push_back(b);
        };
        for (size_t t = 0; t < nt; t++) box({idx[3 * t], idx[3 * t + 1], idx[3 * t + 2]});
        triGrid.build(lo, hi);
        lo.clear(); hi.clear();
        for (size_t i = 0; i < v.size(); i++) box({(uint32_t)i});
        vtxGrid.build(lo, hi);
        lo.clear(); hi.clear();
        for (const MeshEdge& e : edges) box({e.i1, e.i2});
        edgeGrid.build(lo, hi);
    }
};

}  // namespace oracle
