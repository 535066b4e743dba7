// TEST INFRASTRUCTURE ONLY — the CPU oracle for the GPU cloth substep.
// Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg;
// never by the product library.
//
// A restatement of SolidCloth's position-based-dynamics substep, templated on
// the real type (float = parity oracle, double = FP64 rerun). Each function
// cites the reference lines it follows. Where the reference's result depends
// on unspecified behaviour the restatement fixes a canonical choice:
//   * std::sort in Mesh::ConstructEdges (Geometry/Mesh.cpp:110) is unstable,
//     so which of an interior edge's two triangles becomes t1 depends on the
//     standard library. The restatement makes the same std::sort call on the
//     same keys in the same initial order, so with libstdc++ it reproduces the
//     compiled reference exactly (pinned by tests/golden).
//   * Contacts are generated in collidable-major, vertex-minor order, i.e. the
//     reference's order with OMP_NUM_THREADS=1 (ClothCollision.cpp:440-457).
//     The detection loops run on OpenMP threads as the reference's do, in
//     blocks whose outputs are appended in block order, so the lists are the
//     same for every thread count; the solver is serial, as ClothLib's is.
//   * Triangle::invDet is left uninitialised by the reference when det == 0
//     (ClothModel.cpp:67-68); here it is 0.
//   * The Gauss-Seidel loop may be given a constraint permutation (the GPU's
//     colour order) and a fixed-iteration switch that removes the early exit
//     (ClothPBD.cpp:356-358), which otherwise fires after one iteration in FEM
//     mode because the link error stays 0.
#include "oracle_geom.h"
#include "oracle_vec.h"

#include <omp.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

namespace oracle {

enum { CF_VERTICES = 1, CF_TRIANGLES = 2, CF_SELF = 4, CF_WALLS = 8, CF_EDGES = 16 };
enum { CT_WALLS = 1, CT_SPHERE = 3, CT_MESH = 4, CT_CAPSULE = 6, CT_SDF = 7 };

// Runs body(begin, end, part) over consecutive blocks of [0, n) on OpenMP
// threads, each block writing its own part, and returns the parts in block
// order: appending them in turn gives the serial loop's output for any
// thread count.
template <class Part, class F>
std::vector<Part> ordered_blocks(size_t n, F body)
{
    const size_t blocks = std::min<size_t>(n, 1024);
    std::vector<Part> parts(blocks);
#pragma omp parallel for schedule(dynamic, 1) if (n >= 512)
    for (long b = 0; b < (long)blocks; b++) body(n * b / blocks, n * (b + 1) / blocks, parts[b]);
    return parts;
}

template <class T>
void append_parts(std::vector<T>& out, std::vector<std::vector<T>>& parts)
{
    for (auto& p : parts) out.insert(out.end(), p.begin(), p.end());
}

struct Params {
    // defaults: Physics/ClothModel.cpp:10-36
    float unit = 10.f, thickness = 0.3f, tolerance = 0.2f, shear = 5e8f, bend = 5.f, stretch = 2e10f;
    float poisson = 0.5f, friction = 0.4f, area = 10.f, drag = 1.f, selfShrink = 0.9f, selfCollThreshold = 0.003f;
    // continuous self-collision thresholds, the reference's absolute constants
    // (SelfCollisionDetector.cpp:235, 296, 313-317, 446): coplanarity volume,
    // bisection stop, vertex-triangle distance
    float ccdVolume = 0.001f, ccdBisect = 0.01f, ccdVtDistance = 0.1f;
    int iterations = 50, collFlags = 0, bendFreq = 2, numSteps = 1;
    // 1: keep the last detection's self candidate pairs (selfCandVT / EE)
    int recordCandidates = 0;
    // self edge-edge displacement weights: 0 as ClothCollisionHandler.cpp:
    // 531-534 writes them (the first edge's ends moved by each other's
    // barycentric weight, likewise the second's), 1 each end moved by its own
    int selfEdgeWeights = 0;
    // 1: broad phases through uniform grids (same results as the exhaustive
    // scans, for timing the oracle at large sizes)
    int accel = 0;
    // Physics/ClothModel.h:23-30 SolverType: 0 Jacobi, 1 Gauss-Seidel
    // (ClothModel.cpp:24 default), any other the conjugate-residual solver
    int solver = 1;
    bool dihedral = true, fem = false, fixedIterations = false;
};

struct Stats {
    int iterations = 0;
    float bendError = 0, contactError = 0, selfVTError = 0, selfEEError = 0, linkError = 0;
    int selfEECount = 0;
};

template <class R>
struct Tri {
    uint32_t i1, i2, i3;
    R du1, du2, dv1, dv2, det, invDet, lu0, lv0, area, lu2, lv2;
};

struct Bend {
    uint32_t i1, i2, i3, i4;
};

// Physics/ClothTypes.h:94-99 (Link: reference length and stiffness)
template <class R>
struct LinkC {
    uint32_t i1, i2;
    R len, stiffness;
};

struct Edge {
    uint32_t i1, i2;
    int t1 = -1, t2 = -1;
};

template <class R>
struct Contact {
    uint32_t idx;
    int coll;  // index of the collidable that produced it
    V3<R> point, normal, vel;
    R lambda = 0, lambdaF = 0, depth = 0;
    int feature = -1;  // obstacle triangle (mesh contacts)
    int mesh = -1;     // obstacle mesh index, -1 for walls and SDFs
};

// Physics/ClothTypes.h:110-120: an obstacle edge against a cloth edge, an
// obstacle vertex against a cloth triangle
template <class R>
struct EdgeContact {
    uint32_t i1, i2;
    int clothEdge, feature, mesh;
    V3<R> point, normal;
    R w1, w2, lambda = 0, lambdaF = 0;
};

template <class R>
struct TriContact {
    uint32_t i1, i2, i3;
    int clothTri, feature, mesh;
    V3<R> point, normal;
    R w1, w2, w3, lambda = 0, lambdaF = 0;
};

// Physics/ClothTypes.h:122-129: a vertex-triangle (i4 against i1 i2 i3) or
// edge-edge (i1 i2 against i3 i4) self contact. `key` is (triangle << 32 |
// vertex) or (edge1 << 32 | edge2), the canonical list order.
template <class R>
struct SelfContact {
    uint32_t i1, i2, i3, i4;
    R w1 = 0, w2 = 0, w3 = 0;
    V3<R> normal{0, 0, 0};
    R side = R(FLT_MAX), lambda = 0, error = 0;
    uint64_t key = 0;
};

// Node-sampled grid SDF: Geometry/SDF.h:72-155, Geometry/SDF.cpp:10-40,210-256
template <class R>
struct GridSDF {
    int nc[3];  // cells per side
    int nv[3];  // values per side
    V3<R> bmin, bmax, steps;
    std::vector<R> vals;
    std::vector<V3<R>> grads;

    int index(int x, int y, int z) const { return x + y * nv[0] + z * nv[0] * nv[1]; }
    bool oob(int x, int y, int z) const
    {
        return x > nc[0] || y > nc[1] || z > nc[2] || x < 0 || y < 0 || z < 0;
    }
    R value(int x, int y, int z) const { return oob(x, y, z) ? R(0) : vals[index(x, y, z)]; }
    V3<R> grad(int x, int y, int z) const { return oob(x, y, z) ? V3<R>{0, 0, 0} : grads[index(x, y, z)]; }

    void create(int nx, int ny, int nz, V3<R> mn, V3<R> mx)
    {
        nc[0] = nx; nc[1] = ny; nc[2] = nz;
        nv[0] = nx + 1; nv[1] = ny + 1; nv[2] = nz + 1;
        bmin = mn; bmax = mx;
        vals.assign((size_t)nv[0] * nv[1] * nv[2], R(0));
        // SDF.cpp:37-39: steps = extent scaled by the reciprocal cell counts
        steps = bmax - bmin;
        V3<R> inv{R(1.f) / R(nx), R(1.f) / R(ny), R(1.f) / R(nz)};
        steps.x *= inv.x; steps.y *= inv.y; steps.z *= inv.z;
    }

    // SDF.cpp:210-256 forward differences, last layer copied
    void compute_gradient()
    {
        grads.assign(vals.size(), V3<R>{0, 0, 0});
        auto set = [&](int x, int y, int z, V3<R> g) {
            if (!oob(x, y, z)) grads[index(x, y, z)] = g;
        };
        for (int x = 0; x < nc[0]; x++) {
            for (int y = 0; y < nc[1]; y++) {
                for (int z = 0; z < nc[2]; z++) {
                    R v = value(x, y, z);
                    R dfdx = (value(x + 1, y, z) - v) / steps.x;
                    R dfdy = (value(x, y + 1, z) - v) / steps.y;
                    R dfdz = (value(x, y, z + 1) - v) / steps.z;
                    set(x, y, z, V3<R>{dfdx, dfdy, dfdz});
                }
                int z = nc[2];
                set(x, y, z, grad(x, y, z - 1));
            }
            int y = nc[1];
            for (int z = 0; z <= nc[2]; z++) set(x, y, z, grad(x, y - 1, z));
        }
        int x = nc[0];
        for (int y = 0; y <= nc[1]; y++)
            for (int z = 0; z <= nc[2]; z++) set(x, y, z, grad(x - 1, y, z));
    }

    // SDF.h:105-117. The reference truncates floorf() to int and treats a
    // negative or too-large cell as outside; NaN and out-of-int-range values
    // land outside as well (x86 cvttss2si yields INT_MIN for them).
    bool coords(V3<R> p, int& x, int& y, int& z, R& fx, R& fy, R& fz) const
    {
        R rx = (p.x - bmin.x) / steps.x;
        R ry = (p.y - bmin.y) / steps.y;
        R rz = (p.z - bmin.z) / steps.z;
        R flx = std::floor(rx), fly = std::floor(ry), flz = std::floor(rz);
        if (!(flx >= R(0) && flx <= R(nc[0]) && fly >= R(0) && fly <= R(nc[1]) && flz >= R(0) && flz <= R(nc[2])))
            return false;
        x = (int)flx; y = (int)fly; z = (int)flz;
        fx = std::fabs(rx - R(x));
        fy = std::fabs(ry - R(y));
        fz = std::fabs(rz - R(z));
        return true;
    }

    // SDF.h:138-155
    R value_at(V3<R> p) const
    {
        int x, y, z;
        R fx, fy, fz;
        if (!coords(p, x, y, z, fx, fy, fz)) return R(FLT_MAX);
        const R one = 1;
        return (one - fx) * (one - fy) * (one - fz) * value(x, y, z) +
               (one - fx) * (one - fy) * fz * value(x, y, z + 1) +
               (one - fx) * fy * (one - fz) * value(x, y + 1, z) +
               (one - fx) * fy * fz * value(x, y + 1, z + 1) +
               fx * (one - fy) * (one - fz) * value(x + 1, y, z) +
               fx * (one - fy) * fz * value(x + 1, y, z + 1) +
               fx * fy * (one - fz) * value(x + 1, y + 1, z) +
               fx * fy * fz * value(x + 1, y + 1, z + 1);
    }

    // SDF.h:119-136
    V3<R> grad_at(V3<R> p) const
    {
        int x, y, z;
        R fx, fy, fz;
        if (!coords(p, x, y, z, fx, fy, fz)) return V3<R>{0, 0, 0};
        const R one = 1;
        return (one - fx) * (one - fy) * (one - fz) * grad(x, y, z) +
               (one - fx) * (one - fy) * fz * grad(x, y, z + 1) +
               (one - fx) * fy * (one - fz) * grad(x, y + 1, z) +
               (one - fx) * fy * fz * grad(x, y + 1, z + 1) +
               fx * (one - fy) * (one - fz) * grad(x + 1, y, z) +
               fx * (one - fy) * fz * grad(x + 1, y, z + 1) +
               fx * fy * (one - fz) * grad(x + 1, y + 1, z) +
               fx * fy * fz * grad(x + 1, y + 1, z + 1);
    }

    // Physics/ClothCollisionSDF.h:130-139 (GridSDF::QueryPoint)
    R query_point(V3<R> p, V3<R>& cp, V3<R>& n) const
    {
        R val = value_at(p);
        V3<R> g = grad_at(p);
        normalize(g);
        n = g;
        cp = p - val * g;
        return val;
    }

    // Physics/ClothCollisionSDF.h:148-205 (GridSDF::QueryTriangle): Frank-Wolfe
    // iterations from 0.33 (a + b + c) with step 2 / (i + 2), stopped when the
    // value changes by less than 1e-4 relative, or when the point leaves the grid.
    R query_triangle(V3<R> a, V3<R> b, V3<R> c, V3<R>& closest, V3<R>& n, V3<R>& bar) const
    {
        V3<R> cp = R(0.33f) * ((a + b) + c);
        R distOld = 0;
        for (int i = 1; i < 100; i++) {
            const R dist = value_at(cp);
            if (dist == R(FLT_MAX)) return R(FLT_MAX);
            const V3<R> g = grad_at(cp);
            const R dotA = dot(a, g), dotB = dot(b, g), dotC = dot(c, g);
            R d = dotA;
            V3<R> sv = a;
            if (dotB < d) { d = dotB; sv = b; }
            if (dotC < d) { d = dotC; sv = c; }
            if (distOld > R(0)) {
                const R rel = std::fabs(distOld - dist) / distOld;
                if (rel < R(0.0001f)) break;
            }
            distOld = dist;
            const R alpha = R(2.f) / (R(float(i)) + R(2.f));
            cp = cp + alpha * (sv - cp);
        }
        const R val = value_at(cp);
        V3<R> g = grad_at(cp);
        normalize(g);
        n = g;
        closest = cp - val * g;
        bar = barycentric(a, b, c, cp);
        return val;
    }
};

template <class R>
struct Collidable {
    int type;
    V3<R> bmin, bmax;  // walls box
    V3<R> center;      // sphere
    R radius = 0;      // sphere, capsule
    V3<R> axisP, axisQ;  // capsule axis ends
    int sdf = -1;      // index into sdfs
    int mesh = -1;     // index into meshes
};

template <class R>
class Cloth {
public:
    Params prm;
    Stats stats;
    std::vector<V3<R>> pos, prev, vel;
    std::vector<R> invMass;
    std::vector<R> uvx, uvy;
    std::vector<Tri<R>> tris;
    std::vector<Bend> bends;
    std::vector<LinkC<R>> links;
    std::vector<Edge> edges;
    std::vector<uint32_t> meshIndices;
    std::vector<Contact<R>> contacts;
    std::vector<Collidable<R>> world;
    std::vector<GridSDF<R>> sdfs;
    std::vector<ObstacleMesh> meshes;
    std::vector<std::pair<int, int>> selfCandVT, selfCandEE;  // (triangle, vertex), (edge1, edge2)
    std::vector<EdgeContact<R>> edgeContacts;
    std::vector<TriContact<R>> triContacts;
    std::vector<int> triOrder, bendOrder, linkOrder;
    // processing order of edge and triangle contacts by cloth edge / triangle
    // (empty: ascending index, the reference's list order)
    std::vector<int> edgeContactRank, triContactRank;
    std::vector<SelfContact<R>> selfTris, selfEdges;
    // processing order of the next substep's self contacts, as keys (empty:
    // the canonical key order); consumed by substep_solve
    std::vector<uint64_t> selfTriOrder, selfEdgeOrder;

    void clear()
    {
        pos.clear(); prev.clear(); vel.clear(); invMass.clear(); uvx.clear(); uvy.clear();
        tris.clear(); bends.clear(); links.clear(); edges.clear(); meshIndices.clear(); contacts.clear();
        triOrder.clear(); bendOrder.clear(); linkOrder.clear();
    }

    // Physics/ClothModel.inl:12-31 (AddLink)
    void add_link(uint32_t i1, uint32_t i2, float stiffness)
    {
        if (invMass[i1] == R(0) && invMass[i2] == R(0)) return;
        if (stiffness == 0.f) return;
        links.push_back(LinkC<R>{i1, i2, length(pos[i1] - pos[i2]), R(stiffness)});
    }

    // Physics/ClothModel.cpp:54-92
    void add_triangle(uint32_t i1, uint32_t i2, uint32_t i3)
    {
        add_triangle_uv(i1, i2, i3, uvx[i1], uvy[i1], uvx[i2], uvy[i2], uvx[i3], uvy[i3]);
    }

    void add_triangle_uv(uint32_t i1, uint32_t i2, uint32_t i3, R u1x, R u1y, R u2x, R u2y, R u3x, R u3y)
    {
        Tri<R> t;
        t.i1 = i1; t.i2 = i2; t.i3 = i3;
        R d1x = u2x - u1x, d1y = u2y - u1y;
        R d2x = u3x - u1x, d2y = u3y - u1y;
        t.du1 = d1x; t.du2 = d2x; t.dv1 = d1y; t.dv2 = d2y;
        t.det = t.du1 * t.dv2 - t.du2 * t.dv1;
        t.invDet = (t.det != R(0)) ? R(1.f) / t.det : R(0);
        V3<R> dx1 = pos[i2] - pos[i1];
        V3<R> dx2 = pos[i3] - pos[i1];
        V3<R> wu = t.invDet * (t.dv2 * dx1 - t.dv1 * dx2);
        V3<R> wv = t.invDet * (-t.du2 * dx1 + t.du1 * dx2);
        t.lu0 = length(wu);
        t.lv0 = length(wv);
        t.area = R(0.5f) * t.det;
        t.lu2 = t.lu0 * t.lu0;
        t.lv2 = t.lv0 * t.lv0;
        tris.push_back(t);
        meshIndices.push_back(i1); meshIndices.push_back(i2); meshIndices.push_back(i3);
    }

    // Geometry/Mesh.cpp:65-147 (unique sorted edges with winged triangles)
    void construct_edges()
    {
        edges.clear();
        int ne = (int)meshIndices.size();
        std::vector<Edge> dup(ne);
        for (int i = 0; i < ne; i++) {
            uint32_t a = meshIndices[i];
            uint32_t b = ((i + 1) % 3 == 0) ? meshIndices[i - 2] : meshIndices[i + 1];
            if (a > b) std::swap(a, b);
            dup[i].i1 = a; dup[i].i2 = b; dup[i].t1 = i / 3;
        }
        std::sort(dup.begin(), dup.end(), [](const Edge& e1, const Edge& e2) {
            return e1.i1 < e2.i1 || (e1.i1 == e2.i1 && e1.i2 < e2.i2);
        });
        if (dup.empty()) return;
        Edge cur = dup[0];
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

    static uint32_t opposite(const Tri<R>& t, uint32_t a, uint32_t b)
    {
        // ClothTypes.h:54-61
        if (t.i1 != a && t.i1 != b) return t.i1;
        if (t.i2 != a && t.i2 != b) return t.i2;
        return t.i3;
    }

    // Physics/ClothModel.cpp:94-126 (dihedral branch; theta0 is unused by the solver)
    void add_bend(const Edge& e)
    {
        if (e.t2 < 0) return;
        Bend b;
        b.i1 = e.i1; b.i2 = e.i2;
        b.i3 = opposite(tris[e.t1], e.i1, e.i2);
        b.i4 = opposite(tris[e.t2], e.i1, e.i2);
        bends.push_back(b);
    }

    // Physics/ClothPatch.cpp:43-133 (ClothPatch::Init from a mesh): welding
    // by the quadratic squared-distance-0 search, triangles from every third
    // corner (the last one after the loop), links / edges / bends from the
    // sorted corner edges, the last triangle's closing edge left out as the
    // reference's corner loop stops one short. map receives vertex -> particle.
    void init_mesh(const std::vector<V3<float>>& v, const std::vector<uint32_t>& idx, const float* uv,
                   const float* colorsX, V3<float> off, bool attached, std::vector<int>& map)
    {
        clear();
        const size_t nv = v.size(), n = idx.size();
        map.assign(nv, -1);
        for (size_t i = 0; i < nv; i++) {
            if (map[i] >= 0) continue;
            const int np = (int)pos.size();
            for (size_t j = i + 1; j < nv; j++) {
                const V3<float> d = v[i] - v[j];
                if (dot(d, d) == 0.f) map[j] = np;
            }
            map[i] = np;
            const V3<float> p = v[i] + off;
            pos.push_back(V3<R>{R(p.x), R(p.y), R(p.z)});
            prev.push_back(pos.back());
            vel.push_back(V3<R>{0, 0, 0});
            invMass.push_back((attached && colorsX && colorsX[i] != 0.f) ? R(0) : R(1));
            uvx.push_back(uv ? R(uv[2 * i]) : R(0));
            uvy.push_back(uv ? R(uv[2 * i + 1]) : R(0));
        }
        auto tri = [&](size_t a, size_t b, size_t c, uint32_t i1, uint32_t i2, uint32_t i3) {
            if (uv)
                add_triangle_uv(i1, i2, i3, R(uv[2 * a]), R(uv[2 * a + 1]), R(uv[2 * b]), R(uv[2 * b + 1]),
                                R(uv[2 * c]), R(uv[2 * c + 1]));
            else
                add_triangle(i1, i2, i3);
        };
        std::vector<Edge> dup(n - 1);
        for (size_t i = 0; i + 1 < n; i++) {
            uint32_t i1 = (uint32_t)map[idx[i]], i2 = (uint32_t)map[idx[i + 1]];
            if ((i + 1) % 3 == 0) {
                i2 = (uint32_t)map[idx[i - 2]];
                tri(idx[i - 2], idx[i - 1], idx[i], i2, (uint32_t)map[idx[i - 1]], i1);
            }
            if (i1 > i2) std::swap(i1, i2);
            Edge e;
            e.i1 = i1; e.i2 = i2; e.t1 = (int)(i / 3);
            dup[i] = e;
        }
        tri(idx[n - 3], idx[n - 2], idx[n - 1], (uint32_t)map[idx[n - 3]], (uint32_t)map[idx[n - 2]],
            (uint32_t)map[idx[n - 1]]);
        std::sort(dup.begin(), dup.end(), [](const Edge& a, const Edge& b) {
            return a.i1 < b.i1 || (a.i1 == b.i1 && a.i2 < b.i2);
        });
        const float stretch = prm.stretch;
        add_link(dup[0].i1, dup[0].i2, stretch);
        Edge cur = dup[0];
        for (size_t i = 1; i < dup.size(); i++) {
            if (dup[i].i1 == dup[i - 1].i1 && dup[i].i2 == dup[i - 1].i2) {
                cur.t2 = dup[i].t1;
                continue;
            }
            add_link(dup[i].i1, dup[i].i2, stretch);
            edges.push_back(cur);
            add_bend(cur);
            cur = dup[i];
        }
        edges.push_back(cur);
        add_bend(cur);
    }

    // Physics/ClothPatch.cpp:135-246 (grid branch)
    void init_grid(int divX, int divY, float incF, V3<float> offF, bool horizontal, bool attached, float inclineF)
    {
        clear();
        const int n = divX * divY;
        pos.assign(n, V3<R>{0, 0, 0}); prev = pos; vel = pos;
        invMass.assign(n, R(1)); uvx.assign(n, R(0)); uvy.assign(n, R(0));
        const R inc = incF, incline = inclineF, scale = R(1.f);
        const V3<R> offset{R(offF.x), R(offF.y), R(offF.z)};
        R x = R(-0.5f) * R(divX) * inc;
        const R startY = x;
        const float stretch = prm.stretch, shear = prm.shear, bend = prm.bend;
        const bool dihedral = prm.dihedral;
        for (int j = 0, base = 0; j < divX; j++, base += divY) {
            pos[base] = horizontal ? (V3<R>{x, 0, startY} + offset) * scale : (V3<R>{x, startY, 0} + offset) * scale;
            prev[base] = pos[base];
            vel[base] = V3<R>{0, 0, 0};
            invMass[base] = 1;
            uvx[base] = x; uvy[base] = 0;
            if (base >= divY) {
                add_link(base - divY, base, stretch);
                add_link(base - divY + 1, base, shear);
                if (base >= 2 * divY && !dihedral) add_link(base - 2 * divY, base, bend);
            }
            for (int i = 1; i < divY; i++) {
                int idx = base + i;
                if (horizontal)
                    pos[idx] = (V3<R>{x, 0, startY + inc * R(i)} + offset) * scale;
                else
                    pos[idx] = (V3<R>{x, startY + inc * R(i), R(i) * incline} + offset) * scale;
                prev[idx] = pos[idx];
                vel[idx] = V3<R>{0, 0, 0};
                invMass[idx] = 1;
                uvx[idx] = x; uvy[idx] = inc * R(i);
                add_link(idx - 1, idx, stretch);
                if (idx > base + 1 && !dihedral) add_link(idx - 2, idx, bend);
                if (idx >= divY) {
                    add_link(idx - divY, idx, stretch);
                    if (i < divY - 1) add_link(idx - divY + 1, idx, shear);
                    if (j > 1 && !dihedral) add_link(idx - 2 * divY, idx, bend);
                }
                if (idx > divY) add_link(idx - divY - 1, idx, shear);
                if (idx > divY) {
                    if ((i & 1) == (j & 1)) {
                        add_triangle(idx - divY - 1, idx, idx - divY);
                        add_triangle(idx - divY - 1, idx - 1, idx);
                    } else {
                        add_triangle(idx - divY, idx - 1, idx);
                        add_triangle(idx - 1, idx - divY, idx - divY - 1);
                    }
                }
            }
            if (attached && (j == 0 || j == divX - 1)) invMass[base + divY - 1] = 0;
            x += inc;
        }
        construct_edges();
        for (const Edge& e : edges) add_bend(e);
    }

    // Physics/ClothCollision.cpp:199-220
    void wall_collisions(const Collidable<R>& w, int ci)
    {
        const R r = R(prm.thickness) + R(prm.tolerance);
        for (size_t i = 0; i < pos.size(); i++) {
            const V3<R>& p = pos[i];
            if (p.y - r <= w.bmin.y) add_contact(i, V3<R>{p.x, w.bmin.y, p.z}, V3<R>{0, 1, 0}, ci);
            if (p.x - r <= w.bmin.x) add_contact(i, V3<R>{w.bmin.x, p.y, p.z}, V3<R>{1, 0, 0}, ci);
            if (p.x + r >= w.bmax.x) add_contact(i, V3<R>{w.bmax.x, p.y, p.z}, V3<R>{-1, 0, 0}, ci);
            if (p.z - r <= w.bmin.z) add_contact(i, V3<R>{p.x, p.y, w.bmin.z}, V3<R>{0, 0, 1}, ci);
            if (p.z + r >= w.bmax.z) add_contact(i, V3<R>{p.x, p.y, w.bmax.z}, V3<R>{0, 0, -1}, ci);
        }
    }

    // Physics/ClothModel.inl:33-47
    // a vertex contact found by a parallel detection loop, added in order
    struct Hit {
        size_t i;
        V3<R> p, n;
        int feature;
    };

    void add_contact(size_t idx, V3<R> p, V3<R> n, int ci)
    {
        Contact<R> c;
        c.idx = (uint32_t)idx;
        c.coll = ci;
        c.point = p; c.normal = n; c.vel = V3<R>{0, 0, 0};
        R len = dot(n, pos[idx] - p);
        c.depth = R(prm.thickness) - len;
        contacts.push_back(c);
    }

    // Physics/ClothCollisionSDF.h:14-24 (SphereSDF::QueryPoint)
    static R sphere_query(const Collidable<R>& s, V3<R> p, V3<R>& cp, V3<R>& n)
    {
        V3<R> d = p - s.center;
        R len = length(d);
        scale(d, R(1) / len);
        n = d;
        cp = s.center + s.radius * d;
        return len - s.radius;
    }

    // Physics/ClothCollisionSDF.h:57-122 (CapsuleSDF): QueryPoint,
    // QuerySegment and QueryTriangle against the axis [P, Q] grown by r
    static R capsule_query(const Collidable<R>& k, V3<R> x, V3<R>& cp, V3<R>& n)
    {
        R t;
        const V3<R> s = closest_pt_segm(x, k.axisP, k.axisQ, t);
        V3<R> d = x - s;
        R len = length(d);
        scale(d, R(1) / len);
        n = d;
        cp = s + k.radius * d;
        return len - k.radius;
    }

    static R capsule_query_segment(const Collidable<R>& k, V3<R> p, V3<R> q, V3<R>& cp, V3<R>& n, R& w1, R& w2)
    {
        R s, t;
        V3<R> c1, c2;
        closest_pt_segm_segm(p, q, k.axisP, k.axisQ, s, t, c1, c2);
        V3<R> d = c1 - c2;
        R len = length(d);
        scale(d, R(1) / len);
        n = d;
        cp = c2 + k.radius * d;
        w1 = R(1.0f) - s;
        w2 = s;
        return len - k.radius;
    }

    static R capsule_query_triangle(const Collidable<R>& k, V3<R> a, V3<R> b, V3<R> c, V3<R>& cp, V3<R>& n,
                                    V3<R>& bar)
    {
        R t;
        V3<R> ce, ct;
        const R len = closest_pt_segm_triangle(k.axisP, k.axisQ, a, b, c, t, bar, ce, ct);
        V3<R> d = ct - ce;
        scale(d, R(1) / len);
        n = d;
        cp = ce + k.radius * d;
        return len - k.radius;
    }

    // Physics/ClothCollision.cpp:430-470 (vertex branch of SDFCollisions; the
    // defaults handleInside = handleCurrent = false)
    template <class Q>
    void sdf_vertex_collisions(Q query, int ci)
    {
        if (!(prm.collFlags & CF_VERTICES)) return;
        auto parts = ordered_blocks<std::vector<Hit>>(pos.size(), [&](size_t b, size_t e, std::vector<Hit>& out) {
            for (size_t i = b; i < e; i++) {
                V3<R> p, n;
                R dist = query(prev[i], p, n);
                R tv = length(prev[i] - pos[i]);
                R radTol = R(prm.thickness) + smax(R(prm.tolerance), tv * R(1.1f));
                if (dist < radTol && dist >= R(0)) out.push_back({i, p, n, -1});
            }
        });
        for (auto& part : parts)
            for (const Hit& h : part) add_contact(h.i, h.p, h.n, ci);
    }

    // Physics/ClothCollisionSDF.h:26-41 (SphereSDF::QuerySegment)
    static R sphere_query_segment(const Collidable<R>& s, V3<R> p, V3<R> q, V3<R>& cp, V3<R>& n, R& w1, R& w2)
    {
        R t;
        V3<R> d = closest_pt_segm(s.center, p, q, t) - s.center;
        R len = length(d);
        scale(d, R(1) / len);
        n = d;
        cp = s.center + s.radius * d;
        w1 = R(1.0f) - t;
        w2 = t;
        return len - s.radius;
    }

    // Physics/ClothCollisionSDF.h:43-55 (SphereSDF::QueryTriangle)
    static R sphere_query_triangle(const Collidable<R>& s, V3<R> a, V3<R> b, V3<R> c, V3<R>& cp, V3<R>& n,
                                   V3<R>& bar)
    {
        V3<R> pt;
        closest_pt_point_triangle(s.center, a, b, c, pt, bar);
        V3<R> d = pt - s.center;
        R len = length(d);
        scale(d, R(1) / len);
        n = d;
        cp = s.center + s.radius * d;
        return len - s.radius;
    }

    // Physics/ClothCollision.cpp:472-579: the edge and triangle branches of
    // SDFCollisions with its defaults (handleInside = handleCurrent = false,
    // handleBorder = false, widened when vertex / edge contacts are off).
    // Queries use the previous positions; the radius uses 1.1 x the largest
    // travel of the primitive's vertices. `seg` returns false for a query
    // that never yields a contact.
    template <class Seg, class Tri>
    void sdf_feature_collisions(Seg seg, Tri tri, int ci)
    {
        bool handleBorder = (prm.collFlags & CF_VERTICES) == 0;
        auto tv = [&](uint32_t i) { return length(prev[i] - pos[i]); };
        if (prm.collFlags & CF_EDGES) {
            const R eps = R(0.02f);
            auto parts = ordered_blocks<std::vector<EdgeContact<R>>>(
                edges.size(), [&](size_t b, size_t end, std::vector<EdgeContact<R>>& out) {
                    for (size_t e = b; e < end; e++) {
                        const uint32_t i1 = edges[e].i1, i2 = edges[e].i2;
                        V3<R> cp, n;
                        R w1 = 0, w2 = 0;
                        R dist;
                        if (!seg(prev[i1], prev[i2], cp, n, w1, w2, dist)) continue;
                        const R t = smax(tv(i1), tv(i2));
                        const R radTol = R(prm.thickness) + smax(R(prm.tolerance), t * R(1.1f));
                        if (dist < radTol && dist >= R(0) &&
                            ((w1 > eps && w1 < R(1.f) - eps && w2 > eps && w2 < R(1.f) - eps) || handleBorder)) {
                            EdgeContact<R> ec;
                            ec.i1 = i1; ec.i2 = i2;
                            ec.clothEdge = (int)e; ec.feature = -1; ec.mesh = -1;
                            ec.point = cp; ec.normal = n;
                            ec.w1 = w1; ec.w2 = w2;
                            out.push_back(ec);
                        }
                    }
                });
            append_parts(edgeContacts, parts);
        }
        handleBorder = handleBorder || (prm.collFlags & CF_EDGES) == 0;
        if (prm.collFlags & CF_TRIANGLES) {
            const R eps = R(0.001f);
            auto parts = ordered_blocks<std::vector<TriContact<R>>>(
                tris.size(), [&](size_t b, size_t end, std::vector<TriContact<R>>& out) {
                    for (size_t j = b; j < end; j++) {
                        const uint32_t i1 = tris[j].i1, i2 = tris[j].i2, i3 = tris[j].i3;
                        V3<R> cp, n, bar;
                        const R dist = tri(prev[i1], prev[i2], prev[i3], cp, n, bar);
                        const R t = smax(tv(i1), smax(tv(i2), tv(i3)));
                        const R radTol = R(prm.thickness) + smax(R(prm.tolerance), t * R(1.1f));
                        if (dist < radTol && dist >= R(0) &&
                            ((bar.x > eps && bar.y > eps && bar.z > eps) || handleBorder)) {
                            TriContact<R> tc;
                            tc.i1 = i1; tc.i2 = i2; tc.i3 = i3;
                            tc.clothTri = (int)j; tc.feature = -1; tc.mesh = -1;
                            tc.point = cp; tc.normal = n;
                            tc.w1 = bar.x; tc.w2 = bar.y; tc.w3 = bar.z;
                            out.push_back(tc);
                        }
                    }
                });
            append_parts(triContacts, parts);
        }
    }

    // Physics/ClothCollision.cpp:17-87 (external collidables; the cloth tree it
    // also rebuilds is consumed by nothing on this path)
    void detect_collisions()
    {
        contacts.clear();
        edgeContacts.clear();
        triContacts.clear();
        for (size_t ci = 0; ci < world.size(); ci++) {
            const Collidable<R>& c = world[ci];
            if (c.type == CT_WALLS && (prm.collFlags & CF_WALLS)) {
                wall_collisions(c, (int)ci);
            } else if (c.type == CT_SPHERE) {
                sdf_vertex_collisions([&](V3<R> p, V3<R>& cp, V3<R>& n) { return sphere_query(c, p, cp, n); }, (int)ci);
                sdf_feature_collisions(
                    [&](V3<R> p, V3<R> q, V3<R>& cp, V3<R>& n, R& w1, R& w2, R& d) {
                        d = sphere_query_segment(c, p, q, cp, n, w1, w2);
                        return true;
                    },
                    [&](V3<R> a, V3<R> b, V3<R> cc, V3<R>& cp, V3<R>& n, V3<R>& bar) {
                        return sphere_query_triangle(c, a, b, cc, cp, n, bar);
                    },
                    (int)ci);
            } else if (c.type == CT_CAPSULE) {
                sdf_vertex_collisions([&](V3<R> p, V3<R>& cp, V3<R>& n) { return capsule_query(c, p, cp, n); },
                                      (int)ci);
                sdf_feature_collisions(
                    [&](V3<R> p, V3<R> q, V3<R>& cp, V3<R>& n, R& w1, R& w2, R& d) {
                        d = capsule_query_segment(c, p, q, cp, n, w1, w2);
                        return true;
                    },
                    [&](V3<R> a, V3<R> b, V3<R> cc, V3<R>& cp, V3<R>& n, V3<R>& bar) {
                        return capsule_query_triangle(c, a, b, cc, cp, n, bar);
                    },
                    (int)ci);
            } else if (c.type == CT_SDF) {
                const GridSDF<R>& g = sdfs[c.sdf];
                sdf_vertex_collisions([&](V3<R> p, V3<R>& cp, V3<R>& n) { return g.query_point(p, cp, n); }, (int)ci);
                // GridSDF::QuerySegment is a stub returning 0 with coordinates
                // (0, 0): with vertex contacts on, the border test rejects it
                // (ClothCollision.cpp:474,497-498); without, the step refuses.
                sdf_feature_collisions(
                    [&](V3<R>, V3<R>, V3<R>&, V3<R>&, R&, R&, R&) { return false; },
                    [&](V3<R> a, V3<R> b, V3<R> cc, V3<R>& cp, V3<R>& n, V3<R>& bar) {
                        return g.query_triangle(a, b, cc, cp, n, bar);
                    },
                    (int)ci);
            } else if (c.type == CT_MESH) {
                mesh_collisions(meshes[c.mesh], (int)ci, c.mesh);
            }
        }
        if (prm.collFlags & CF_SELF) self_collisions();
        auto byRank = [](const std::vector<int>& rank) {
            return [&rank](int a, int b) { return rank[a] < rank[b]; };
        };
        if (!edgeContactRank.empty())
            std::stable_sort(edgeContacts.begin(), edgeContacts.end(),
                             [&](const EdgeContact<R>& a, const EdgeContact<R>& b) {
                                 return byRank(edgeContactRank)(a.clothEdge, b.clothEdge);
                             });
        if (!triContactRank.empty())
            std::stable_sort(triContacts.begin(), triContacts.end(), [&](const TriContact<R>& a, const TriContact<R>& b) {
                return byRank(triContactRank)(a.clothTri, b.clothTri);
            });
    }

    static V3<R> vR(const V3<float>& a) { return V3<R>{R(a.x), R(a.y), R(a.z)}; }

    // Physics/ClothCollision.cpp:89-197 with MeshClosestPoints::ClosestPoints
    // (Geometry/ClosestPointsBetweenMeshes.cpp:595-695): nearest obstacle
    // triangle per cloth vertex, nearest obstacle vertex per cloth triangle,
    // nearest obstacle edge per cloth edge, all from the cloth's previous
    // positions. The reference finds the nearest feature by a branch-and-bound
    // tree walk that keeps the first strictly nearer candidate in its walk
    // order; here the search is exhaustive and ties at equal FP32 distance go
    // to the lowest feature index. Distance 0 is skipped as in the reference.
    // ComputeTriangleNormal (ClosestPointOnMesh.cpp:37-62) as the contact
    // solve calls it (ClothCollisionHandler.cpp:159: no pseudo-normal, mesh
    // normals): the vertex normals interpolated with the barycentric
    // coordinates, normalised
    V3<R> mesh_triangle_normal(const ObstacleMesh& M, int tri, V3<R> co) const
    {
        const uint32_t i0 = M.idx[3 * tri], i1 = M.idx[3 * tri + 1], i2 = M.idx[3 * tri + 2];
        V3<R> nn = co.x * vR(M.vn[i0]) + co.y * vR(M.vn[i1]) + co.z * vR(M.vn[i2]);
        normalize(nn);
        return nn;
    }

    // ComputeEdgeNormal (ClosestPointOnMesh.cpp:483-497): the edge's
    // pseudo-normal, or an end's vertex normal at parameter exactly 0 or 1
    V3<R> mesh_edge_normal(const ObstacleMesh& M, int e, R pm) const
    {
        V3<float> nf = M.en[e];
        if (pm == R(0)) nf = M.vn[M.edges[e].i1];
        if (pm == R(1)) nf = M.vn[M.edges[e].i2];
        return vR(nf);
    }

    void mesh_collisions(const ObstacleMesh& M, int ci, int mi)
    {
        const size_t n = pos.size();
        std::vector<R> tvs(n);
        R maxTv = 0;
        for (size_t i = 0; i < n; i++) {
            tvs[i] = length(pos[i] - prev[i]);
            if (tvs[i] > maxTv) maxTv = tvs[i];
        }
        const R radTol = R(prm.thickness) + smax(R(prm.tolerance), maxTv * R(1.5f));
        const size_t nt = M.idx.size() / 3;
        // the features whose boxes meet the box of radius r around [a, b]
        // (every feature nearer than r does), or all of them
        auto candidates = [&](std::vector<int>& cand, const BoxGrid& g, size_t count, V3<R> a, V3<R> b, R r) {
            cand.clear();
            if (!prm.accel) {
                for (size_t k = 0; k < count; k++) cand.push_back((int)k);
                return;
            }
            const double pad = double(r) * (1.0 + 1e-6) + 1e-6 * (std::fabs(double(a.x)) + std::fabs(double(a.y)) +
                                                                  std::fabs(double(a.z)) + 1.0);
            const V3<float> lo{float(std::min(double(a.x), double(b.x)) - pad),
                               float(std::min(double(a.y), double(b.y)) - pad),
                               float(std::min(double(a.z), double(b.z)) - pad)};
            const V3<float> hi{float(std::max(double(a.x), double(b.x)) + pad),
                               float(std::max(double(a.y), double(b.y)) + pad),
                               float(std::max(double(a.z), double(b.z)) + pad)};
            g.query(lo, hi, cand);
        };
        if (prm.collFlags & CF_VERTICES) {
            auto parts = ordered_blocks<std::vector<Hit>>(n, [&](size_t b, size_t end, std::vector<Hit>& out) {
                std::vector<int> cand;
                for (size_t i = b; i < end; i++) {
                    const V3<R> q = prev[i];
                    R best = R(FLT_MAX);
                    int tri = -1;
                    V3<R> bcp{0, 0, 0}, bn{0, 0, 0};
                    candidates(cand, M.triGrid, nt, q, q, radTol);
                    for (const int ct : cand) {
                        const size_t t = (size_t)ct;
                        V3<R> pt, bar;
                        closest_pt_point_triangle(q, vR(M.v[M.idx[3 * t]]), vR(M.v[M.idx[3 * t + 1]]),
                                                  vR(M.v[M.idx[3 * t + 2]]), pt, bar);
                        V3<R> delta = q - pt;
                        R d = length(delta);
                        if (d == R(0)) continue;
                        if (d < best) {
                            best = d;
                            tri = (int)t;
                            bcp = pt;
                            bn = (R(1.f) / d) * delta;
                        }
                    }
                    if (tri >= 0 && best < radTol && best >= R(0)) out.push_back({i, bcp, bn, tri});
                }
            });
            for (auto& part : parts)
                for (const Hit& h : part) {
                    add_contact(h.i, h.p, h.n, ci);
                    contacts.back().feature = h.feature;
                    contacts.back().mesh = mi;
                }
        }
        if (prm.collFlags & CF_TRIANGLES) {
            auto parts = ordered_blocks<std::vector<TriContact<R>>>(
                tris.size(), [&](size_t b, size_t end, std::vector<TriContact<R>>& out) {
            std::vector<int> cand;
            for (size_t j = b; j < end; j++) {
                const uint32_t i1 = tris[j].i1, i2 = tris[j].i2, i3 = tris[j].i3;
                const V3<R> a = prev[i1], b = prev[i2], c = prev[i3];
                R best = R(FLT_MAX), dist = R(-FLT_MAX);
                int vtx = -1;
                V3<R> bpt{0, 0, 0}, bn{0, 0, 0}, bbar{0, 0, 0};
                {
                    const R tvm = smax(tvs[i1], smax(tvs[i2], tvs[i3]));
                    const R rtq = R(prm.thickness) + smax(R(prm.tolerance), tvm * R(1.2f));
                    const V3<R> lo3{std::min({a.x, b.x, c.x}), std::min({a.y, b.y, c.y}), std::min({a.z, b.z, c.z})};
                    const V3<R> hi3{std::max({a.x, b.x, c.x}), std::max({a.y, b.y, c.y}), std::max({a.z, b.z, c.z})};
                    candidates(cand, M.vtxGrid, M.v.size(), lo3, hi3, rtq);
                }
                for (const int cv : cand) {
                    const size_t v = (size_t)cv;
                    V3<R> cp, bar;
                    const V3<R> mv = vR(M.v[v]);
                    closest_pt_point_triangle(mv, a, b, c, cp, bar);
                    V3<R> delta = cp - mv;
                    R d = length(delta);
                    if (d == R(0)) continue;
                    if (d < best) {
                        best = d;
                        dist = d;
                        vtx = (int)v;
                        bpt = mv;
                        bbar = bar;
                        bn = (R(1.f) / d) * delta;
                    }
                }
                const R tv = smax(tvs[i1], smax(tvs[i2], tvs[i3]));
                const R rt = R(prm.thickness) + smax(R(prm.tolerance), tv * R(1.2f));
                if (dist < rt && dist >= R(0)) {
                    TriContact<R> tc;
                    tc.i1 = i1; tc.i2 = i2; tc.i3 = i3;
                    tc.clothTri = (int)j; tc.feature = vtx; tc.mesh = mi;
                    tc.point = bpt; tc.normal = bn;
                    tc.w1 = bbar.x; tc.w2 = bbar.y; tc.w3 = bbar.z;
                    out.push_back(tc);
                }
            }
                });
            append_parts(triContacts, parts);
        }
        if (prm.collFlags & CF_EDGES) {
            auto parts = ordered_blocks<std::vector<EdgeContact<R>>>(
                edges.size(), [&](size_t b, size_t end, std::vector<EdgeContact<R>>& out) {
            std::vector<int> cand;
            for (size_t e1 = b; e1 < end; e1++) {
                const uint32_t i1 = edges[e1].i1, i2 = edges[e1].i2;
                const V3<R> p = prev[i1], q = prev[i2];
                R best = R(FLT_MAX), dist = R(-FLT_MAX);
                int edge = -1;
                V3<R> bcp{0, 0, 0}, bn{0, 0, 0};
                R bws = 0;
                candidates(cand, M.edgeGrid, M.edges.size(), p, q, radTol);
                for (const int ce : cand) {
                    const size_t e = (size_t)ce;
                    R ps, pm;
                    V3<R> cs, cm;
                    closest_pt_segm_segm(p, q, vR(M.v[M.edges[e].i1]), vR(M.v[M.edges[e].i2]), ps, pm, cs, cm);
                    V3<R> delta = cm - cs;
                    R d = length(delta);
                    if (d == R(0)) continue;
                    if (d < best) {
                        best = d;
                        dist = d;
                        edge = (int)e;
                        bcp = cm;
                        bws = ps;
                        bn = mesh_edge_normal(M, (int)e, pm);
                    }
                }
                if (dist < radTol && dist >= R(0)) {
                    EdgeContact<R> ec;
                    ec.i1 = i1; ec.i2 = i2;
                    ec.clothEdge = (int)e1; ec.feature = edge; ec.mesh = mi;
                    ec.point = bcp; ec.normal = bn;
                    ec.w1 = R(1.0f) - bws; ec.w2 = bws;
                    out.push_back(ec);
                }
            }
                });
            append_parts(edgeContacts, parts);
        }
    }

    // Physics/ClothCollisionHandler.cpp:289-362
    R solve_edge_contacts()
    {
        const R mu = prm.friction, len0 = prm.thickness;
        R error = 0;
        for (EdgeContact<R>& c : edgeContacts) {
            V3<R>& x1 = pos[c.i1];
            V3<R>& x2 = pos[c.i2];
            const R m1 = invMass[c.i1], m2 = invMass[c.i2];
            V3<R> p = c.w1 * x1 + c.w2 * x2;
            if (c.mesh >= 0) {
                const ObstacleMesh& M = meshes[c.mesh];
                R s1, t;
                closest_pt_segm_segm(vR(M.v[M.edges[c.feature].i1]), vR(M.v[M.edges[c.feature].i2]), x1, x2, s1, t,
                                     c.point, p);
                c.w1 = R(1.0f) - t;
                c.w2 = t;
            }
            R len = dot(c.normal, p - c.point);
            if (len > len0) continue;
            R s = R(1.f) / (c.w1 * c.w1 / m1 + c.w2 * c.w2 / m2);
            R depth = len0 - len;
            if (depth > error) error = depth;
            R dLambda = s * depth;
            c.lambda += dLambda;
            V3<R> disp = dLambda * c.normal;
            if (mu != R(0)) {
                V3<R> p0 = c.w1 * prev[c.i1] + c.w2 * prev[c.i2];
                V3<R> v12 = p - p0;
                R vnrel = dot(c.normal, v12);
                V3<R> vt = v12 - vnrel * c.normal;
                R vtrel = length(vt);
                const R limit = mu * c.lambda;
                if (vtrel > R(0.001f)) {
                    R dLF = vtrel;
                    R l0 = c.lambdaF;
                    c.lambdaF = l0 + dLF;
                    if (c.lambdaF >= limit) c.lambdaF = limit;
                    dLambda = c.lambdaF - l0;
                    scale(vt, R(1.f) / vtrel);
                    V3<R> pf = (-s * dLambda) * vt;
                    x1 += pf * (c.w1 * m1);
                    x2 += pf * (c.w2 * m2);
                }
            }
            x1 += disp * (c.w1 * m1);
            x2 += disp * (c.w2 * m2);
        }
        return error;
    }

    // Physics/ClothCollisionHandler.cpp:204-287
    R solve_tri_contacts()
    {
        const R mu = prm.friction, len0 = prm.thickness;
        R error = 0;
        for (TriContact<R>& c : triContacts) {
            V3<R>& x1 = pos[c.i1];
            V3<R>& x2 = pos[c.i2];
            V3<R>& x3 = pos[c.i3];
            const R m1 = invMass[c.i1], m2 = invMass[c.i2], m3 = invMass[c.i3];
            V3<R> p, coords;
            closest_pt_point_triangle(c.point, x1, x2, x3, p, coords);
            c.w1 = coords.x; c.w2 = coords.y; c.w3 = coords.z;
            V3<R> n = -cross(x2 - x1, x3 - x1);
            normalize(n);
            c.normal = n;
            if (c.mesh >= 0 && c.feature >= 0) {
                const V3<R> mn = vR(meshes[c.mesh].vn[c.feature]);
                if (dot(n, mn) < R(0)) c.normal = -n;
            }
            R len = dot(c.normal, p - c.point);
            if (len > len0) continue;
            R s = R(1.f) / (c.w1 * c.w1 / m1 + c.w2 * c.w2 / m2 + c.w3 * c.w3 / m3);
            R depth = len0 - len;
            if (depth > error) error = depth;
            R dLambda = s * depth;
            c.lambda += dLambda;
            V3<R> disp = dLambda * c.normal;
            if (mu != R(0)) {
                V3<R> p0 = c.w1 * prev[c.i1] + c.w2 * prev[c.i2] + c.w3 * prev[c.i3];
                V3<R> v12 = p - p0;
                R vnrel = dot(c.normal, v12);
                V3<R> vt = v12 - vnrel * c.normal;
                R vtrel = length(vt);
                const R limit = mu * c.lambda;
                if (vtrel > R(0.001f)) {
                    R dLF = vtrel;
                    R l0 = c.lambdaF;
                    c.lambdaF = l0 + dLF;
                    if (c.lambdaF >= limit) c.lambdaF = limit;
                    dLambda = c.lambdaF - l0;
                    scale(vt, R(1.f) / vtrel);
                    V3<R> pf = (-s * dLambda) * vt;
                    x1 += pf * (c.w1 * m1);
                    x2 += pf * (c.w2 * m2);
                    x3 += pf * (c.w3 * m3);
                }
            }
            x1 += disp * (c.w1 * m1);
            x2 += disp * (c.w2 * m2);
            x3 += disp * (c.w3 * m3);
        }
        return error;
    }

    // Physics/ClothPBD.cpp:52-108
    void project_triangle(Tri<R>& t, R h2, R epsilonUV)
    {
        V3<R>& x1 = pos[t.i1];
        V3<R>& x2 = pos[t.i2];
        V3<R>& x3 = pos[t.i3];
        const R m1 = invMass[t.i1], m2 = invMass[t.i2], m3 = invMass[t.i3];
        V3<R> dx1 = x2 - x1;
        V3<R> dx2 = x3 - x1;
        V3<R> wu = t.invDet * (t.dv2 * dx1 - t.dv1 * dx2);
        V3<R> wv = t.invDet * (-t.du2 * dx1 + t.du1 * dx2);
        R err, diag, lambda;
        V3<R> q1, q2, q3;
        // warp
        q2 = (t.invDet * t.dv2) * wu;
        q3 = (-t.invDet * t.dv1) * wu;
        q1 = -q2 - q3;
        diag = dot(m1 * q1, q1) + dot(m2 * q2, q2) + dot(m3 * q3, q3);
        err = R(0.5f) * (dot(wu, wu) - t.lu2);
        lambda = -err / diag;
        x1 += (m1 * lambda) * q1;
        x2 += (m2 * lambda) * q2;
        x3 += (m3 * lambda) * q3;
        // weft
        q2 = (-t.invDet * t.du2) * wv;
        q3 = (t.invDet * t.du1) * wv;
        q1 = -q2 - q3;
        diag = dot(m1 * q1, q1) + dot(m2 * q2, q2) + dot(m3 * q3, q3);
        err = R(0.5f) * (dot(wv, wv) - t.lv2);
        lambda = -err / diag;
        x1 += (m1 * lambda) * q1;
        x2 += (m2 * lambda) * q2;
        x3 += (m3 * lambda) * q3;
        // shear
        q2 = t.dv2 * wv - t.du2 * wu;
        q3 = -t.dv1 * wv + t.du1 * wu;
        q1 = -q2 - q3;
        diag = t.invDet * (dot(m1 * q1, q1) + dot(m2 * q2, q2) + dot(m3 * q3, q3));
        err = dot(wu, wv);
        R s = R(1.f) / (R(1.f) + t.lu0 * t.lv0 * epsilonUV / (diag * h2 * t.area));
        lambda = -s * err / diag;
        x1 += (m1 * lambda) * q1;
        x2 += (m2 * lambda) * q2;
        x3 += (m3 * lambda) * q3;
    }

    void solve_triangles(R h)
    {
        const R h2 = h * h;
        const R epsilonUV = R(2) * (R(1) + R(prm.poisson)) / (smax(R(prm.shear), R(0.1f)) * R(prm.unit) * R(prm.unit));
        if (triOrder.empty()) {
            for (auto& t : tris) project_triangle(t, h2, epsilonUV);
        } else {
            for (int k : triOrder) project_triangle(tris[k], h2, epsilonUV);
        }
    }

    // Physics/ClothPBD.cpp:179-224 with EvaluateAngle 110-139, EvaluateGradient 141-177
    R project_bend(const Bend& bc, R sStiff, R& error)
    {
        const R m1 = invMass[bc.i1], m2 = invMass[bc.i2], m3 = invMass[bc.i3], m4 = invMass[bc.i4];
        if (m1 + m2 + m3 + m4 == R(0)) return error;
        const V3<R> x1 = pos[bc.i1], x2 = pos[bc.i2], x3 = pos[bc.i3], x4 = pos[bc.i4];
        // EvaluateAngle
        V3<R> x12 = x2 - x1;
        R l = length(x12);
        V3<R> x13 = x3 - x1;
        V3<R> n1 = cross(x12, x13);
        R l1 = length(n1);
        V3<R> x14 = x4 - x1;
        V3<R> n2 = cross(x14, x12);
        R l2 = length(n2);
        R cosine, theta;
        if (l1 == R(0) || l2 == R(0)) {
            cosine = 1;
            theta = 0;
        } else {
            R il1 = R(1.f) / l1, il2 = R(1.f) / l2;
            V3<R> n1n = il1 * n1, n2n = il2 * n2;
            cosine = dot(n1n, n2n);
            cosine = smin(R(1.f), smax(R(-1.f), cosine));
            theta = acos_as(cosine);
        }
        if (std::fabs(cosine) >= R(1 - 0.001f)) return error;
        // EvaluateGradient (recomputes the same intermediate quantities)
        R h1 = l1 / l, h2 = l2 / l;
        R he = R(0.5f) * (h1 + h2);
        R il1 = R(1.f) / l1, il2 = R(1.f) / l2;
        V3<R> n1n = il1 * n1, n2n = il2 * n2;
        V3<R> n1p = n1n - n2n * cosine;
        V3<R> n2p = n2n - n1n * cosine;
        R sine = rsqrt_len(R(1) - cosine * cosine);
        V3<R> q3 = (il1 / sine / he) * cross(x12, n2p);
        V3<R> q4 = (-il2 / sine / he) * cross(x12, n1p);
        V3<R> q2 = (-il1 / sine / he) * cross(x13, n2p) + (il2 / sine / he) * cross(x14, n1p);
        V3<R> q1 = -q2 - q3 - q4;
        R diag = m1 * dot(q1, q1) + m2 * dot(q2, q2) + m3 * dot(q3, q3) + m4 * dot(q4, q4);
        if (diag == R(0)) return error;
        R e = sStiff * theta;
        if (std::fabs(e) > error) error = std::fabs(e);
        R lambda = -sStiff * theta / he / diag;
        pos[bc.i1] += (lambda * m1) * q1;
        pos[bc.i2] += (lambda * m2) * q2;
        pos[bc.i3] += (lambda * m3) * q3;
        pos[bc.i4] += (lambda * m4) * q4;
        return error;
    }

    R solve_bends()
    {
        R error = 0;
        const R s = R(0.0005f) * R(prm.bend);
        if (bendOrder.empty()) {
            for (const Bend& b : bends) project_bend(b, s, error);
        } else {
            for (int k : bendOrder) project_bend(bends[k], s, error);
        }
        return error;
    }

    // Physics/ClothCollisionHandler.cpp:134-202 (external vertex contacts;
    // no mesh re-projection for walls and SDFs)
    R solve_contacts(R h)
    {
        const R invH = R(1.f) / h;
        const R mu = prm.friction;
        const R thickness = prm.thickness;
        R error = 0;
        for (Contact<R>& c : contacts) {
            V3<R>& p = pos[c.idx];
            if (c.mesh >= 0) {
                // UPDATE_CLOSEST_POINT (ClothCollisionHandler.cpp:145-160):
                // re-project onto the same obstacle triangle; interpolated
                // vertex normal (ComputeTriangleNormal, ClosestPointOnMesh.cpp:37-62)
                const ObstacleMesh& M = meshes[c.mesh];
                const uint32_t i0 = M.idx[3 * c.feature], i1 = M.idx[3 * c.feature + 1], i2 = M.idx[3 * c.feature + 2];
                V3<R> cp, co;
                closest_pt_point_triangle(p, vR(M.v[i0]), vR(M.v[i1]), vR(M.v[i2]), cp, co);
                c.point = cp;
                c.normal = mesh_triangle_normal(M, (int)c.feature, co);
            }
            V3<R> delta = p - c.point;
            R len = dot(delta, c.normal);
            V3<R> disp = c.normal;
            if (len > thickness) continue;
            R depth = thickness - len;
            if (std::fabs(depth) > error) error = std::fabs(depth);
            R dLambda = depth;
            c.lambda += dLambda;
            scale(disp, dLambda);
            V3<R> v12 = p - prev[c.idx] - c.vel;
            R vnrel = dot(c.normal, v12);
            V3<R> vt = v12 - vnrel * c.normal;
            R vtrel = length(vt);
            const R limit = mu * c.lambda;
            if (vtrel > R(0.001f)) {
                R dLambdaF = vtrel;
                R lambdaF0 = c.lambdaF;
                c.lambdaF = lambdaF0 + dLambdaF;
                if (c.lambdaF >= limit) c.lambdaF = limit;
                dLambdaF = c.lambdaF - lambdaF0;
                scale(vt, R(1.0f) / vtrel);
                disp += -dLambdaF * vt;
            }
            p += disp;
            vel[c.idx] += invH * disp;
        }
        return error;
    }

    // ------------------------------------------------------ self collision
    // Swept boxes of ClothModel::ComputeTree (ClothCollision.cpp:240-315):
    // prev and pos of the primitive's vertices, extruded by
    // thickness + 0.5 tolerance; closed-interval overlap (Aabb3.h:72-87).
    struct Box {
        V3<float> lo, hi;
    };
    Box swept_box(std::initializer_list<uint32_t> ids, float tol) const
    {
        Box b{V3<float>{1e7f, 1e7f, 1e7f}, V3<float>{-1e7f, -1e7f, -1e7f}};
        auto add = [&](V3<R> p) {
            const V3<float> q{float(p.x), float(p.y), float(p.z)};
            b.lo = V3<float>{std::min(b.lo.x, q.x), std::min(b.lo.y, q.y), std::min(b.lo.z, q.z)};
            b.hi = V3<float>{std::max(b.hi.x, q.x), std::max(b.hi.y, q.y), std::max(b.hi.z, q.z)};
        };
        for (uint32_t i : ids) add(prev[i]);
        for (uint32_t i : ids) add(pos[i]);
        b.lo = V3<float>{b.lo.x - tol, b.lo.y - tol, b.lo.z - tol};
        b.hi = V3<float>{b.hi.x + tol, b.hi.y + tol, b.hi.z + tol};
        return b;
    }
    static bool overlap(const Box& a, const Box& b)
    {
        return !(a.hi.x < b.lo.x || a.hi.y < b.lo.y || a.hi.z < b.lo.z || b.hi.x < a.lo.x || b.hi.y < a.lo.y ||
                 b.hi.z < a.lo.z);
    }

    // SelfCollisionDetector.cpp:263-276, in FP64
    static double coplanarity(double t, const V3<double> x[4], const V3<double> v[4])
    {
        const V3<double> z1 = x[0] + t * v[0], z2 = x[1] + t * v[1], z3 = x[2] + t * v[2], z4 = x[3] + t * v[3];
        return dot(cross(z2 - z1, z3 - z1), z4 - z1);
    }
    static double triple(V3<double> a, V3<double> b, V3<double> c) { return dot(c, cross(a, b)); }

    // SelfCollisionDetector.cpp:414-488: bisection for the sign change of the
    // coplanarity volume on [ta, tb]
    bool check_interval(double ta, double tb, const V3<double> x[4], const V3<double> v[4], double& tx) const
    {
        const double eps = double(prm.ccdBisect);
        double tx0 = ta, tx1 = tb;
        double fx0 = coplanarity(tx0, x, v), fx1 = coplanarity(tx1, x, v);
        if (fx0 * fx1 > 0) return false;
        tx = tx0;
        for (int l = 0; l < 40; l++) {
            const double tmid = 0.5 * (tx0 + tx1);
            const double fmid = coplanarity(tmid, x, v);
            if (std::fabs(fmid) < eps) {
                tx = tmid;
                break;
            }
            if (fx0 * fmid < 0) {
                tx1 = tmid;
                fx1 = fmid;
                tx = tx0;
            } else if (fmid * fx1 < 0) {
                tx0 = tmid;
                fx0 = fmid;
                tx = tmid;
            } else {
                return false;
            }
        }
        return true;
    }

    // SelfCollisionDetector.cpp:278-412 (SolveCubic, CheckDegenerateCase,
    // FindInflexionPoints), evaluated in FP64: the time in [0, 1] at which the
    // four moving points become coplanar.
    bool solve_cubic(const V3<R> xf[4], const V3<R> vf[4], double& toi) const
    {
        V3<double> x[4], v[4];
        for (int k = 0; k < 4; k++) {
            x[k] = V3<double>{double(xf[k].x), double(xf[k].y), double(xf[k].z)};
            v[k] = V3<double>{double(vf[k].x), double(vf[k].y), double(vf[k].z)};
        }
        const double volEps = double(prm.ccdVolume);
        const double f0 = coplanarity(0.0, x, v);
        if (std::fabs(f0) < volEps) return false;
        const V3<double> v10 = v[0] - v[3], v20 = v[1] - v[3], v30 = v[2] - v[3];
        const V3<double> x10 = x[0] - x[3], x20 = x[1] - x[3], x30 = x[2] - x[3];
        const double a = triple(v10, v20, v30);
        const double b = triple(x10, v20, v30) + triple(v10, x20, v30) + triple(v10, v20, x30);
        const double c = triple(x10, x20, v30) + triple(x10, v20, x30) + triple(v10, x20, x30);
        if (std::fabs(a) < volEps) {
            if (std::fabs(b) < volEps) return false;
            const double d = triple(x10, x20, x30);
            const double discr = c * c - 4 * b * d;
            if (discr < 0) return false;
            if (discr == 0) {
                const double t = -0.5 * c / b;
                if (t < 0 || t > 1) return false;
                toi = t;
                return true;
            }
            const double sq = std::sqrt(discr);
            const double t1 = 0.5 * (-c - sq) / b, t2 = 0.5 * (-c + sq) / b;
            const double t = std::min(t1, t2);
            if (t < 0 || t > 1) return false;
            toi = t;
            return true;
        }
        double t1 = 0, t2 = 0;
        int np;
        const double discr = b * b - 3 * a * c;
        if (discr < 0) {
            np = 0;
        } else if (discr == 0) {
            t1 = t2 = -b / (3 * a);
            np = 1;
        } else {
            const double sq = std::sqrt(discr);
            t1 = (-b - sq) / (3 * a);
            t2 = (-b + sq) / (3 * a);
            if (t2 < t1) std::swap(t1, t2);
            np = 2;
        }
        auto ci = [&](double ta, double tb) { return check_interval(ta, tb, x, v, toi); };
        if (np == 0) return ci(0, 1);
        if (np == 1) {
            if (t1 < 0 || t1 > 1) return ci(0, 1);
            return ci(0, t1) || ci(t1, 1);
        }
        if (t1 > 1 || t2 < 0 || (t1 < 0 && t2 > 1)) return ci(0, 1);
        if (t1 < 0 && t2 < 1) return ci(0, t2) || ci(t2, 1);
        if (t1 < 1 && t2 > 1) return ci(0, t1) || ci(t1, 1);
        if (t1 > 0 && t2 < 1) return ci(0, t1) || ci(t1, t2) || ci(t2, 1);
        return false;
    }

    // SelfCollisionDetector.cpp:137-261: vertex i4 against triangle (i1 i2 i3)
    bool vertex_triangle(uint32_t i1, uint32_t i2, uint32_t i3, uint32_t i4, SelfContact<R>& out) const
    {
        const V3<R> x[4] = {prev[i1], prev[i2], prev[i3], prev[i4]};
        const V3<R> v[4] = {pos[i1] - prev[i1], pos[i2] - prev[i2], pos[i3] - prev[i3], pos[i4] - prev[i4]};
        double toid;
        if (!solve_cubic(x, v, toid)) return false;
        const R toi = R(float(toid));
        const V3<R> z1 = x[0] + toi * v[0], z2 = x[1] + toi * v[1], z3 = x[2] + toi * v[2], z4 = x[3] + toi * v[3];
        V3<R> p, coords;
        closest_pt_point_triangle(z4, z1, z2, z3, p, coords);
        const V3<R> delta = z4 - p;
        const R eps = R(prm.ccdVtDistance);
        if (!(dot(delta, delta) < eps * eps)) return false;
        V3<R> triNormal = cross(x[1] - x[0], x[2] - x[0]);
        normalize(triNormal);
        const R side = dot(triNormal, x[3] - x[0]);
        V3<R> n = cross(z2 - z1, z3 - z1);
        normalize(n);
        if (side < R(0)) n = -n;
        out.i1 = i1; out.i2 = i2; out.i3 = i3; out.i4 = i4;
        out.w1 = coords.x; out.w2 = coords.y; out.w3 = coords.z;
        out.normal = n;
        out.side = signed_volume(x[3], x[0], x[1], x[2]);
        return true;
    }

    // SelfCollisionDetector.cpp:491-606: edge (i1 i2) against edge (i3 i4)
    bool edge_edge(uint32_t i1, uint32_t i2, uint32_t i3, uint32_t i4, SelfContact<R>& out) const
    {
        const V3<R> x[4] = {prev[i1], prev[i2], prev[i3], prev[i4]};
        const V3<R> v[4] = {pos[i1] - prev[i1], pos[i2] - prev[i2], pos[i3] - prev[i3], pos[i4] - prev[i4]};
        double toid;
        if (!solve_cubic(x, v, toid)) return false;
        const R toi = R(float(toid));
        const V3<R> z1 = x[0] + toi * v[0], z2 = x[1] + toi * v[1], z3 = x[2] + toi * v[2], z4 = x[3] + toi * v[3];
        R s, t;
        V3<R> c1, c2;
        closest_pt_segm_segm(z1, z2, z3, z4, s, t, c1, c2);
        const V3<R> d = c1 - c2;
        const R eps = R(0.5f) * R(prm.thickness);
        if (!(dot(d, d) <= eps * eps)) return false;
        out.i1 = i1; out.i2 = i2; out.i3 = i3; out.i4 = i4;
        out.w1 = s; out.w2 = t; out.w3 = 0;
        R s0, t0;
        closest_pt_segm_segm(x[0], x[1], x[2], x[3], s0, t0, c1, c2);
        V3<R> n = c2 - c1;
        normalize(n);
        out.normal = n;
        out.side = signed_volume(x[0], x[1], x[2], x[3]);
        return true;
    }

    // SelfCollisionDetector.cpp:60-135 over the exact candidate set: every
    // (triangle, vertex) pair, vertex not of the triangle and not pinned, and
    // every (edge1 < edge2) pair without a shared vertex, whose swept boxes
    // overlap. The reference's tree duplicates straddling primitives into
    // several leaves and so may repeat a pair; each pair is tested once here.
    void self_collisions()
    {
        selfTris.clear();
        selfEdges.clear();
        const float tol = prm.thickness + prm.tolerance * 0.5f;
        const size_t nt = tris.size(), nv = pos.size(), ne = edges.size();
        std::vector<Box> tb(nt), vb(nv), eb(ne);
        for (size_t t = 0; t < nt; t++) tb[t] = swept_box({tris[t].i1, tris[t].i2, tris[t].i3}, tol);
        for (size_t i = 0; i < nv; i++) vb[i] = swept_box({(uint32_t)i}, tol);
        for (size_t e = 0; e < ne; e++) eb[e] = swept_box({edges[e].i1, edges[e].i2}, tol);
        // candidate pairs from grids (accel) or all pairs, in ascending order
        // either way, so the contact lists come out the same
        BoxGrid vg, eg;
        if (prm.accel) {
            std::vector<V3<float>> lo(nv), hi(nv);
            std::vector<char> pinned(nv);
            for (size_t i = 0; i < nv; i++) {
                lo[i] = vb[i].lo;
                hi[i] = vb[i].hi;
                pinned[i] = invMass[i] == R(0);
            }
            vg.build(lo, hi, &pinned);
            lo.resize(ne);
            hi.resize(ne);
            for (size_t e = 0; e < ne; e++) {
                lo[e] = eb[e].lo;
                hi[e] = eb[e].hi;
            }
            eg.build(lo, hi);
        }
        auto each = [&](std::vector<int>& cand, const BoxGrid& g, const Box& q, size_t count, size_t first, auto f) {
            if (prm.accel) {
                g.query(q.lo, q.hi, cand);
                for (const int i : cand)
                    if ((size_t)i >= first) f((size_t)i);
            } else {
                for (size_t i = first; i < count; i++) f(i);
            }
        };
        selfCandVT.clear();
        selfCandEE.clear();
        struct Part {
            std::vector<SelfContact<R>> contacts;
            std::vector<std::pair<int, int>> cand;
        };
        auto merge = [](std::vector<Part>& parts, std::vector<SelfContact<R>>& list,
                        std::vector<std::pair<int, int>>& cands) {
            for (Part& p : parts) {
                list.insert(list.end(), p.contacts.begin(), p.contacts.end());
                cands.insert(cands.end(), p.cand.begin(), p.cand.end());
            }
        };
        auto vt = ordered_blocks<Part>(nt, [&](size_t b, size_t end, Part& out) {
            std::vector<int> cand;
            for (size_t t = b; t < end; t++) {
                const uint32_t j1 = tris[t].i1, j2 = tris[t].i2, j3 = tris[t].i3;
                each(cand, vg, tb[t], nv, 0, [&](size_t i) {
                    if (i == j1 || i == j2 || i == j3 || invMass[i] == R(0)) return;
                    if (!overlap(tb[t], vb[i])) return;
                    if (prm.recordCandidates) out.cand.push_back({(int)t, (int)i});
                    SelfContact<R> c;
                    if (vertex_triangle(j1, j2, j3, (uint32_t)i, c)) {
                        c.key = (uint64_t(t) << 32) | uint64_t(i);
                        out.contacts.push_back(c);
                    }
                });
            }
        });
        merge(vt, selfTris, selfCandVT);
        auto ee = ordered_blocks<Part>(ne ? ne - 1 : 0, [&](size_t b, size_t end, Part& out) {
            std::vector<int> cand;
            for (size_t e1 = b; e1 < end; e1++) {
                each(cand, eg, eb[e1], ne, e1 + 1, [&](size_t e2) {
                    const uint32_t i1 = edges[e1].i1, i2 = edges[e1].i2, i3 = edges[e2].i1, i4 = edges[e2].i2;
                    if (i1 == i3 || i2 == i4 || i1 == i4 || i2 == i3) return;
                    if (!overlap(eb[e1], eb[e2])) return;
                    if (prm.recordCandidates) out.cand.push_back({(int)e1, (int)e2});
                    SelfContact<R> c;
                    if (edge_edge(i1, i2, i3, i4, c)) {
                        c.key = (uint64_t(e1) << 32) | uint64_t(e2);
                        out.contacts.push_back(c);
                    }
                });
            }
        });
        merge(ee, selfEdges, selfCandEE);
    }

    // processing order of the self contacts, given as their keys
    void order_self(std::vector<SelfContact<R>>& list, const std::vector<uint64_t>& keys)
    {
        std::unordered_map<uint64_t, size_t> rank;
        for (size_t k = 0; k < keys.size(); k++) rank[keys[k]] = k;
        std::stable_sort(list.begin(), list.end(), [&](const SelfContact<R>& a, const SelfContact<R>& b) {
            auto ra = rank.find(a.key), rb = rank.find(b.key);
            const size_t xa = ra == rank.end() ? SIZE_MAX : ra->second;
            const size_t xb = rb == rank.end() ? SIZE_MAX : rb->second;
            return xa < xb;
        });
    }

    // ClothCollisionHandler.cpp:364-436
    R solve_self_tris(R h)
    {
        const R shrink = R(prm.selfShrink), soften = R(.9f);
        R error = 0;
        for (SelfContact<R>& c : selfTris) {
            R len0 = R(prm.thickness) * shrink;
            V3<R>& x1 = pos[c.i1];
            V3<R>& x2 = pos[c.i2];
            V3<R>& x3 = pos[c.i3];
            V3<R>& x4 = pos[c.i4];
            const R m1 = invMass[c.i1], m2 = invMass[c.i2], m3 = invMass[c.i3], m4 = invMass[c.i4];
            V3<R> p, coords;
            closest_pt_point_triangle(x4, x1, x2, x3, p, coords);
            // a collapsed triangle gives no closest point (the reference's
            // coordinates become NaN and spread to the whole cloth): skipped
            if (!std::isfinite(double(coords.x + coords.y + coords.z))) continue;
            c.w1 = coords.x; c.w2 = coords.y; c.w3 = coords.z;
            const V3<R> delta = x4 - p;
            V3<R> n = cross(x2 - x1, x3 - x1);
            if (length(n) == R(0)) continue;
            normalize(n);
            const R side = signed_volume(x4, x1, x2, x3);
            if (c.side * side <= R(0)) len0 = R(prm.thickness) * R(0.1f);
            if (dot(n, c.normal) < R(0)) n = -n;
            const R len = dot(n, delta);
            if (len > len0) continue;
            const R im = c.w1 * c.w1 * m1 + c.w2 * c.w2 * m2 + c.w3 * c.w3 * m3 + m4;
            const R s = R(1.f) / im;
            const R err = len - len0;
            c.error = err;
            const R absErr = smax(R(0.f), -err);
            if (absErr > error) error = absErr;
            const R dLambda = soften * s * err;
            const V3<R> disp = dLambda * n;
            c.lambda += -dLambda / h;
            x1 += disp * (c.w1 * m1);
            x2 += disp * (c.w2 * m2);
            x3 += disp * (c.w3 * m3);
            x4 -= disp * m4;
        }
        return error;
    }

    // ClothCollisionHandler.cpp:439-541
    R solve_self_edges(R h, int& count)
    {
        const R shrink = R(prm.selfShrink), soften = R(0.9f);
        R error = 0;
        count = 0;
        for (SelfContact<R>& c : selfEdges) {
            R thick = R(prm.thickness) * shrink;
            V3<R>& x1 = pos[c.i1];
            V3<R>& x2 = pos[c.i2];
            V3<R>& x3 = pos[c.i3];
            V3<R>& x4 = pos[c.i4];
            const R m1 = invMass[c.i1], m2 = invMass[c.i2], m3 = invMass[c.i3], m4 = invMass[c.i4];
            const R side = signed_volume(x1, x2, x3, x4);
            V3<R> p, q;
            R t1, t2;
            closest_pt_segm_segm(x1, x2, x3, x4, t1, t2, p, q);
            c.w1 = R(1) - t1;
            c.w2 = R(1) - t2;
            const R omw1 = R(1.f) - c.w1, omw2 = R(1.f) - c.w2;
            const V3<R> delta = q - p;
            V3<R> n = delta;
            // edges that touch exactly have no direction between them (the
            // reference normalises the zero vector into NaN, which spreads to
            // the whole cloth): use the contact's normal instead
            if (length(delta) == R(0)) n = c.normal;
            else normalize(n);
            if (!std::isfinite(double(n.x + n.y + n.z))) continue;
            if (dot(n, c.normal) < R(0.0f)) n = -n;
            if (side * c.side < R(0)) thick = R(prm.thickness) * R(0.1f);
            const R len = dot(n, delta);
            if (len > thick) continue;
            const R im = c.w1 * c.w1 * m1 + omw1 * omw1 * m2 + c.w2 * c.w2 * m3 + omw2 * omw2 * m4;
            const R s = R(1.f) / im;
            const R err = len - thick;
            c.error = err;
            const R absErr = smax(R(0.f), -err);
            if (absErr > error) error = absErr;
            const R dLambda = -soften * s * err;
            c.lambda += dLambda / h;
            const V3<R> disp = dLambda * n;
            if (prm.selfEdgeWeights == 0) {
                x1 -= disp * (omw1 * m1);
                x2 -= disp * (c.w1 * m2);
                x3 += disp * (omw2 * m3);
                x4 += disp * (c.w2 * m4);
            } else {
                x1 -= disp * (c.w1 * m1);
                x2 -= disp * (omw1 * m2);
                x3 += disp * (c.w2 * m3);
                x4 += disp * (omw2 * m4);
            }
            count++;
        }
        return error;
    }

    // Physics/ClothPBD.cpp:9-50 (omega = 1)
    void project_link(LinkC<R>& link, R k, R& error)
    {
        const R m1 = invMass[link.i1], m2 = invMass[link.i2];
        if (m1 + m2 == R(0)) return;
        const R mu = R(1.f) / (m1 + m2);
        V3<R> delta = pos[link.i1] - pos[link.i2];
        const R len0 = link.len;
        const R len = length(delta);
        if (len == R(0) || std::isnan(len) || std::isinf(len)) return;
        const R err = len - len0;
        scale(delta, R(1) / len);
        const R epsilon = len0 / (k * link.stiffness);
        const R s = R(1.f) / (R(1.f) + mu * epsilon);
        const R lambda = err * mu * s;
        const R e = s * err;
        if (std::fabs(e) > error) error = std::fabs(e);
        const V3<R> fc = (R(1.f) * lambda) * delta;
        pos[link.i1] -= fc * m1;
        pos[link.i2] += fc * m2;
    }

    R solve_links(R h)
    {
        const R h2 = h * h;
        const R k = h2 * R(prm.area) * R(prm.unit);
        R error = 0;
        if (linkOrder.empty()) {
            for (auto& l : links) project_link(l, k, error);
        } else {
            for (int i : linkOrder) project_link(links[i], k, error);
        }
        return error;
    }

    // Physics/ClothPBD.cpp:324-369
    void solve_gs(R h)
    {
        R cb = 0, cv = 0, sct = 0, sce = 0, cl = 0;
        int k = 0, sceCount = 0;
        while (true) {
            if (prm.fem) solve_triangles(h);
            else cl = solve_links(h);
            if (prm.dihedral && k % prm.bendFreq == 0) cb = solve_bends();
            cv = solve_contacts(h);
            const R ce = solve_edge_contacts();
            const R ct = solve_tri_contacts();
            sct = solve_self_tris(h);
            sce = solve_self_edges(h, sceCount);
            // the FEM model reports no link error: cl stays 0 in the exit test
            k++;
            const float eps = 0.01f;
            if (!prm.fixedIterations && cl < R(0.05f) && cb < R(0.005f) && sce < R(prm.selfCollThreshold) &&
                sct < R(prm.selfCollThreshold) && cv < R(eps) && ce < R(eps) && ct < R(eps))
                break;
            if (k >= prm.iterations) break;
        }
        stats.iterations = k;
        stats.bendError = (float)cb;
        stats.contactError = (float)cv;
        stats.selfVTError = (float)sct;
        stats.selfEEError = (float)sce;
        stats.selfEECount = sceCount;
        stats.linkError = (float)cl;
    }

    // Physics/ClothPBD.cpp:616-651 (SolveLinksJacobi): every link's
    // displacement from the positions at the start of the pass, then all
    // applied in link order
    void solve_links_jacobi(R h, R omega)
    {
        const R h2 = h * h;
        const R k = h2 * R(prm.area) * R(prm.unit);
        std::vector<V3<R>> disp(links.size());
        for (size_t i = 0; i < links.size(); i++) {
            const LinkC<R>& l = links[i];
            const R mu = R(1.f) / (invMass[l.i1] + invMass[l.i2]);
            V3<R> delta = pos[l.i1] - pos[l.i2];
            const R len0 = l.len;
            const R len = length(delta);
            scale(delta, R(1) / len);
            const R err = len - len0;
            const R epsilon = len0 / (k * l.stiffness);
            const R s = R(1.f) / (R(1.f) + mu * epsilon);
            const R lambda = err * mu * s;
            disp[i] = (omega * lambda) * delta;
        }
        for (size_t i = 0; i < links.size(); i++) {
            const LinkC<R>& l = links[i];
            pos[l.i1] -= disp[i] * invMass[l.i1];
            pos[l.i2] += disp[i] * invMass[l.i2];
        }
    }

    // Physics/ClothPBD.cpp:653-662 (SolveJacobi)
    void solve_jacobi(R h)
    {
        for (int k = 0; k < prm.iterations; ++k) solve_links_jacobi(h, R(0.5f));
    }

    // Physics/ClothPBD.cpp:371-478 (ComputeErrorAndNormal and SolveCR as
    // compiled: SKIP_ALPHA and SKIP_BETA are defined, :381-382)
    void solve_cr(R h, R alpha0)
    {
        const size_t nl = links.size();
        if (nl == 0) return;
        std::vector<R> r(nl), d(nl, R(0));
        std::vector<V3<R>> normal(nl);
        const int N = prm.iterations;
        for (int iter = 0; iter < N; iter++) {
            for (size_t i = 0; i < nl; i++) {
                const LinkC<R>& l = links[i];
                V3<R> delta = pos[l.i2] - pos[l.i1];
                const R c = (length(delta) - l.len) * l.stiffness * R(prm.unit);
                scale(delta, R(1.f) / length(delta));  // Vector3::Normalize
                normal[i] = delta;
                r[i] = -c;
            }
            // the beta schedule in FP32 with powf, as the reference's floats
            float betaf = (float)iter / (float)(N - 1);
            betaf = std::pow(betaf, 0.6f);
            const R beta = R(betaf);
            for (size_t i = 0; i < nl; i++) d[i] = r[i] + beta * d[i];
            const R alpha = alpha0;
            for (size_t i = 0; i < nl; i++) {
                const LinkC<R>& l = links[i];
                const R dLambda = alpha * d[i];
                const V3<R> disp = dLambda * normal[i];
                pos[l.i1] -= invMass[l.i1] * disp;
                pos[l.i2] += invMass[l.i2] * disp;
            }
        }
    }

    // Physics/ClothPBD.cpp:255-322
    void substep(R h)
    {
        substep_detect(h);
        substep_solve(h);
    }

    // The mouse spring (ClothTypes.h:85-91; ClothModel.inl:86-98 sets its
    // rest length 0 and stiffness 100)
    struct MouseSpring {
        bool active = false;
        int i[3] = {0, 0, 0};
        float a[3] = {0, 0, 0};
        V3<float> point{0, 0, 0};
        float len = 0.f, stiffness = 100.f;
    } mouse;

    // Physics/ClothPBD.cpp:267-282: pulls the barycentric point of the
    // spring's triangle towards the mouse point, pinned particles included
    void apply_mouse_spring(R h)
    {
        const MouseSpring& m = mouse;
        const V3<R> p = (R(m.a[0]) * pos[m.i[0]] + R(m.a[1]) * pos[m.i[1]]) + R(m.a[2]) * pos[m.i[2]];
        V3<R> delta = p - V3<R>{R(m.point.x), R(m.point.y), R(m.point.z)};
        const R len = length(delta);
        scale(delta, R(1) / len);
        const V3<R> disp = (R(m.stiffness) * (len - R(m.len))) * delta;
        for (int k = 0; k < 3; k++) {
            pos[m.i[k]] -= (h * R(m.a[k])) * disp;
            vel[m.i[k]] -= R(m.a[k]) * disp;
        }
    }

    // prediction, collision detection, constraint reset
    void substep_detect(R h)
    {
        const V3<R> gravity{R(0), R(-9.8f), R(0)};
        const R unit = prm.unit, drag = prm.drag;
        for (size_t i = 0; i < pos.size(); i++) {
            prev[i] = pos[i];
            if (invMass[i] == R(0)) continue;
            vel[i] += (h * gravity) * unit;
            scale(vel[i], drag);
            pos[i] += h * vel[i];
        }
        if (mouse.active) apply_mouse_spring(h);
        detect_collisions();
        for (auto& c : contacts) { c.lambda = 0; c.lambdaF = 0; }
    }

    // Gauss-Seidel iterations and velocity update
    void substep_solve(R h)
    {
        if (!selfTriOrder.empty()) order_self(selfTris, selfTriOrder);
        if (!selfEdgeOrder.empty()) order_self(selfEdges, selfEdgeOrder);
        selfTriOrder.clear();
        selfEdgeOrder.clear();
        if (prm.solver == 1) solve_gs(h);
        else if (prm.solver == 0) solve_jacobi(h);
        else solve_cr(h, R(0.1f));
        const R invH = R(1.f) / h;
        for (size_t i = 0; i < pos.size(); i++) {
            if (invMass[i] == R(0)) continue;
            vel[i] = invH * (pos[i] - prev[i]);
        }
    }

    // Physics/ClothPatch.cpp:323-335
    int step(float dt)
    {
        // grid-SDF edge queries without vertex contacts yield undefined
        // contacts in the reference (a stub query accepted at the border)
        for (const auto& c : world)
            if (c.type == CT_SDF && (prm.collFlags & CF_EDGES) && !(prm.collFlags & CF_VERTICES)) return -2;
        const int n = prm.numSteps;
        const R h = R(dt / n);
        for (int i = 0; i < n; i++) substep(h);
        return 0;
    }
};

struct Handle {
    bool real64;
    std::unique_ptr<Cloth<float>> f;
    std::unique_ptr<Cloth<double>> d;
};

template <class F>
static auto with(void* hp, F&& fn)
{
    Handle* h = (Handle*)hp;
    if (h->real64) return fn(*h->d);
    return fn(*h->f);
}

}  // namespace oracle

using namespace oracle;

extern "C" {

// OpenMP threads for the detection loops of later calls from this thread
void oc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int oc_get_threads() { return omp_get_max_threads(); }

void* oc_create(int real64)
{
    Handle* h = new Handle;
    h->real64 = real64 != 0;
    if (h->real64) h->d.reset(new Cloth<double>);
    else h->f.reset(new Cloth<float>);
    return h;
}

void oc_free(void* h) { delete (Handle*)h; }

int oc_set_param_f(void* h, const char* name, float v)
{
    return with(h, [&](auto& c) {
        std::string n(name);
        Params& p = c.prm;
        if (n == "unit") p.unit = v;
        else if (n == "thickness") p.thickness = v;
        else if (n == "tolerance") p.tolerance = v;
        else if (n == "shear") p.shear = v;
        else if (n == "bend") p.bend = v;
        else if (n == "stretch") p.stretch = v;
        else if (n == "poisson") p.poisson = v;
        else if (n == "friction") p.friction = v;
        else if (n == "area") p.area = v;
        else if (n == "drag") p.drag = v;
        else if (n == "selfShrink") p.selfShrink = v;
        else if (n == "selfThreshold") p.selfCollThreshold = v;
        else if (n == "ccdVolume") p.ccdVolume = v;
        else if (n == "ccdBisect") p.ccdBisect = v;
        else if (n == "ccdVtDistance") p.ccdVtDistance = v;
        else return -1;
        return 0;
    });
}

// scAddMouseSpring / scUpdateMouseSpring / scRemoveMouseSpring
// (ClothModel.h:150-152): active 0 removes the spring
void oc_set_mouse_spring(void* h, int active, int i1, int i2, int i3, float a1, float a2, float a3, float px, float py,
                         float pz)
{
    with(h, [&](auto& c) {
        c.mouse.active = active != 0;
        c.mouse.i[0] = i1; c.mouse.i[1] = i2; c.mouse.i[2] = i3;
        c.mouse.a[0] = a1; c.mouse.a[1] = a2; c.mouse.a[2] = a3;
        c.mouse.point = V3<float>{px, py, pz};
        c.mouse.len = 0.f;
        c.mouse.stiffness = 100.f;
        return 0;
    });
}

int oc_set_param_i(void* h, const char* name, int v)
{
    return with(h, [&](auto& c) {
        std::string n(name);
        Params& p = c.prm;
        if (n == "iterations") p.iterations = v;
        else if (n == "collFlags") p.collFlags = v;
        else if (n == "bendFreq") p.bendFreq = v;
        else if (n == "numSteps") p.numSteps = v;
        else if (n == "dihedral") p.dihedral = v != 0;
        else if (n == "fem") p.fem = v != 0;
        else if (n == "accel") p.accel = v;
        else if (n == "solver") p.solver = v;
        else if (n == "fixedIterations") p.fixedIterations = v != 0;
        else if (n == "selfEdgeWeights") p.selfEdgeWeights = v;
        else if (n == "recordCandidates") p.recordCandidates = v;
        else return -1;
        return 0;
    });
}

void oc_init_grid(void* h, int divX, int divY, float inc, float ox, float oy, float oz, int horizontal, int attached,
                  float incline)
{
    with(h, [&](auto& c) { c.init_grid(divX, divY, inc, V3<float>{ox, oy, oz}, horizontal != 0, attached != 0, incline); });
}

// A cloth from arrays: particles (positions, material coordinates, inverse
// masses) and triangles, each added with ClothModel::AddTriangle's rest-state
// precompute (ClothModel.cpp:54-92), then edges and dihedral bends as
// ClothPatch::Init does (ClothPatch.cpp:232-241).
void oc_init_arrays(void* h, int nv, const double* pos, const double* uv, const double* invMass, int nt,
                    const int* tri)
{
    with(h, [&](auto& c) {
        using R = typename std::remove_reference<decltype(c.pos[0].x)>::type;
        c.clear();
        c.pos.resize(nv); c.prev.resize(nv); c.vel.assign(nv, V3<R>{0, 0, 0});
        c.invMass.resize(nv); c.uvx.resize(nv); c.uvy.resize(nv);
        for (int i = 0; i < nv; i++) {
            c.pos[i] = V3<R>{R(pos[3 * i]), R(pos[3 * i + 1]), R(pos[3 * i + 2])};
            c.prev[i] = c.pos[i];
            c.invMass[i] = R(invMass[i]);
            c.uvx[i] = R(uv[2 * i]); c.uvy[i] = R(uv[2 * i + 1]);
        }
        for (int t = 0; t < nt; t++) c.add_triangle(tri[3 * t], tri[3 * t + 1], tri[3 * t + 2]);
        c.construct_edges();
        for (const Edge& e : c.edges) c.add_bend(e);
    });
}

int oc_add_walls(void* h, const float* box6)
{
    return with(h, [&](auto& c) {
        using R = typename std::remove_reference<decltype(c.pos[0].x)>::type;
        Collidable<R> w;
        w.type = CT_WALLS;
        w.bmin = V3<R>{R(box6[0]), R(box6[1]), R(box6[2])};
        w.bmax = V3<R>{R(box6[3]), R(box6[4]), R(box6[5])};
        c.world.push_back(w);
        return (int)c.world.size() - 1;
    });
}

int oc_add_sphere(void* h, float cx, float cy, float cz, float r)
{
    return with(h, [&](auto& c) {
        using R = typename std::remove_reference<decltype(c.pos[0].x)>::type;
        Collidable<R> s;
        s.type = CT_SPHERE;
        s.center = V3<R>{R(cx), R(cy), R(cz)};
        s.radius = r;
        c.world.push_back(s);
        return (int)c.world.size() - 1;
    });
}

// Math/Quaternion.cpp:116-147: qRotate(q, v) = (q (v q*)).v, the product of
// a vector and a quaternion taking the vector as (0, v); FP32 as the
// reference's Capsule stores its data.
static V3<float> quat_rotate(float qs, V3<float> qv, V3<float> v)
{
    const V3<float> iv{-qv.x, -qv.y, -qv.z};
    const float ts = 0.f * qs - dot(v, iv);
    const V3<float> tv = (0.f * iv + qs * v) + cross(v, iv);
    return (qs * tv + ts * qv) + cross(qv, tv);
}

// Capsule (Physics/Common.h; ClothCollisionSDF.h:62-68): axis ends
// centre +- qRotate(rot, (0, hh, 0)); rotation (qs; qx, qy, qz)
int oc_add_capsule(void* h, float cx, float cy, float cz, float r, float hh, float qs, float qx, float qy, float qz,
                   float* axisOut)
{
    return with(h, [&](auto& c) {
        using R = typename std::remove_reference<decltype(c.pos[0].x)>::type;
        const V3<float> ctr{cx, cy, cz};
        const V3<float> half = quat_rotate(qs, V3<float>{qx, qy, qz}, V3<float>{0.f, hh, 0.f});
        const V3<float> P = ctr + half, Q = ctr - half;
        if (axisOut) {
            axisOut[0] = P.x; axisOut[1] = P.y; axisOut[2] = P.z;
            axisOut[3] = Q.x; axisOut[4] = Q.y; axisOut[5] = Q.z;
        }
        Collidable<R> k;
        k.type = CT_CAPSULE;
        k.center = V3<R>{R(cx), R(cy), R(cz)};
        k.radius = r;
        k.axisP = V3<R>{R(P.x), R(P.y), R(P.z)};
        k.axisQ = V3<R>{R(Q.x), R(Q.y), R(Q.z)};
        c.world.push_back(k);
        return (int)c.world.size() - 1;
    });
}

// values: (nx+1)(ny+1)(nz+1) node samples, x fastest (Geometry/SDF.h:49-52)
int oc_add_grid_sdf(void* h, int nx, int ny, int nz, const float* box6, const float* values)
{
    return with(h, [&](auto& c) {
        using R = typename std::remove_reference<decltype(c.pos[0].x)>::type;
        GridSDF<R> g;
        g.create(nx, ny, nz, V3<R>{R(box6[0]), R(box6[1]), R(box6[2])}, V3<R>{R(box6[3]), R(box6[4]), R(box6[5])});
        for (size_t i = 0; i < g.vals.size(); i++) g.vals[i] = R(values[i]);
        g.compute_gradient();
        c.sdfs.push_back(g);
        Collidable<R> s;
        s.type = CT_SDF;
        s.sdf = (int)c.sdfs.size() - 1;
        c.world.push_back(s);
        return (int)c.world.size() - 1;
    });
}

// Query a grid SDF the way GridSDF::QueryPoint does; out = [val, grad(3), normal(3), cp(3)]
void oc_sdf_query(void* h, int sdf, float px, float py, float pz, double* out)
{
    with(h, [&](auto& c) {
        using R = typename std::remove_reference<decltype(c.pos[0].x)>::type;
        const GridSDF<R>& g = c.sdfs[sdf];
        V3<R> p{R(px), R(py), R(pz)};
        V3<R> gr = g.grad_at(p);
        V3<R> cp, n;
        R v = g.query_point(p, cp, n);
        double o[10] = {(double)v, (double)gr.x, (double)gr.y, (double)gr.z, (double)n.x, (double)n.y, (double)n.z,
                        (double)cp.x, (double)cp.y, (double)cp.z};
        memcpy(out, o, sizeof(o));
    });
}

int oc_sdf_node_count(void* h, int sdf)
{
    int n = 0;
    with(h, [&](auto& c) { n = (int)c.sdfs[sdf].grads.size(); });
    return n;
}

void oc_get_sdf_grads(void* h, int sdf, double* out)
{
    with(h, [&](auto& c) {
        const auto& g = c.sdfs[sdf];
        for (size_t i = 0; i < g.grads.size(); i++) {
            out[3 * i] = g.grads[i].x; out[3 * i + 1] = g.grads[i].y; out[3 * i + 2] = g.grads[i].z;
        }
    });
}

void oc_set_order(void* h, int kind, const int* order, int n)
{
    with(h, [&](auto& c) {
        std::vector<int>& o = kind == 0 ? c.triOrder : kind == 1 ? c.bendOrder : c.linkOrder;
        o.assign(order, order + n);
    });
}

void oc_get_counts(void* h, int* out)
{
    with(h, [&](auto& c) {
        out[0] = (int)c.pos.size();
        out[1] = (int)c.tris.size();
        out[2] = (int)c.edges.size();
        out[3] = (int)c.bends.size();
        out[4] = (int)c.contacts.size();
        out[5] = (int)c.links.size();
    });
}

// idx: T*3 ints; rest: T*11 doubles (du1 du2 dv1 dv2 det invDet lu0 lv0 area lu2 lv2)
void oc_get_triangles(void* h, int* idx, double* rest)
{
    with(h, [&](auto& c) {
        for (size_t k = 0; k < c.tris.size(); k++) {
            const auto& t = c.tris[k];
            idx[3 * k] = t.i1; idx[3 * k + 1] = t.i2; idx[3 * k + 2] = t.i3;
            double r[11] = {(double)t.du1, (double)t.du2, (double)t.dv1, (double)t.dv2, (double)t.det, (double)t.invDet,
                            (double)t.lu0, (double)t.lv0, (double)t.area, (double)t.lu2, (double)t.lv2};
            memcpy(rest + 11 * k, r, sizeof(r));
        }
    });
}

// out: L*2 ints; rest: L*2 doubles (len, stiffness)
void oc_get_links(void* h, int* out, double* rest)
{
    with(h, [&](auto& c) {
        for (size_t k = 0; k < c.links.size(); k++) {
            out[2 * k] = c.links[k].i1; out[2 * k + 1] = c.links[k].i2;
            rest[2 * k] = (double)c.links[k].len; rest[2 * k + 1] = (double)c.links[k].stiffness;
        }
    });
}

void oc_get_bends(void* h, int* out)
{
    with(h, [&](auto& c) {
        for (size_t k = 0; k < c.bends.size(); k++) {
            out[4 * k] = c.bends[k].i1; out[4 * k + 1] = c.bends[k].i2;
            out[4 * k + 2] = c.bends[k].i3; out[4 * k + 3] = c.bends[k].i4;
        }
    });
}

void oc_get_edges(void* h, int* out)
{
    with(h, [&](auto& c) {
        for (size_t k = 0; k < c.edges.size(); k++) {
            out[4 * k] = c.edges[k].i1; out[4 * k + 1] = c.edges[k].i2;
            out[4 * k + 2] = c.edges[k].t1; out[4 * k + 3] = c.edges[k].t2;
        }
    });
}

// which: 0 pos, 1 prev, 2 vel, 3 invMass (V doubles)
void oc_get_state(void* h, int which, double* out)
{
    with(h, [&](auto& c) {
        size_t n = c.pos.size();
        if (which == 3) {
            for (size_t i = 0; i < n; i++) out[i] = c.invMass[i];
            return;
        }
        const auto& a = which == 0 ? c.pos : which == 1 ? c.prev : c.vel;
        for (size_t i = 0; i < n; i++) {
            out[3 * i] = a[i].x; out[3 * i + 1] = a[i].y; out[3 * i + 2] = a[i].z;
        }
    });
}

void oc_set_state(void* h, int which, const double* in)
{
    with(h, [&](auto& c) {
        using R = typename std::remove_reference<decltype(c.pos[0].x)>::type;
        size_t n = c.pos.size();
        if (which == 3) {
            for (size_t i = 0; i < n; i++) c.invMass[i] = R(in[i]);
            return;
        }
        auto& a = which == 0 ? c.pos : which == 1 ? c.prev : c.vel;
        for (size_t i = 0; i < n; i++) a[i] = V3<R>{R(in[3 * i]), R(in[3 * i + 1]), R(in[3 * i + 2])};
    });
}

// contacts: idx (int), coll (int), point(3), normal(3), lambda, lambdaF -> 8 doubles each
void oc_get_contacts(void* h, int* ids, double* data)
{
    with(h, [&](auto& c) {
        for (size_t k = 0; k < c.contacts.size(); k++) {
            const auto& ct = c.contacts[k];
            ids[2 * k] = ct.idx; ids[2 * k + 1] = ct.coll;
            double r[8] = {(double)ct.point.x, (double)ct.point.y, (double)ct.point.z, (double)ct.normal.x,
                           (double)ct.normal.y, (double)ct.normal.z, (double)ct.lambda, (double)ct.lambdaF};
            memcpy(data + 8 * k, r, sizeof(r));
        }
    });
}

void oc_get_stats(void* h, double* out)
{
    with(h, [&](auto& c) {
        out[0] = c.stats.iterations;
        out[1] = c.stats.bendError;
        out[2] = c.stats.contactError;
        out[3] = c.stats.selfVTError;
        out[4] = c.stats.selfEEError;
        out[5] = c.stats.selfEECount;
        out[6] = c.stats.linkError;
    });
}

int oc_step(void* h, float dt)
{
    return with(h, [&](auto& c) { return c.step(dt); });
}

void oc_substep(void* h, float hstep)
{
    with(h, [&](auto& c) {
        using R = typename std::remove_reference<decltype(c.pos[0].x)>::type;
        c.substep(R(hstep));
    });
}

// The two halves of a substep, so that a test can hand the self contacts'
// processing order in between.
void oc_substep_detect(void* h, float hstep)
{
    with(h, [&](auto& c) {
        using R = typename std::remove_reference<decltype(c.pos[0].x)>::type;
        c.substep_detect(R(hstep));
    });
}
void oc_substep_solve(void* h, float hstep)
{
    with(h, [&](auto& c) {
        using R = typename std::remove_reference<decltype(c.pos[0].x)>::type;
        c.substep_solve(R(hstep));
    });
}

// one pass over the self contacts (kind 0: vertex-triangle, 1: edge-edge)
double oc_solve_self(void* h, int kind, float hstep)
{
    return with(h, [&](auto& c) {
        using R = typename std::remove_reference<decltype(c.pos[0].x)>::type;
        int count = 0;
        return kind == 0 ? (double)c.solve_self_tris(R(hstep)) : (double)c.solve_self_edges(R(hstep), count);
    });
}

// kind 0: vertex-triangle, 1: edge-edge self contacts; processing order as keys
void oc_set_self_order(void* h, int kind, const uint64_t* keys, int n)
{
    with(h, [&](auto& c) { (kind == 0 ? c.selfTriOrder : c.selfEdgeOrder).assign(keys, keys + n); });
}

// self contacts in list order: keys, ids (i1 i2 i3 i4), data (w1 w2 w3,
// normal 3, side, lambda)
int oc_get_self_contacts(void* h, int kind, uint64_t* keys, int* ids, double* data, int cap)
{
    return with(h, [&](auto& c) {
        const auto& list = kind == 0 ? c.selfTris : c.selfEdges;
        const int n = (int)list.size();
        for (int k = 0; k < n && k < cap; k++) {
            const auto& x = list[k];
            keys[k] = x.key;
            ids[4 * k] = x.i1; ids[4 * k + 1] = x.i2; ids[4 * k + 2] = x.i3; ids[4 * k + 3] = x.i4;
            const double r[8] = {(double)x.w1, (double)x.w2, (double)x.w3, (double)x.normal.x,
                                 (double)x.normal.y, (double)x.normal.z, (double)x.side, (double)x.lambda};
            memcpy(data + 8 * k, r, sizeof(r));
        }
        return n;
    });
}

// Pieces of the substep, for tests that pin one stage at a time.
void oc_detect(void* h) { with(h, [&](auto& c) { c.detect_collisions(); }); }
void oc_solve_triangles(void* h, float hstep)
{
    with(h, [&](auto& c) {
        using R = typename std::remove_reference<decltype(c.pos[0].x)>::type;
        c.solve_triangles(R(hstep));
    });
}
double oc_solve_bends(void* h) { return with(h, [&](auto& c) { return (double)c.solve_bends(); }); }
double oc_solve_links(void* h, float step) { return with(h, [&](auto& c) { return (double)c.solve_links(step); }); }
double oc_solve_contacts(void* h, float hstep)
{
    return with(h, [&](auto& c) {
        using R = typename std::remove_reference<decltype(c.pos[0].x)>::type;
        return (double)c.solve_contacts(R(hstep));
    });
}

double oc_solve_edge_contacts(void* h) { return with(h, [&](auto& c) { return (double)c.solve_edge_contacts(); }); }
double oc_solve_tri_contacts(void* h) { return with(h, [&](auto& c) { return (double)c.solve_tri_contacts(); }); }

float oc_acos(float x) { return acos_as(x); }

// A static triangle-mesh obstacle (CollisionMesh, Physics/Common.h:113-138)
int oc_add_mesh(void* h, int nv, const float* verts, int ni, const unsigned* idx)
{
    return with(h, [&](auto& c) {
        using R = typename std::remove_reference<decltype(c.pos[0].x)>::type;
        ObstacleMesh M;
        M.v.resize(nv);
        for (int i = 0; i < nv; i++) M.v[i] = V3<float>{verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]};
        M.idx.assign(idx, idx + ni);
        M.build();
        c.meshes.push_back(M);
        Collidable<R> col;
        col.type = CT_MESH;
        col.mesh = (int)c.meshes.size() - 1;
        c.world.push_back(col);
        return (int)c.world.size() - 1;
    });
}

// The obstacle mesh `mesh`'s interpolated triangle normals at (tris[k],
// coords[3k..3k+2]) and edge normals at (edges[k], params[k]); edgesOut
// receives the mesh's edges (i1, i2) when not null. Returns the edge count.
int oc_mesh_normals_at(void* h, int mesh, const int* tris, const float* coords, int nt, const int* edges,
                       const float* params, int ne, float* triOut, float* edgeOut, int* edgesOut)
{
    return with(h, [&](auto& c) {
        using R = typename std::remove_reference<decltype(c.pos[0].x)>::type;
        const ObstacleMesh& M = c.meshes[mesh];
        for (int k = 0; k < nt; k++) {
            const V3<R> n = c.mesh_triangle_normal(M, tris[k], V3<R>{R(coords[3 * k]), R(coords[3 * k + 1]),
                                                                      R(coords[3 * k + 2])});
            triOut[3 * k] = (float)n.x; triOut[3 * k + 1] = (float)n.y; triOut[3 * k + 2] = (float)n.z;
        }
        for (int k = 0; k < ne; k++) {
            const V3<R> n = c.mesh_edge_normal(M, edges[k], R(params[k]));
            edgeOut[3 * k] = (float)n.x; edgeOut[3 * k + 1] = (float)n.y; edgeOut[3 * k + 2] = (float)n.z;
        }
        if (edgesOut)
            for (size_t e = 0; e < M.edges.size(); e++) {
                edgesOut[2 * e] = (int)M.edges[e].i1;
                edgesOut[2 * e + 1] = (int)M.edges[e].i2;
            }
        return (int)M.edges.size();
    });
}

// The last self detection's candidate pairs of one kind (0: (triangle,
// vertex), 1: (edge1, edge2)) in enumeration order; needs recordCandidates.
// Returns the count, writing at most cap pairs.
long oc_get_self_candidates(void* h, int kind, int* out, long cap)
{
    long n = 0;
    with(h, [&](auto& c) {
        const auto& v = kind == 0 ? c.selfCandVT : c.selfCandEE;
        n = (long)v.size();
        for (long k = 0; k < n && k < cap; k++) {
            out[2 * k] = v[k].first;
            out[2 * k + 1] = v[k].second;
        }
    });
    return n;
}

// Physics/ClothPatch.cpp:43-133; uv (2 nv floats) and colorsX (nv red
// components) may be null; mapOut (nv ints) receives vertex -> particle
void oc_init_mesh(void* h, int nv, const float* verts, int ni, const unsigned* idx, const float* uv,
                  const float* colorsX, float ox, float oy, float oz, int attached, int* mapOut)
{
    with(h, [&](auto& c) {
        std::vector<V3<float>> v(nv);
        for (int i = 0; i < nv; i++) v[i] = V3<float>{verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]};
        std::vector<uint32_t> ix(idx, idx + ni);
        std::vector<int> map;
        c.init_mesh(v, ix, uv, colorsX, V3<float>{ox, oy, oz}, attached != 0, map);
        if (mapOut) std::copy(map.begin(), map.end(), mapOut);
        return 0;
    });
}

// Processing order of edge (kind 2) or triangle (kind 3) contacts, given as
// the cloth triangle / edge indices in order.
void oc_set_contact_order(void* h, int kind, const int* order, int n)
{
    with(h, [&](auto& c) {
        std::vector<int>& rank = kind == 3 ? c.triContactRank : c.edgeContactRank;
        rank.assign(n, 0);
        for (int k = 0; k < n; k++) rank[order[k]] = k;
    });
}

// edge contacts: ids (i1, i2, clothEdge, feature), data (point 3, normal 3, w1, w2, lambda, lambdaF)
int oc_get_edge_contacts(void* h, int* ids, double* data, int cap)
{
    return with(h, [&](auto& c) {
        int n = (int)c.edgeContacts.size();
        for (int k = 0; k < n && k < cap; k++) {
            const auto& e = c.edgeContacts[k];
            ids[4 * k] = e.i1; ids[4 * k + 1] = e.i2; ids[4 * k + 2] = e.clothEdge; ids[4 * k + 3] = e.feature;
            double r[10] = {(double)e.point.x, (double)e.point.y, (double)e.point.z, (double)e.normal.x,
                            (double)e.normal.y, (double)e.normal.z, (double)e.w1, (double)e.w2, (double)e.lambda,
                            (double)e.lambdaF};
            memcpy(data + 10 * k, r, sizeof(r));
        }
        return n;
    });
}

// triangle contacts: ids (clothTri, feature), data (point 3, normal 3, w1, w2, w3, lambda, lambdaF)
int oc_get_tri_contacts(void* h, int* ids, double* data, int cap)
{
    return with(h, [&](auto& c) {
        int n = (int)c.triContacts.size();
        for (int k = 0; k < n && k < cap; k++) {
            const auto& t = c.triContacts[k];
            ids[2 * k] = t.clothTri; ids[2 * k + 1] = t.feature;
            double r[11] = {(double)t.point.x, (double)t.point.y, (double)t.point.z, (double)t.normal.x,
                            (double)t.normal.y, (double)t.normal.z, (double)t.w1, (double)t.w2, (double)t.w3,
                            (double)t.lambda, (double)t.lambdaF};
            memcpy(data + 11 * k, r, sizeof(r));
        }
        return n;
    });
}

// contact features of the vertex contacts: (mesh, feature) per contact
void oc_get_contact_features(void* h, int* out)
{
    with(h, [&](auto& c) {
        for (size_t k = 0; k < c.contacts.size(); k++) {
            out[2 * k] = c.contacts[k].mesh;
            out[2 * k + 1] = c.contacts[k].feature;
        }
    });
}

// Geometry primitives in FP32 for the golden-vector tests
int oc_closest_pt_point_triangle(const float* p, const float* a, const float* b, const float* c, float* out)
{
    V3<float> cp, bar;
    int r = closest_pt_point_triangle(V3<float>{p[0], p[1], p[2]}, V3<float>{a[0], a[1], a[2]},
                                      V3<float>{b[0], b[1], b[2]}, V3<float>{c[0], c[1], c[2]}, cp, bar);
    float o[6] = {cp.x, cp.y, cp.z, bar.x, bar.y, bar.z};
    memcpy(out, o, sizeof(o));
    return r;
}

int oc_closest_pt_segm_segm(const float* p1, const float* q1, const float* p2, const float* q2, float* out)
{
    float s, t;
    V3<float> c1, c2;
    int r = closest_pt_segm_segm(V3<float>{p1[0], p1[1], p1[2]}, V3<float>{q1[0], q1[1], q1[2]},
                                 V3<float>{p2[0], p2[1], p2[2]}, V3<float>{q2[0], q2[1], q2[2]}, s, t, c1, c2);
    float o[8] = {s, t, c1.x, c1.y, c1.z, c2.x, c2.y, c2.z};
    memcpy(out, o, sizeof(o));
    return r;
}

// out = t, closest point (3)
void oc_closest_pt_segm(const float* p, const float* a, const float* b, float* out)
{
    float t = 0.f;
    const V3<float> cp =
        closest_pt_segm(V3<float>{p[0], p[1], p[2]}, V3<float>{a[0], a[1], a[2]}, V3<float>{b[0], b[1], b[2]}, t);
    out[0] = t; out[1] = cp.x; out[2] = cp.y; out[3] = cp.z;
}

// out = dist, param, bar (3), closest on the segment (3), closest on the triangle (3)
void oc_closest_pt_segm_triangle(const float* p, const float* q, const float* a, const float* b, const float* c,
                                 float* out)
{
    float param = 0.f;
    V3<float> bar{0, 0, 0}, ce{0, 0, 0}, ct{0, 0, 0};
    const float d = closest_pt_segm_triangle(V3<float>{p[0], p[1], p[2]}, V3<float>{q[0], q[1], q[2]},
                                             V3<float>{a[0], a[1], a[2]}, V3<float>{b[0], b[1], b[2]},
                                             V3<float>{c[0], c[1], c[2]}, param, bar, ce, ct);
    float o[11] = {d, param, bar.x, bar.y, bar.z, ce.x, ce.y, ce.z, ct.x, ct.y, ct.z};
    memcpy(out, o, sizeof(o));
}

float oc_signed_volume(const float* a0, const float* a1, const float* a2, const float* a3)
{
    return signed_volume(V3<float>{a0[0], a0[1], a0[2]}, V3<float>{a1[0], a1[1], a1[2]}, V3<float>{a2[0], a2[1], a2[2]},
                         V3<float>{a3[0], a3[1], a3[2]});
}

// ObstacleMesh::build normals (vertex normals, nv*3 floats)
void oc_mesh_normals(int nv, const float* verts, int ni, const unsigned* idx, float* out)
{
    ObstacleMesh M;
    M.v.resize(nv);
    for (int i = 0; i < nv; i++) M.v[i] = V3<float>{verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]};
    M.idx.assign(idx, idx + ni);
    M.build();
    for (int i = 0; i < nv; i++) { out[3 * i] = M.vn[i].x; out[3 * i + 1] = M.vn[i].y; out[3 * i + 2] = M.vn[i].z; }
}

// Mesh::ConstructEdges on an arbitrary index list; out = (i1, i2, t1, t2) per
// unique edge. Returns the number of edges.
int oc_construct_edges(const unsigned* idx, int ni, int* out, int cap)
{
    Cloth<float> c;
    c.meshIndices.assign(idx, idx + ni);
    c.construct_edges();
    int n = (int)c.edges.size();
    for (int e = 0; e < n && e < cap; e++) {
        out[4 * e] = c.edges[e].i1; out[4 * e + 1] = c.edges[e].i2;
        out[4 * e + 2] = c.edges[e].t1; out[4 * e + 3] = c.edges[e].t2;
    }
    return n;
}

// SDF::Create + SDF::CreateFromMesh (Geometry/SDF.cpp:9-39, 58-76) over an
// (nx, ny, nz)-cell grid of box6 = (min, max): node (x, y, z) sits at
// (x, y, z) * steps + min (SDF.h:64-70) with steps = (max - min) * (1 / n)
// (SDF.cpp:37-38). Its value is ClosestPointOnMeshToPointAcc's distance, the
// smallest FP32 distance to a triangle (ClosestTriKernel,
// ClosestPointOnMesh.cpp:99-125; distance 0 skipped), FLT_MAX when there is
// none. The normal that query returns is the unit delta, so SDF.cpp:72-74's
// sign test never flips the value: the field is unsigned. The tree only
// prunes, so the minimum is taken over all triangles here. Nodes x-fastest,
// (nx + 1) (ny + 1) (nz + 1) values.
void oc_sdf_from_mesh(int nv, const float* verts, int ni, const unsigned* idx, int nx, int ny, int nz,
                      const float* box6, float* out)
{
    std::vector<V3<float>> v(nv);
    for (int i = 0; i < nv; i++) v[i] = V3<float>{verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]};
    const V3<float> mn{box6[0], box6[1], box6[2]}, mx{box6[3], box6[4], box6[5]};
    const V3<float> ext = mx - mn;
    const V3<float> steps{ext.x * (1.f / (float)nx), ext.y * (1.f / (float)ny), ext.z * (1.f / (float)nz)};
    const long vx = nx + 1, vy = ny + 1, count = vx * vy * (nz + 1);
    const int nt = ni / 3;
#pragma omp parallel for schedule(static)
    for (long k = 0; k < count; k++) {
        const long x = k % vx, w = k / vx, y = w % vy, z = w / vy;
        const V3<float> org{(float)x * steps.x + mn.x, (float)y * steps.y + mn.y, (float)z * steps.z + mn.z};
        float best = FLT_MAX;
        for (int t = 0; t < nt; t++) {
            V3<float> pt, bar;
            closest_pt_point_triangle(org, v[idx[3 * t]], v[idx[3 * t + 1]], v[idx[3 * t + 2]], pt, bar);
            const float d = length(org - pt);
            if (d == 0.f) continue;
            if (d < best) best = d;
        }
        out[k] = best;
    }
}

}  // extern "C"
