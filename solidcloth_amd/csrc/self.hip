// gfx950 kernels for continuous self collision of the cloth:
//   swept boxes -> LBVH refit -> per-vertex / per-edge overlap walks with the
//   vertex-triangle / edge-edge CCD test (Physics/SelfCollisionDetector.cpp,
//   coplanarity cubic in FP64) -> deterministic colouring of the contacts ->
//   coloured Gauss-Seidel passes (Physics/ClothCollisionHandler.cpp:364-541).
// FP32 arithmetic follows the reference's operation order; the cubic follows
// it in FP64, as the CPU oracle does.
#include "devutil.h"
#include "geom.h"
#include "kernels.h"

namespace sc {

static constexpr int kBlock = 256;
static constexpr int kSoloBlock = 1024;
static constexpr int kCcdBlocks = 4096;

// --------------------------------------------------------------- FP64 ----

struct d3 {
    double x, y, z;
};
__device__ __forceinline__ d3 dd(f3 a) { return d3{(double)a.x, (double)a.y, (double)a.z}; }
__device__ __forceinline__ d3 dadd(d3 a, d3 b) { return d3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ d3 dsub(d3 a, d3 b) { return d3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ d3 dmul(double s, d3 a) { return d3{s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ double ddot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ d3 dcross(d3 a, d3 b)
{
    return d3{a.y * b.z - b.y * a.z, -a.x * b.z + a.z * b.x, a.x * b.y - b.x * a.y};
}
// Math/Vector3.h:289-292: triple(a, b, c) = c . (a x b)
__device__ __forceinline__ double dtriple(d3 a, d3 b, d3 c) { return ddot(c, dcross(a, b)); }

// SelfCollisionDetector.cpp:263-276: volume spanned by the four moving points at t
__device__ __forceinline__ double coplanarity(double t, const d3* x, const d3* v)
{
    const d3 z1 = dadd(x[0], dmul(t, v[0])), z2 = dadd(x[1], dmul(t, v[1]));
    const d3 z3 = dadd(x[2], dmul(t, v[2])), z4 = dadd(x[3], dmul(t, v[3]));
    return ddot(dcross(dsub(z2, z1), dsub(z3, z1)), dsub(z4, z1));
}

// SelfCollisionDetector.cpp:414-488
__device__ bool check_interval(double ta, double tb, const d3* x, const d3* v, double eps, double& tx)
{
    double tx0 = ta, tx1 = tb;
    double fx0 = coplanarity(tx0, x, v), fx1 = coplanarity(tx1, x, v);
    if (fx0 * fx1 > 0) return false;
    tx = tx0;
    for (int l = 0; l < 40; l++) {
        const double tmid = 0.5 * (tx0 + tx1);
        const double fmid = coplanarity(tmid, x, v);
        if (fabs(fmid) < eps) {
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

// SelfCollisionDetector.cpp:278-412: first time in [0, 1] at which the four
// points become coplanar (cubic in t, split at its critical points).
__device__ bool solve_cubic(const f3* xf, const f3* vf, const SelfParams& sp, double& toi)
{
    d3 x[4], v[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        x[k] = dd(xf[k]);
        v[k] = dd(vf[k]);
    }
    const double volEps = sp.volume;
    const double f0 = coplanarity(0.0, x, v);
    if (fabs(f0) < volEps) return false;
    const d3 v10 = dsub(v[0], v[3]), v20 = dsub(v[1], v[3]), v30 = dsub(v[2], v[3]);
    const d3 x10 = dsub(x[0], x[3]), x20 = dsub(x[1], x[3]), x30 = dsub(x[2], x[3]);
    const double a = dtriple(v10, v20, v30);
    const double b = dtriple(x10, v20, v30) + dtriple(v10, x20, v30) + dtriple(v10, v20, x30);
    const double c = dtriple(x10, x20, v30) + dtriple(x10, v20, x30) + dtriple(v10, x20, x30);
    if (fabs(a) < volEps) {
        if (fabs(b) < volEps) return false;
        const double d = dtriple(x10, x20, x30);
        const double discr = c * c - 4 * b * d;
        if (discr < 0) return false;
        if (discr == 0) {
            const double t = -0.5 * c / b;
            if (t < 0 || t > 1) return false;
            toi = t;
            return true;
        }
        const double sq = sqrt(discr);
        const double t1 = 0.5 * (-c - sq) / b, t2 = 0.5 * (-c + sq) / b;
        const double t = t2 < t1 ? t2 : t1;
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
        const double sq = sqrt(discr);
        t1 = (-b - sq) / (3 * a);
        t2 = (-b + sq) / (3 * a);
        if (t2 < t1) {
            const double tmp = t1;
            t1 = t2;
            t2 = tmp;
        }
        np = 2;
    }
    const double eps = sp.bisect;
    if (np == 0) return check_interval(0, 1, x, v, eps, toi);
    if (np == 1) {
        if (t1 < 0 || t1 > 1) return check_interval(0, 1, x, v, eps, toi);
        return check_interval(0, t1, x, v, eps, toi) || check_interval(t1, 1, x, v, eps, toi);
    }
    if (t1 > 1 || t2 < 0 || (t1 < 0 && t2 > 1)) return check_interval(0, 1, x, v, eps, toi);
    if (t1 < 0 && t2 < 1) return check_interval(0, t2, x, v, eps, toi) || check_interval(t2, 1, x, v, eps, toi);
    if (t1 < 1 && t2 > 1) return check_interval(0, t1, x, v, eps, toi) || check_interval(t1, 1, x, v, eps, toi);
    if (t1 > 0 && t2 < 1)
        return check_interval(0, t1, x, v, eps, toi) || check_interval(t1, t2, x, v, eps, toi) ||
               check_interval(t2, 1, x, v, eps, toi);
    return false;
}

// ------------------------------------------------------------ detection ----

__device__ __forceinline__ bool overlap(f3 alo, f3 ahi, float4 blo, float4 bhi)
{
    return !(ahi.x < blo.x || ahi.y < blo.y || ahi.z < blo.z || bhi.x < alo.x || bhi.y < alo.y || bhi.z < alo.z);
}

// overlap without short-circuit evaluation (no branches on the way)
__device__ __forceinline__ bool overlap_nb(f3 alo, f3 ahi, float4 blo, float4 bhi)
{
    return !((ahi.x < blo.x) | (ahi.y < blo.y) | (ahi.z < blo.z) | (bhi.x < alo.x) | (bhi.y < alo.y) |
             (bhi.z < alo.z));
}

// ClothCollision.cpp:256-315: prev and pos of the primitive's vertices,
// extruded by tol (Aabb3::Extrude).
__device__ __forceinline__ void swept(f3 p, f3& lo, f3& hi)
{
    lo = f3{fminf(lo.x, p.x), fminf(lo.y, p.y), fminf(lo.z, p.z)};
    hi = f3{fmaxf(hi.x, p.x), fmaxf(hi.y, p.y), fmaxf(hi.z, p.z)};
}

__global__ void __launch_bounds__(kBlock) k_self_boxes(const float4* __restrict__ pos, const float4* __restrict__ prev,
                                                       const int4* __restrict__ cTris, int nt,
                                                       const int2* __restrict__ cEdges, int ne, float tol,
                                                       float4* triLo, float4* triHi, float4* edgeLo, float4* edgeHi,
                                                       int nv, float4* vtxLo, float4* vtxHi)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (vtxLo && i >= nt + ne && i < nt + ne + nv) {
        const int v = i - nt - ne;
        f3 lo{1e7f, 1e7f, 1e7f}, hi{-1e7f, -1e7f, -1e7f};
        swept(xyz(prev[v]), lo, hi);
        swept(xyz(pos[v]), lo, hi);
        vtxLo[v] = make_float4(lo.x - tol, lo.y - tol, lo.z - tol, 0.f);
        vtxHi[v] = make_float4(hi.x + tol, hi.y + tol, hi.z + tol, 0.f);
        return;
    }
    f3 lo{1e7f, 1e7f, 1e7f}, hi{-1e7f, -1e7f, -1e7f};
    if (i < nt) {
        const int4 t = cTris[i];
        swept(xyz(prev[t.x]), lo, hi);
        swept(xyz(prev[t.y]), lo, hi);
        swept(xyz(prev[t.z]), lo, hi);
        swept(xyz(pos[t.x]), lo, hi);
        swept(xyz(pos[t.y]), lo, hi);
        swept(xyz(pos[t.z]), lo, hi);
        triLo[i] = make_float4(lo.x - tol, lo.y - tol, lo.z - tol, 0.f);
        triHi[i] = make_float4(hi.x + tol, hi.y + tol, hi.z + tol, 0.f);
    } else if (i < nt + ne) {
        const int2 e = cEdges[i - nt];
        swept(xyz(prev[e.x]), lo, hi);
        swept(xyz(prev[e.y]), lo, hi);
        swept(xyz(pos[e.x]), lo, hi);
        swept(xyz(pos[e.y]), lo, hi);
        edgeLo[i - nt] = make_float4(lo.x - tol, lo.y - tol, lo.z - tol, 0.f);
        edgeHi[i - nt] = make_float4(hi.x + tol, hi.y + tol, hi.z + tol, 0.f);
    }
}

// Appends one contact; the calling lanes of a wave share one atomic on the
// kind's counter (the leader reserves their places).
__device__ __forceinline__ void self_append(SelfBuf b, int kind, SolverCtl* ctl, int4 ids, float4 W, float4 N,
                                            unsigned long long key)
{
    const unsigned long long act = __ballot(true);
    const int lane = __lane_id(), leader = __ffsll(act) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(&ctl->selfCount[kind], __popcll(act));
    const int k = __builtin_amdgcn_readlane(base, leader) + __popcll(act & ((1ull << lane) - 1ull));
    if (k >= b.cap) {
        atomicExch(&ctl->selfOverflow, 1);
        return;
    }
    b.ids[k] = ids;
    b.W[k] = W;
    b.N[k] = N;
    b.key[k] = key;
}

// SelfCollisionDetector.cpp:214-261 (VertexTriangleCCD + VertexTriangleTest)
__device__ bool vertex_triangle(const float4* pos, const float4* prev, int i1, int i2, int i3, int i4,
                                const SelfParams& sp, float4& W, float4& N)
{
    const f3 x[4] = {xyz(prev[i1]), xyz(prev[i2]), xyz(prev[i3]), xyz(prev[i4])};
    const f3 v[4] = {sub(xyz(pos[i1]), x[0]), sub(xyz(pos[i2]), x[1]), sub(xyz(pos[i3]), x[2]),
                     sub(xyz(pos[i4]), x[3])};
    double toid;
    if (!solve_cubic(x, v, sp, toid)) return false;
    const float toi = (float)toid;
    const f3 z1 = add(x[0], mul(toi, v[0])), z2 = add(x[1], mul(toi, v[1]));
    const f3 z3 = add(x[2], mul(toi, v[2])), z4 = add(x[3], mul(toi, v[3]));
    f3 p, co;
    closest_pt_point_triangle(z4, z1, z2, z3, p, co);
    const f3 delta = sub(z4, p);
    const float eps = sp.vtDistance;
    if (!(dot(delta, delta) < eps * eps)) return false;
    const f3 triNormal = normalized(cross(sub(x[1], x[0]), sub(x[2], x[0])));
    const float side = dot(triNormal, sub(x[3], x[0]));
    f3 n = normalized(cross(sub(z2, z1), sub(z3, z1)));
    if (side < 0.f) n = neg(n);
    W = make_float4(co.x, co.y, co.z, signed_volume(x[3], x[0], x[1], x[2]));
    N = f4(n, 0.f);
    return true;
}

// SelfCollisionDetector.cpp:491-606 (EdgeEdgeTest + EdgeEdgeCCD)
__device__ bool edge_edge(const float4* pos, const float4* prev, int i1, int i2, int i3, int i4,
                          const SelfParams& sp, float4& W, float4& N)
{
    const f3 x[4] = {xyz(prev[i1]), xyz(prev[i2]), xyz(prev[i3]), xyz(prev[i4])};
    const f3 v[4] = {sub(xyz(pos[i1]), x[0]), sub(xyz(pos[i2]), x[1]), sub(xyz(pos[i3]), x[2]),
                     sub(xyz(pos[i4]), x[3])};
    double toid;
    if (!solve_cubic(x, v, sp, toid)) return false;
    const float toi = (float)toid;
    const f3 z1 = add(x[0], mul(toi, v[0])), z2 = add(x[1], mul(toi, v[1]));
    const f3 z3 = add(x[2], mul(toi, v[2])), z4 = add(x[3], mul(toi, v[3]));
    float s, t;
    f3 c1, c2;
    closest_pt_segm_segm(z1, z2, z3, z4, s, t, c1, c2);
    const f3 d = sub(c1, c2);
    const float eps = 0.5f * sp.thickness;
    if (!(dot(d, d) <= eps * eps)) return false;
    float s0, t0;
    closest_pt_segm_segm(x[0], x[1], x[2], x[3], s0, t0, c1, c2);
    W = make_float4(s, t, 0.f, signed_volume(x[0], x[1], x[2], x[3]));
    N = f4(normalized(sub(c2, c1)), 0.f);
    return true;
}

// The hardware hands workgroups to the eight XCDs round-robin (workgroup b
// to XCD b mod 8), and each XCD has its own L2. This maps hardware workgroup
// b to a logical one so that every XCD gets a contiguous run of logical
// workgroups: Morton-ordered queries of one XCD then share tree nodes in its
// L2. Bijective on [0, nb).
__device__ __forceinline__ int xcd_logical_block(int b, int nb)
{
    const int x = b & 7, k = b >> 3;
    int start = 0;
    for (int y = 0; y < x; y++) start += (nb - y + 7) >> 3;
    return start + k;
}

// Broad phase (SelfCollisionDetector.cpp:60-135 over the trees
// ClothModel::ComputeTree builds, ClothCollision.cpp:240-315): every vertex
// queries the triangle tree with its swept box, every edge the edge tree; a
// pair whose boxes overlap (closed intervals, as Aabb3::Intersect) is a
// candidate unless the vertex belongs to the triangle or the edges share a
// vertex (and, like the reference's pass over leaf pairs, a pinned vertex is
// never a candidate). Queries run in Morton order (the trees' leaf orders).
// An edge query pairs only with the edge tree's later leaves, so each edge
// pair is found once, and is written with the lower edge index first.
//
// The 64 queries of a wave walk the tree together: one node sequence for the
// wave (depth-first, left-first), a node entered when any lane's box overlaps
// it (and, for edges, any such lane still wants one of its leaves), each lane
// testing its own box at every leaf the walk reaches. Each lane thus meets
// exactly the leaves its own walk would, in the same order, and the control
// flow stays uniform. The records lie in that depth-first order, so the walk
// only moves forward: to the next record when it enters a node or leaves a
// leaf, past the subtree when it skips one. The wave holds two windows of 16
// consecutive records in registers (lane 4 r + q holds float4 q of record r)
// and reads a step's fields from them with lane reads; the window after the
// current one is requested when the walk enters the current one, so a walk
// waits on memory only where a skip lands beyond both. Pairs go to flat pools
// in chunks of kCandChunk slots: the lanes that find a pair at a leaf take
// consecutive slots of the wave's chunk (its base and fill are wave-uniform,
// in scalar registers), and a fresh chunk costs one atomic. The continuous
// tests run in k_self_ccd, so that the walk and the tests' FP64 registers do
// not meet in one kernel.
constexpr int kWalkWindow = 16;  // records per window (one float4 per lane)

template <bool kEdges>
__device__ __forceinline__ void wave_candidates(const BvhView& bv, bool active, f3 lo, f3 hi, int self, int first,
                                                int2 me, int2* pool, int* fill, int cap, int* reserved,
                                                SolverCtl* ctl)
{
    const int lane = __lane_id();
    int cbase = -1, cused = kCandChunk;  // the wave's chunk (no chunk yet)
    auto take = [&](bool hit, unsigned long long mask, int2 p) {
        if (!mask) return;
        const int cnt = __popcll(mask);
        if (cused + cnt > kCandChunk) {
            const int lead = __ffsll(mask) - 1;
            int got = 0;
            if (lane == lead) {
                if (cbase >= 0 && cbase < cap) fill[cbase / kCandChunk] = cused;
                got = atomicAdd(reserved, kCandChunk);
            }
            cbase = __builtin_amdgcn_readlane(got, lead);
            cused = 0;
        }
        if (hit) {
            const int slot = cbase + cused + __popcll(mask & ((1ull << lane) - 1ull));
            if (slot < cap)
                pool[slot] = p;
            else
                atomicExch(&ctl->selfOverflow, 1);
        }
        cused += cnt;
    };
    const int end = 2 * bv.n - 1;  // records
    if (bv.n >= 1 && __ballot(active)) {
        // this lane's share of a window starting at record `base` (records
        // past the end repeat the last one; the walk never reads them)
        const int sub = lane >> 2, part = lane & 3;
        auto window = [&](int base) { return bv.dfs[4 * (size_t)min(base + sub, end - 1) + part]; };
        auto fl = [](float v, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); };
        auto in = [](float v, int l) { return __builtin_amdgcn_readlane(__float_as_int(v), l); };
        // A holds records [wa, wa + 16), B the next 16 (perhaps in flight).
        // Waiting for A alone (vmcnt 1: B, the youngest, may stay in flight)
        // before the steps tells the compiler that A is in: otherwise the
        // pair stores of the steps, counted by the same counter, would make
        // it wait for B as well.
        constexpr int kWaitAllButOne = 0x0F71;  // vmcnt(1), expcnt and lgkmcnt not waited on
        int wa = 0, p = 0;
        float4 A = window(0), B = window(kWalkWindow);
        __builtin_amdgcn_s_waitcnt(kWaitAllButOne);
        // every step moves forward (an escape below p + 1 is taken as p + 1),
        // so there are at most `end` steps
        while (p < end) {
            // the steps inside window A: no memory access
            for (const int lim = min(wa + kWalkWindow, end); p < lim;) {
                const int r = 4 * (p - wa);
                // (min, primitive or -1), (max, escape), payload, (index, ...)
                const float4 L = make_float4(fl(A.x, r), fl(A.y, r), fl(A.z, r), 0.f);
                const float4 H = make_float4(fl(A.x, r + 1), fl(A.y, r + 1), fl(A.z, r + 1), 0.f);
                const int prim = in(A.w, r);
                const int idx = in(A.x, r + 3);  // leaf: its sorted index; node: the last one below it
                const bool ov = (int)active & (int)overlap_nb(lo, hi, L, H);
                if (prim < 0) {
                    const bool want = (int)ov & (int)(!kEdges | (idx > first));
                    p = __ballot(want) ? p + 1 : max(p + 1, in(A.w, r + 1));
                } else {
                    const int vx = in(A.x, r + 2), vy = in(A.y, r + 2), vz = in(A.z, r + 2);
                    bool hit = ov;
                    int2 pr;
                    if (kEdges) {
                        hit = (int)hit & (int)(idx > first) &
                              (int)!((me.x == vx) | (me.y == vy) | (me.x == vy) | (me.y == vx));
                        pr = self < prim ? make_int2(self, prim) : make_int2(prim, self);
                    } else {
                        hit = (int)hit & (int)!((self == vx) | (self == vy) | (self == vz));
                        pr = make_int2(prim, self);
                    }
                    take(hit, __ballot(hit), pr);
                    p = p + 1;
                }
                p = __builtin_amdgcn_readfirstlane(p);  // (the same on every lane)
            }
            if (p >= end) break;
            // the next window: B if the walk landed in it, else a fresh one
            if (p < wa + 2 * kWalkWindow) {
                wa += kWalkWindow;
                A = B;
            } else {
                wa = p;
                A = window(wa);
            }
            B = window(wa + kWalkWindow);
            __builtin_amdgcn_s_waitcnt(kWaitAllButOne);
        }
    }
    if (lane == 0 && cbase >= 0 && cbase < cap) fill[cbase / kCandChunk] = cused;
}

__global__ void __launch_bounds__(kBlock) k_self_candidates(const float4* __restrict__ pos,
                                                            const float4* __restrict__ prev,
                                                            const int4* __restrict__ cTris,
                                                            const int2* __restrict__ cEdges, int nv, int ne,
                                                            int vertexBlocks, BvhView triBvh, BvhView edgeBvh,
                                                            const int* __restrict__ vertexOrder, float tol,
                                                            SelfCandidates cand, SolverCtl* ctl)
{
    const int blk = xcd_logical_block(blockIdx.x, gridDim.x);
    const bool vertexPass = blk < vertexBlocks;  // uniform over the workgroup
    const int j = (vertexPass ? blk : blk - vertexBlocks) * blockDim.x + threadIdx.x;
    f3 lo{1e7f, 1e7f, 1e7f}, hi{-1e7f, -1e7f, -1e7f};
    if (vertexPass) {
        bool active = false;
        int v = 0;
        if (j < nv) {
            v = vertexOrder ? vertexOrder[j] : j;
            const float4 P = pos[v];
            if (P.w != 0.f) {
                active = true;
                swept(xyz(prev[v]), lo, hi);
                swept(xyz(P), lo, hi);
                lo = f3{lo.x - tol, lo.y - tol, lo.z - tol};
                hi = f3{hi.x + tol, hi.y + tol, hi.z + tol};
            }
        }
        wave_candidates<false>(triBvh, active, lo, hi, v, 0, make_int2(0, 0), cand.vt,
                               cand.fillVT, cand.capVT, &ctl->candSlots[0], ctl);
    } else {
        bool active = false;
        int e1 = 0;
        int2 ev = make_int2(-1, -1);
        if (j < ne) {
            active = true;
            e1 = edgeBvh.leafPrim[j];
            ev = cEdges[e1];
            swept(xyz(prev[ev.x]), lo, hi);
            swept(xyz(prev[ev.y]), lo, hi);
            swept(xyz(pos[ev.x]), lo, hi);
            swept(xyz(pos[ev.y]), lo, hi);
            lo = f3{lo.x - tol, lo.y - tol, lo.z - tol};
            hi = f3{hi.x + tol, hi.y + tol, hi.z + tol};
        }
        // query j (sorted leaf j) pairs with the overlapping leaves after it
        wave_candidates<true>(edgeBvh, active, lo, hi, e1, j, ev, cand.ee, cand.fillEE,
                              cand.capEE, &ctl->candSlots[1], ctl);
    }
}

__device__ __forceinline__ void ccd_vt(const float4* pos, const float4* prev, const int4* cTris, int2 c,
                                       const SelfParams& sp, SelfBuf vt, SolverCtl* ctl)
{
    const int4 tv = cTris[c.x];
    float4 W, N;
    if (vertex_triangle(pos, prev, tv.x, tv.y, tv.z, c.y, sp, W, N))
        self_append(vt, 0, ctl, make_int4(tv.x, tv.y, tv.z, c.y), W, N, ((unsigned long long)c.x << 32) | (unsigned)c.y);
}

__device__ __forceinline__ void ccd_ee(const float4* pos, const float4* prev, const int2* cEdges, int2 c,
                                       const SelfParams& sp, SelfBuf ee, SolverCtl* ctl)
{
    const int2 ev = cEdges[c.x], fv = cEdges[c.y];
    float4 W, N;
    if (edge_edge(pos, prev, ev.x, ev.y, fv.x, fv.y, sp, W, N))
        self_append(ee, 1, ctl, make_int4(ev.x, ev.y, fv.x, fv.y), W, N, ((unsigned long long)c.x << 32) | (unsigned)c.y);
}

// SelfCollisionDetector.cpp:137-190, 23-58: the continuous tests of the
// candidate pairs, one slot of the pools per thread (grid-stride; the
// reserved slot counts are read on the device, and a slot past its chunk's
// fill holds no pair).
__global__ void __launch_bounds__(kBlock) k_self_ccd(const float4* __restrict__ pos, const float4* __restrict__ prev,
                                                     const int4* __restrict__ cTris, const int2* __restrict__ cEdges,
                                                     SelfParams sp, SelfCandidates cand, SelfBuf vt, SelfBuf ee,
                                                     SolverCtl* ctl)
{
    const int nvt = min(ctl->candSlots[0], cand.capVT), nee = min(ctl->candSlots[1], cand.capEE);
    const int stride = gridDim.x * blockDim.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nvt + nee; i += stride) {
        if (i < nvt) {
            if (i % kCandChunk < cand.fillVT[i / kCandChunk]) ccd_vt(pos, prev, cTris, cand.vt[i], sp, vt, ctl);
        } else {
            const int k = i - nvt;
            if (k % kCandChunk < cand.fillEE[k / kCandChunk]) ccd_ee(pos, prev, cEdges, cand.ee[k], sp, ee, ctl);
        }
    }
}

// ------------------------------------------------------------ colouring ----

// Bid of a contact: a bijective 64-bit mix (splitmix64's finaliser) of its
// key, so that bids are unique and spatially neighbouring contacts do not
// form long chains of decreasing bids (plain keys would make the colouring
// take as many rounds as such a chain is long). No valid key maps to 0.
__device__ __forceinline__ unsigned long long bid_of(unsigned long long key)
{
    unsigned long long z = key ^ 0xffffffffffffffffull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// Jones-Plassmann colouring of one kind's contacts, by one workgroup. In
// round r every uncoloured contact bids at its four vertices; a contact that
// holds the highest bid at all four takes colour r. Winners of a round share
// no vertex, and the highest uncoloured bid always wins, so every round makes
// progress and there are at most as many colours as contacts; the result
// depends only on the set of contacts. The contacts are then listed colour by
// colour: counts per colour in groupOff (global memory, capacity + 1
// entries), an exclusive prefix sum over them, and a scatter through cursors
// kept in the (by then unused) round list.
//
// First the kind's distinct vertices are listed (vlist) and each contact's
// vertices renamed to their places in that list (lid): every vertex is
// claimed by the least (contact, corner) that names it (atomic minimum in
// vloc), its claimant appends it, and vloc holds its place until the
// entries are released. With at most kColourPerThread contacts per thread
// and kColourLdsVerts distinct vertices, the rounds then run on registers
// and LDS alone (the bids of the listed vertices in LDS); otherwise on the
// per-vertex bids in global memory (vmax), which this one workgroup owns
// (zero before and after it), with atomics, loads and stores scoped to the
// workgroup. Both give the same colours.
constexpr int kColourPerThread = 8;
constexpr int kColourLdsVerts = 12288;
constexpr int kVlocFree = 0x7f7f7f7f, kVlocSlot = 0x40000000;

__device__ __forceinline__ int corner(const int4 q, int j) { return j == 0 ? q.x : j == 1 ? q.y : j == 2 ? q.z : q.w; }

__global__ void __launch_bounds__(kSoloBlock) k_self_colour(SelfBuf bvt, SelfBuf bee, unsigned long long* vmax,
                                                            int* vloc, int nv, SolverCtl* ctl)
{
    // workgroup 0 colours the vertex-triangle contacts, workgroup 1 the
    // edge-edge ones, each with its own half of the per-vertex tables
    const int kind = blockIdx.x;
    const SelfBuf b = kind == 0 ? bvt : bee;
    vmax += (size_t)kind * nv;
    vloc += (size_t)kind * nv;
    __shared__ unsigned long long lbid[kColourLdsVerts];
    __shared__ int nextCount, uniq, rem[2];
    __shared__ ushort4 tailLid[64];  // the last contacts, handed to one wave
    __shared__ unsigned long long tailBid[64];
    __shared__ int tailIdx[64];
    __shared__ int part[kSoloBlock];
    const int n = min(ctl->selfCount[kind], b.cap);
    const int tid = threadIdx.x, nt = blockDim.x;
    auto vatomic_min = [&](int v, int x) {
        __hip_atomic_fetch_min(&vloc[v], x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    };
    auto vload = [&](int v) { return __hip_atomic_load(&vloc[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP); };
    auto vstore = [&](int v, int x) {
        __hip_atomic_store(&vloc[v], x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    };
    // the distinct vertices, in claim order
    if (tid == 0) uniq = 0;
    for (int i = tid; i < 4 * n; i += nt) vatomic_min(corner(b.ids[i >> 2], i & 3), i);
    __syncthreads();
    for (int i = tid; i < 4 * n; i += nt) {
        const int v = corner(b.ids[i >> 2], i & 3);
        if (vload(v) == i) {
            const int slot = atomicAdd(&uniq, 1);
            b.vlist[slot] = v;
            vstore(v, kVlocSlot | slot);
        }
    }
    __syncthreads();
    const int m = uniq;
    if (m <= 65536) {
        for (int c = tid; c < n; c += nt) {
            const int4 q = b.ids[c];
            b.lid[c] = make_ushort4((unsigned short)(vload(q.x) & 0xffff), (unsigned short)(vload(q.y) & 0xffff),
                                    (unsigned short)(vload(q.z) & 0xffff), (unsigned short)(vload(q.w) & 0xffff));
        }
    }
    __syncthreads();
    for (int k = tid; k < m; k += nt) vstore(b.vlist[k], kVlocFree);
    if (tid == 0) ctl->selfUniq[kind] = m;
    int r = 0;
    if (n <= kColourPerThread * kSoloBlock && m <= kColourLdsVerts) {
        // this thread's contacts: local vertices, bid, colour (-1: none yet)
        ushort4 ml[kColourPerThread];
        unsigned long long mb[kColourPerThread];
        int mc[kColourPerThread];
#pragma unroll
        for (int k = 0; k < kColourPerThread; k++) {
            const int c = tid + k * nt;
            mc[k] = -2;
            if (c < n) {
                ml[k] = b.lid[c];
                mb[k] = bid_of(b.key[c]);
                mc[k] = -1;
            }
        }
        for (int k = tid; k < m; k += nt) lbid[k] = 0ull;
        if (tid == 0) rem[0] = rem[1] = 0;
        __syncthreads();
        for (;;) {
#pragma unroll
            for (int k = 0; k < kColourPerThread; k++)
                if (mc[k] == -1) {
                    atomicMax(&lbid[ml[k].x], mb[k]);
                    atomicMax(&lbid[ml[k].y], mb[k]);
                    atomicMax(&lbid[ml[k].z], mb[k]);
                    atomicMax(&lbid[ml[k].w], mb[k]);
                }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < kColourPerThread; k++)
                if (mc[k] == -1 && lbid[ml[k].x] == mb[k] && lbid[ml[k].y] == mb[k] && lbid[ml[k].z] == mb[k] &&
                    lbid[ml[k].w] == mb[k])
                    mc[k] = r;
            __syncthreads();
            // clear this round's bids, count the contacts left (two counters
            // in turn: the one read after this round is cleared in the next)
            int left = 0;
#pragma unroll
            for (int k = 0; k < kColourPerThread; k++)
                if (mc[k] == -1 || mc[k] == r) {
                    lbid[ml[k].x] = 0ull;
                    lbid[ml[k].y] = 0ull;
                    lbid[ml[k].z] = 0ull;
                    lbid[ml[k].w] = 0ull;
                    left += mc[k] == -1;
                }
            if (left) atomicAdd(&rem[r & 1], left);
            if (tid == 0) rem[(r + 1) & 1] = 0;
            __syncthreads();
            const int remaining = rem[r & 1];
            r++;
            if (remaining == 0) break;
            if (remaining <= 64) {
                // The last few contacts go to wave 0, which runs the same
                // rounds (same bids, so the same colours) with no workgroup
                // barrier: a wave's LDS operations take effect in order. The
                // counter rem[r & 1] was cleared before the last barrier.
#pragma unroll
                for (int k = 0; k < kColourPerThread; k++)
                    if (mc[k] == -1) {
                        const int slot = atomicAdd(&rem[r & 1], 1);
                        tailLid[slot] = ml[k];
                        tailBid[slot] = mb[k];
                        tailIdx[slot] = tid + k * nt;
                    }
                __syncthreads();
                if (tid < 64) {
                    const bool have = tid < remaining;
                    const ushort4 l = have ? tailLid[tid] : make_ushort4(0, 0, 0, 0);
                    const unsigned long long bid = have ? tailBid[tid] : 0ull;
                    int col = -1, rr = r;
                    constexpr int kWaitLds = 0xC07F;  // lgkmcnt(0)
                    while (__ballot(have && col < 0)) {
                        if (have && col < 0) {
                            atomicMax(&lbid[l.x], bid);
                            atomicMax(&lbid[l.y], bid);
                            atomicMax(&lbid[l.z], bid);
                            atomicMax(&lbid[l.w], bid);
                        }
                        __builtin_amdgcn_s_waitcnt(kWaitLds);
                        const bool open = have && col < 0;
                        if (open && lbid[l.x] == bid && lbid[l.y] == bid && lbid[l.z] == bid && lbid[l.w] == bid)
                            col = rr;
                        __builtin_amdgcn_s_waitcnt(kWaitLds);
                        if (open) {
                            lbid[l.x] = 0ull;
                            lbid[l.y] = 0ull;
                            lbid[l.z] = 0ull;
                            lbid[l.w] = 0ull;
                        }
                        __builtin_amdgcn_s_waitcnt(kWaitLds);
                        rr++;
                    }
                    if (have) b.colour[tailIdx[tid]] = col;
                    if (tid == 0) rem[0] = rr;
                }
                __syncthreads();
                r = rem[0];
                break;
            }
        }
#pragma unroll
        for (int k = 0; k < kColourPerThread; k++)
            if (mc[k] >= 0) b.colour[tid + k * nt] = mc[k];
        if (n == 0) r = 0;
        __syncthreads();
    } else {
        auto vreset = [&](int c) {
            const int4 q = b.ids[c];
            __hip_atomic_store(&vmax[q.x], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            __hip_atomic_store(&vmax[q.y], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            __hip_atomic_store(&vmax[q.z], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            __hip_atomic_store(&vmax[q.w], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        };
        auto vbid = [&](int v, unsigned long long bid) {
            __hip_atomic_fetch_max(&vmax[v], bid, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        };
        auto vget = [&](int v) {
            return __hip_atomic_load(&vmax[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        };
        // the uncoloured contacts, compacted after every round (two lists)
        int* cur = b.order;
        int* nxt = b.active;
        for (int c = tid; c < n; c += nt) {
            b.colour[c] = -1;
            cur[c] = c;
            vreset(c);
        }
        __syncthreads();
        int mm = n;
        while (mm > 0) {
            if (tid == 0) nextCount = 0;
            for (int i = tid; i < mm; i += nt) {
                const int c = cur[i];
                const int4 q = b.ids[c];
                const unsigned long long bid = bid_of(b.key[c]);
                vbid(q.x, bid);
                vbid(q.y, bid);
                vbid(q.z, bid);
                vbid(q.w, bid);
            }
            __syncthreads();
            for (int i = tid; i < mm; i += nt) {
                const int c = cur[i];
                const int4 q = b.ids[c];
                const unsigned long long bid = bid_of(b.key[c]);
                const bool win = vget(q.x) == bid && vget(q.y) == bid && vget(q.z) == bid && vget(q.w) == bid;
                if (win) b.colour[c] = r;
            }
            __syncthreads();
            for (int i = tid; i < mm; i += nt) {
                const int c = cur[i];
                vreset(c);
                if (b.colour[c] < 0) nxt[atomicAdd(&nextCount, 1)] = c;
            }
            __syncthreads();
            mm = nextCount;
            int* t = cur;
            cur = nxt;
            nxt = t;
            __syncthreads();
            r++;
        }
    }
    // r colours; counts, then offsets
    const int nc = r;
    for (int g = tid; g <= nc; g += nt) b.groupOff[g] = 0;
    __syncthreads();
    for (int c = tid; c < n; c += nt) atomicAdd(&b.groupOff[b.colour[c]], 1);
    __syncthreads();
    // exclusive scan of groupOff[0 .. nc]: thread t owns a contiguous chunk
    const int chunk = (nc + 1 + nt - 1) / nt;
    const int lo = min(tid * chunk, nc + 1), hi = min(lo + chunk, nc + 1);
    int sum = 0;
    for (int g = lo; g < hi; g++) sum += b.groupOff[g];
    part[tid] = sum;
    __syncthreads();
    if (tid == 0) {
        int run = 0;
        for (int t = 0; t < nt; t++) {
            const int k = part[t];
            part[t] = run;
            run += k;
        }
    }
    __syncthreads();
    int run = part[tid];
    for (int g = lo; g < hi; g++) {
        const int k = b.groupOff[g];
        b.groupOff[g] = run;
        b.active[g] = run;  // the scatter's cursor of colour g
        run += k;
    }
    if (tid == 0) ctl->selfColours[kind] = nc;
    __syncthreads();
    for (int c = tid; c < n; c += nt) b.order[atomicAdd(&b.active[b.colour[c]], 1)] = c;
}

// --------------------------------------------------------- projections ----

// A self contact's record as its projection reads it: vertices, weights and
// side, normal and lambda. Only the contact's own projection writes W and N,
// so the record can be read ahead of the contact's colour group.
struct SelfRec {
    int4 id;
    ushort4 lid;  // places in the kind's vertex list (LDS solve)
    float4 W, N;
};

__device__ __forceinline__ SelfRec self_rec(const SelfBuf& b, int c, bool local)
{
    SelfRec r{};
    if (c >= 0) {
        if (local)
            r.lid = b.lid[c];
        else
            r.id = b.ids[c];
        r.W = b.W[c];
        r.N = b.N[c];
    }
    return r;
}

// ClothCollisionHandler.cpp:364-436 (SolveSelfTrisPosition). at(j) is the
// position of the contact's vertex j (global memory or an LDS copy).
template <class At>
__device__ __forceinline__ float self_vt(At at, SelfBuf b, int c, const SelfRec& r, const SelfParams& sp, float h)
{
    const float4 P1 = at(0), P2 = at(1), P3 = at(2), P4 = at(3);
    f3 x1 = xyz(P1), x2 = xyz(P2), x3 = xyz(P3), x4 = xyz(P4);
    const float m1 = P1.w, m2 = P2.w, m3 = P3.w, m4 = P4.w;
    const float4 W = r.W, N = r.N;
    const f3 cn = xyz(N);
    float len0 = sp.thickness * sp.shrink;
    f3 p, co;
    closest_pt_point_triangle(x4, x1, x2, x3, p, co);
    // a collapsed triangle gives no closest point (the reference's
    // coordinates become NaN and spread to the whole cloth): skipped
    if (!isfinite(co.x + co.y + co.z)) return 0.f;
    const float w1 = co.x, w2 = co.y, w3 = co.z;
    const f3 delta = sub(x4, p);
    const f3 nc = cross(sub(x2, x1), sub(x3, x1));
    if (len(nc) == 0.f) return 0.f;
    f3 n = normalized(nc);
    const float side = signed_volume(x4, x1, x2, x3);
    if (W.w * side <= 0.f) len0 = sp.thickness * 0.1f;
    if (dot(n, cn) < 0.f) n = neg(n);
    const float len = dot(n, delta);
    float err = 0.f;
    float lambda = N.w;
    if (!(len > len0)) {
        const float im = w1 * w1 * m1 + w2 * w2 * m2 + w3 * w3 * m3 + m4;
        const float s = 1.f / im;
        const float e = len - len0;
        err = smax(0.f, -e);
        const float dLambda = 0.9f * s * e;
        const f3 disp = mul(dLambda, n);
        lambda += -dLambda / h;
        x1 = add(x1, mul(w1 * m1, disp));
        x2 = add(x2, mul(w2 * m2, disp));
        x3 = add(x3, mul(w3 * m3, disp));
        x4 = sub(x4, mul(m4, disp));
        at(0) = f4(x1, m1);
        at(1) = f4(x2, m2);
        at(2) = f4(x3, m3);
        at(3) = f4(x4, m4);
    }
    b.W[c] = make_float4(w1, w2, w3, W.w);
    b.N[c] = make_float4(N.x, N.y, N.z, lambda);
    return err;
}

// ClothCollisionHandler.cpp:439-541 (SolveSelfEdgesPosition)
template <class At>
__device__ __forceinline__ float self_ee(At at, SelfBuf b, int c, const SelfRec& r, const SelfParams& sp, float h)
{
    const float4 P1 = at(0), P2 = at(1), P3 = at(2), P4 = at(3);
    f3 x1 = xyz(P1), x2 = xyz(P2), x3 = xyz(P3), x4 = xyz(P4);
    const float m1 = P1.w, m2 = P2.w, m3 = P3.w, m4 = P4.w;
    const float4 W = r.W, N = r.N;
    float thick = sp.thickness * sp.shrink;
    const float side = signed_volume(x1, x2, x3, x4);
    f3 p, q;
    float t1, t2;
    closest_pt_segm_segm(x1, x2, x3, x4, t1, t2, p, q);
    const float w1 = 1.f - t1, w2 = 1.f - t2;
    const float omw1 = 1.f - w1, omw2 = 1.f - w2;
    const f3 delta = sub(q, p);
    // edges that touch exactly have no direction between them (the reference
    // normalises the zero vector into NaN, which spreads to the whole cloth):
    // the contact's normal is used instead
    f3 n = len(delta) == 0.f ? xyz(N) : normalized(delta);
    if (!isfinite(n.x + n.y + n.z)) return 0.f;
    if (dot(n, xyz(N)) < 0.0f) n = neg(n);
    if (side * W.w < 0.f) thick = sp.thickness * 0.1f;
    const float len = dot(n, delta);
    float err = 0.f;
    float lambda = N.w;
    if (!(len > thick)) {
        const float im = w1 * w1 * m1 + omw1 * omw1 * m2 + w2 * w2 * m3 + omw2 * omw2 * m4;
        const float s = 1.f / im;
        const float e = len - thick;
        err = smax(0.f, -e);
        const float dLambda = -0.9f * s * e;
        lambda += dLambda / h;
        const f3 disp = mul(dLambda, n);
        if (sp.eeWeights == 0) {
            // as the reference writes it: p1 moves by p2's barycentric weight
            // and p2 by p1's (likewise p3, p4)
            x1 = sub(x1, mul(omw1 * m1, disp));
            x2 = sub(x2, mul(w1 * m2, disp));
            x3 = add(x3, mul(omw2 * m3, disp));
            x4 = add(x4, mul(w2 * m4, disp));
        } else {
            x1 = sub(x1, mul(w1 * m1, disp));
            x2 = sub(x2, mul(omw1 * m2, disp));
            x3 = add(x3, mul(w2 * m3, disp));
            x4 = add(x4, mul(omw2 * m4, disp));
        }
        at(0) = f4(x1, m1);
        at(1) = f4(x2, m2);
        at(2) = f4(x3, m3);
        at(3) = f4(x4, m4);
    }
    b.W[c] = make_float4(w1, w2, 0.f, W.w);
    b.N[c] = make_float4(N.x, N.y, N.z, lambda);
    return err;
}

// One workgroup walks the colour groups in order; a group's contacts share
// no vertex and are projected in parallel. The group offsets sit in LDS and
// each thread reads its contacts ahead: the contact index two groups ahead,
// the record one group ahead. When the kind's distinct vertices fit
// (kSolveLdsVerts), they are staged in LDS for the whole pass and written
// back at its end, so a group waits on no memory at all; otherwise a group
// waits only for the positions the previous group wrote.
constexpr int kSelfGroupsLds = 4096;
constexpr int kSolveLdsVerts = 8192;

__global__ void __launch_bounds__(kSoloBlock) k_self_solve(float4* pos, SelfBuf b, int kind, SelfParams sp, float h,
                                                           SolverCtl* ctl)
{
    __shared__ int goff[kSelfGroupsLds + 1];
    __shared__ float4 sv[kSolveLdsVerts];
    if (ctl->done) return;
    const int nc = ctl->selfColours[kind];
    const int m = ctl->selfUniq[kind];
    const int tid = threadIdx.x;
    const bool inLds = nc <= kSelfGroupsLds;
    const bool local = m <= kSolveLdsVerts;
    if (inLds)
        for (int g = tid; g <= nc; g += kSoloBlock) goff[g] = b.groupOff[g];
    if (local)
        for (int k = tid; k < m; k += kSoloBlock) sv[k] = pos[b.vlist[k]];
    __syncthreads();
    auto project = [&](int c, const SelfRec& r) {
        if (local) {
            auto at = [&](int j) -> float4& {
                return sv[j == 0 ? r.lid.x : j == 1 ? r.lid.y : j == 2 ? r.lid.z : r.lid.w];
            };
            return kind == 0 ? self_vt(at, b, c, r, sp, h) : self_ee(at, b, c, r, sp, h);
        }
        auto at = [&](int j) -> float4& { return pos[j == 0 ? r.id.x : j == 1 ? r.id.y : j == 2 ? r.id.z : r.id.w]; };
        return kind == 0 ? self_vt(at, b, c, r, sp, h) : self_ee(at, b, c, r, sp, h);
    };
    // the pass over the groups, with the offsets read through off(g) (one
    // instance per source, so each reads its own memory directly)
    auto sweep = [&](auto off) {
        // this thread's first contact of group g, or -1
        auto first = [&](int g) {
            if (g >= nc) return -1;
            const int k = off(g) + tid;
            return k < off(g + 1) ? b.order[k] : -1;
        };
        float err = 0.f;
        int c1 = first(0), c2 = first(1);
        SelfRec r1 = self_rec(b, c1, local);
        for (int g = 0; g < nc; g++) {
            const int c3 = first(g + 2);
            const SelfRec r2 = self_rec(b, c2, local);
            if (c1 >= 0) {
                const float x = project(c1, r1);
                err = err < x ? x : err;
            }
            // a group wider than the workgroup: its further contacts in turn
            for (int k = off(g) + tid + kSoloBlock; k < off(g + 1); k += kSoloBlock) {
                const int c = b.order[k];
                const float x = project(c, self_rec(b, c, local));
                err = err < x ? x : err;
            }
            // staged vertices pass from group to group through LDS alone (a
            // contact's own record is read by nothing else in the pass)
            if (local)
                lds_barrier();
            else
                __syncthreads();
            c1 = c2;
            r1 = r2;
            c2 = c3;
        }
        return err;
    };
    const float err = inLds ? sweep([&](int g) { return goff[g]; }) : sweep([&](int g) { return b.groupOff[g]; });
    if (local) {
        __syncthreads();
        for (int k = tid; k < m; k += kSoloBlock) pos[b.vlist[k]] = sv[k];
    }
    wave_max_atomic(kind == 0 ? &ctl->sct : &ctl->sce, err);
}

static inline int blocks(int n) { return (n + kBlock - 1) / kBlock; }

void launch_self_boxes(hipStream_t s, const float4* pos, const float4* prev, const int4* cTris, int nt,
                       const int2* cEdges, int ne, float tol, float4* triLo, float4* triHi, float4* edgeLo,
                       float4* edgeHi, int nv, float4* vtxLo, float4* vtxHi)
{
    const int n = nt + ne + (vtxLo ? nv : 0);
    if (n <= 0) return;
    hipLaunchKernelGGL(k_self_boxes, dim3(blocks(n)), dim3(kBlock), 0, s, pos, prev, cTris, nt, cEdges, ne, tol,
                       triLo, triHi, edgeLo, edgeHi, nv, vtxLo, vtxHi);
}

void launch_self_detect(hipStream_t s, const float4* pos, const float4* prev, const int4* cTris, const int2* cEdges,
                        int nv, int ne, BvhView triBvh, BvhView edgeBvh, const int* vertexOrder, SelfParams sp,
                        SelfCandidates cand, SelfBuf vt, SelfBuf ee, SolverCtl* ctl)
{
    const int vb = blocks(nv), eb = blocks(ne);
    if (vb + eb == 0) return;
    hipLaunchKernelGGL(k_self_candidates, dim3(vb + eb), dim3(kBlock), 0, s, pos, prev, cTris, cEdges, nv, ne, vb,
                       triBvh, edgeBvh, vertexOrder, sp.tol, cand, ctl);
    // enough workgroups to fill the chip several times over; they stride
    hipLaunchKernelGGL(k_self_ccd, dim3(kCcdBlocks), dim3(kBlock), 0, s, pos, prev, cTris, cEdges, sp, cand, vt, ee,
                       ctl);
}

void launch_self_colour(hipStream_t s, SelfBuf vt, SelfBuf ee, unsigned long long* vmax, int* vloc, int nv,
                        SolverCtl* ctl)
{
    hipLaunchKernelGGL(k_self_colour, dim3(2), dim3(kSoloBlock), 0, s, vt, ee, vmax, vloc, nv, ctl);
}

void launch_self_solve(hipStream_t s, float4* pos, SelfBuf b, int kind, SelfParams sp, float h, SolverCtl* ctl)
{
    hipLaunchKernelGGL(k_self_solve, dim3(1), dim3(kSoloBlock), 0, s, pos, b, kind, sp, h, ctl);
}

}  // namespace sc
