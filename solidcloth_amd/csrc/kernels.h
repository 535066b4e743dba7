// Device data layout and kernel launchers for the cloth substep on gfx950.
//
// HBM layout (all FP32, one array per quantity):
//   pos  float4[V]  x, y, z, inverse mass
//   prev float4[V]  x, y, z, 0
//   vel  float4[V]  x, y, z, 0
//   triangles, in Gauss-Seidel colour order, as four parallel streams so a
//   wavefront reads each with one coalesced 16/8-byte load per lane:
//     tidx int4[T] (i1, i2, i3, original index), tA float4[T] (du1, du2, dv1,
//     dv2), tB float4[T] (invDet, lu2, lv2, area), tC float2[T] (lu0, lv0)
//   bends, in colour order: bidx int4[B] (i1, i2, i3, i4)
//   contacts: per-vertex slots, slot-major ([slot * V + v]) so that slot s of
//     consecutive vertices is contiguous: cP float4 (point, lambda),
//     cN float4 (normal, lambdaF), cSrc int (collidable), cCount int[V]
//   grid SDF nodes: float4 (grad x, grad y, grad z, value), x fastest
#pragma once

#include <hip/hip_runtime.h>

#include "lbvh.h"
#include "scmath.h"

namespace sc {

enum { CF_VERTICES = 1, CF_TRIANGLES = 2, CF_SELF = 4, CF_WALLS = 8, CF_EDGES = 16 };
enum { CT_INVALID = 0, CT_WALLS = 1, CT_SLAB = 2, CT_SPHERE = 3, CT_MESH = 4, CT_BOX = 5, CT_CAPSULE = 6, CT_SDF = 7 };

// Per-substep solver control block, written by the kernels. Errors are
// non-negative floats stored by their bit pattern so atomicMax orders them.
// The packed sweeps' error maxima go to kErrParts words per error, one per
// 128-byte line, chosen by workgroup, so that a launch's waves do not queue
// on one address (a same-address atomic costs the launch about 2.6 ns per
// wave); a reader takes the maximum of the parts and the plain field.
constexpr int kErrParts = 64, kErrStride = 32;

struct SolverCtl {
    unsigned done;
    int k;
    unsigned cb;  // bend error of the last bend pass
    unsigned cv;  // vertex-contact error of the current iteration
    int iterations;
    float bendError;
    float contactError;
    unsigned ce;     // edge-contact error of the current iteration
    unsigned ct;     // triangle-contact error of the current iteration
    unsigned maxTv;  // largest |pos - prev| after prediction (mesh contacts)
    unsigned sct;    // self vertex-triangle error of the current iteration
    unsigned sce;    // self edge-edge error of the current iteration
    int selfCount[2];    // self contacts found this substep (vertex-triangle, edge-edge)
    int selfColours[2];  // colour groups of each kind
    int selfOverflow;    // contacts or candidate pairs beyond capacity
    unsigned cl;         // link error of the current iteration (link model)
    int candSlots[2];    // candidate pool slots reserved this substep (vertex, edge; may pass the capacity)
    // largest selfCount[0..1] and candSlots[0..1] of the step's earlier
    // substeps (the host clears them before a step and folds in the last)
    int selfPeak[4];
    int selfUniq[2];  // distinct vertices of each kind's contacts (this substep)
    unsigned errPart[2][kErrParts * kErrStride];  // [0] bend, [1] vertex-contact error parts
};

enum { MAX_COLLIDABLES = 8 };

// A static triangle-mesh obstacle on the device: vertices, triangles (i1, i2,
// i3, 0), angle-weighted vertex normals, edges (i1, i2) with their
// pseudo-normals, and one LBVH each over triangles, vertices and edges.
// `scale` is the largest absolute vertex coordinate (pruning slack).
struct MeshGeo {
    const float4* v;
    const int4* tri;
    const float4* vn;
    const int2* edge;
    const float4* en;
    BvhView triBvh, vtxBvh, edgeBvh;
    float scale;
    int nv, nt, ne;
};

struct CollDev {
    int type;
    int pad;
    float ax, ay, az;  // walls: box min; sphere: centre; sdf: box min; capsule: axis end P
    float bx, by, bz;  // walls: box max; sdf: cell steps; capsule: axis end Q
    float r;           // sphere / capsule radius
    const float4* nodes;
    const MeshGeo* mesh;  // mesh obstacle (device copy)
    int nc0, nc1, nc2;
    int nv0, nv1;
};

struct CollSet {
    int n;
    int maxSlots;
    CollDev c[MAX_COLLIDABLES];
};

struct ContactBuf {
    int* count;
    float4* P;
    float4* N;
    int* src;   // producing collidable's world index, per slot
    int* feat;  // obstacle triangle of a mesh contact, -1 otherwise, per slot
    const MeshGeo* geo;  // per collidable index; null when the world has no mesh
    int nv;
    int slots;
};

// Obstacle-edge contacts per cloth edge and obstacle-vertex contacts per cloth
// triangle (Physics/ClothTypes.h:110-120), slot-major like ContactBuf. Cloth
// edges (triangles) carrying a contact are listed per colour group: group g
// holds up to groupBase[g+1] - groupBase[g] entries, groupCount[g *
// kGroupStride] used (one 128-byte line per counter, so that appends to
// different groups do not queue on one line); the groups are processed in
// order, their members in parallel.
constexpr int kGroupStride = 32;
struct FeatureContactBuf {
    int* count;        // per cloth edge / triangle
    float4* P;         // point, lambda
    float4* N;         // normal, lambdaF
    float4* W;         // w1, w2 (, w3)
    int2* F;           // obstacle feature, collidable
    const int* group;  // colour group of each cloth edge / triangle
    const int* groupBase;
    int* groupCount;
    int* list;
    int n;             // cloth edges / triangles
    int slots;
    int groups;
};

struct TriStreams {
    const int4* idx;
    const float4* a;
    const float4* b;
    const float2* c;
};

// Slab schedule on the device (slabs numbered phase-major): per slab its
// first vertex in `verts`, vertex count, first colour in `colOff` and colour
// count; colour k of a slab spans constraint positions [colOff[k], colOff[k+1]).
// Constraint vertices are given as slab-local indices (ushort4).
struct SlabSet {
    const int4* info;   // per slab: first vertex, vertex count, first colour, colour count
    const int* colOff;  // colour offsets (schedule positions) of all slabs' colours
    const int* first;   // per slab: schedule position of its first constraint
    const int* verts;
    int base;   // first slab of the launched phase
    int count;  // constraints in the schedule (bound of prefetch indices)
};

enum { kSlabBlock = 256, kMaxSlabColours = 63, kSlabBatch = 8 };
// Flags in the high bits of a slab's vertex ids: the vertex's owner (the
// FEM slab that applies staged contacts, or a tile's own vertex) and, in
// bend slabs, the vertex's last slab (which may apply the iteration's
// contacts); ids are below 2^30.
constexpr unsigned kSlabOwner = 0x80000000u, kSlabContacts = 0x40000000u;
constexpr int kSlabIdMask = 0x3fffffff;

// Vertex contacts applied by the slab kernels: enabled 1 (FEM slabs), while
// staging vertices (the previous iteration's contacts, by each vertex's owner
// slab); 2 (FEM tiles), after the last colour (this iteration's contacts of
// the tile's own vertices); 3 (bend slabs), after the last colour (this
// iteration's contacts of the vertices whose last bend slab it is).
struct SlabContacts {
    ContactBuf cb;
    const float4* prev;
    float4* vel;
    float thickness, mu, invH;
    int enabled;
    // error maxima this launch adds to the solver's accumulators: bit 0 the
    // vertex-contact error, bit 1 the bend error (a fixed iteration count
    // reports only the last passes', so earlier passes skip the atomics)
    int report = 3;
    // FEM tiles applying the last iteration's contacts: each tile also sets
    // its own vertices' velocities, v = (pos - prev) / h (ClothPBD.cpp:314-321),
    // in place of the velocity launch
    int velocity = 0;
};

// Packed sweep schedules (one phase of a slab or tile schedule), laid out so
// that every load of a sweep launch is addressed by the workgroup and thread
// alone: each slab has kv * kSlabBlock vertex slots and nc colours of `width`
// constraint slots. Vertex slots hold a global id with flags (kSlabOwner,
// kSlabContacts as in the slab lists, kPackPad for padding, which repeats a
// vertex of the slab); constraint slots hold slab-local vertex ids, and
// padding slots name the spare LDS entry kv * kSlabBlock.
constexpr unsigned kPackPad = 0x20000000u;
constexpr int kPackIdMask = 0x1fffffff;
// FEM tiles: colour c of tile t has slots [(t nc + c) kSlabBlock, +kSlabBlock);
// s0 = (local 1 | local 2 << 10 | local 3 << 20, invDet bits, lu0 bits, lv0
// bits), s1 = (du1, du2, dv1, dv2): 32 bytes a slot. A tile has fewer than
// kv * kSlabBlock vertices (at most 1023: ten-bit local ids), and its last
// vertex slot, always padding, is the spare entry. Owned vertices come first
// in a tile's list, at most one per thread.
struct PackedFem {
    const int* verts;
    const uint4* s0;
    const float4* s1;
    int nc;  // colours in use (the rest of the kernel's colours are padding)
};
// Bend slabs: colour c of slab s has slots [(s nc + c) width, +width), each
// the bend's four slab-local ids.
struct PackedBend {
    const int* verts;
    const uint2* ids;
    int nc, width;
};
// Largest colour count of a packed bend phase, and the vertex slots per
// thread the packed kernels are built for: FEM tiles and bend slabs that
// apply vertex contacts up to kPackMaxKv, other bend slabs up to
// kPackMaxBendKv.
constexpr int kPackMaxBendColours = 16;
constexpr int kPackMaxKv = 4;
constexpr int kPackMaxBendKv = 6;

// Self contacts of one kind (vertex i4 against triangle i1 i2 i3, or edge i1 i2
// against edge i3 i4), in detection order; `order` lists them colour group by
// colour group (group g spans order[groupOff[g] .. groupOff[g+1])), members of
// a group sharing no vertex. key = (triangle << 32 | vertex) or
// (edge1 << 32 | edge2).
struct SelfBuf {
    int4* ids;
    float4* W;  // w1, w2, w3, side
    float4* N;  // normal, lambda
    unsigned long long* key;
    int* colour;
    int* order;
    int* active;  // scratch list for the colouring rounds
    int* groupOff;
    // the kind's distinct contact vertices (selfUniq of them, nv capacity)
    // and each contact's vertices as indices into that list (valid when there
    // are at most 65536 distinct vertices)
    int* vlist;
    ushort4* lid;
    int cap;
    int maxColours;
};

struct SelfParams {
    float tol;       // swept-box extrusion: thickness + 0.5 tolerance
    float thickness;
    float shrink;
    double volume, bisect;  // CCD thresholds (FP64 cubic)
    float vtDistance;
    int eeWeights;  // 0: edge-edge ends moved as ClothCollisionHandler.cpp:531-534 writes it; 1: by their own weights
};

// One cloth of a batch: its own device buffers (the patch's), its greedy
// colour schedule and its substep parameters. The batch kernel runs a whole
// substep of one cloth in one workgroup.
struct ClothDesc {
    float4 *pos, *prev, *vel;
    TriStreams tr;      // triangles in colour order (global vertex ids)
    const int* triCol;  // colour c spans [triCol[c], triCol[c+1])
    const int4* bidx;   // bends in colour order
    const int* bendCol;
    ContactBuf cb;
    const CollSet* cs;  // the cloth's collidables (device copy)
    SolverCtl* ctl;
    int nv, nt, nb, triColours, bendColours;
    float h, h2, epsUV, sStiff, drag, thickness, tolerance, mu, invH, selfThreshold;
    f3 gdv;
    int iterations, fixedIterations, bendFreq, dihedral, collFlags, anyContacts;
    int pad[2];
};

// Launchers (kernels.hip). All are asynchronous on the given stream.
void launch_predict(hipStream_t s, float4* pos, float4* prev, float4* vel, int n, float h, f3 gdv, float drag,
                    SolverCtl* ctl);
void launch_detect(hipStream_t s, const float4* pos, const float4* prev, int n, const CollSet& cs, ContactBuf cb,
                   float thickness, float tolerance, int collFlags, const SolverCtl* ctl);
void launch_fem(hipStream_t s, float4* pos, TriStreams tr, int start, int count, float h2, float epsUV,
                const SolverCtl* ctl);
void launch_bend(hipStream_t s, float4* pos, const int4* bidx, int start, int count, float stiff, SolverCtl* ctl);
void launch_fem_slab(hipStream_t s, float4* pos, TriStreams tr, const ushort4* lidx, SlabSet ss, int numSlabs,
                     int maxVerts, float h2, float epsUV, SolverCtl* ctl, const SlabContacts& kc, bool tiles);
void launch_predict_detect(hipStream_t s, float4* pos, float4* prev, float4* vel, int n, float h, f3 gdv, float drag,
                           SolverCtl* ctl, const CollSet& cs, ContactBuf cb, float thickness, float tolerance,
                           int collFlags);
void launch_contacts_velocity(hipStream_t s, float4* pos, const float4* prev, float4* vel, ContactBuf cb,
                              float thickness, float mu, float invH, SolverCtl* ctl);
void launch_bend_slab(hipStream_t s, float4* pos, const ushort4* lidx, SlabSet ss, int numSlabs, int maxVerts,
                      float stiff, SolverCtl* ctl, const SlabContacts& kc);
// Packed sweeps: numSlabs workgroups; nc (FEM: 8 or 12, colours padded to it)
// and kv (1 .. kPackMaxKv) as packed. kc.enabled 2 (FEM) / 3 (bends) applies
// vertex contacts as the slab kernels do.
void launch_fem_packed(hipStream_t s, float4* pos, PackedFem pk, int numTiles, int nc, int kv, float h2, float epsUV,
                       SolverCtl* ctl, const SlabContacts& kc);
void launch_bend_packed(hipStream_t s, float4* pos, PackedBend pk, int numSlabs, int kv, float stiff, SolverCtl* ctl,
                        const SlabContacts& kc);
void launch_contacts(hipStream_t s, float4* pos, const float4* prev, float4* vel, ContactBuf cb, float thickness,
                     float mu, float invH, SolverCtl* ctl);
void launch_iter_end(hipStream_t s, SolverCtl* ctl, int maxIter, int fixedIter, int bendFreq, int dihedral,
                     float selfThreshold);
// Clears error accumulators before the last pass of a fixed iteration count
// (mask 1: links, contacts, self; 2: bends).
void launch_reset_errors(hipStream_t s, SolverCtl* ctl, int mask);
// Link model (ClothPBD.cpp:9-50): links (i1, i2) with (rest length,
// stiffness) in schedule order; k = h^2 area unit.
void launch_links(hipStream_t s, float4* pos, const int2* lidx, const float2* lrest, int start, int count, float k,
                  SolverCtl* ctl);
void launch_link_slab(hipStream_t s, float4* pos, const ushort4* lidx, const float2* lrest, SlabSet ss, int numSlabs,
                      int maxVerts, float k, SolverCtl* ctl);
void launch_velocity(hipStream_t s, float4* pos, const float4* prev, float4* vel, int n, float invH, SolverCtl* ctl);
void launch_copy(hipStream_t s, const float4* src, float4* dst, long n);
// The mouse spring (Physics/ClothTypes.h:85-91): particles i1..i3 with
// barycentric weights a1..a3, pulled towards (px, py, pz) with force
// stiffness (|p - point| - len). Kept in device memory so that a captured
// substep follows its moves.
struct MouseSpringDev {
    int i1, i2, i3;
    float a1, a2, a3;
    float px, py, pz;
    float len, stiffness;
};
// Physics/ClothPBD.cpp:267-282, one thread; with resetTravel it also clears
// ctl->maxTv for launch_max_travel
void launch_mouse_spring(hipStream_t s, float4* pos, float4* vel, const MouseSpringDev* ms, float h, SolverCtl* ctl,
                         int resetTravel);
// ctl->maxTv = max over all particles of |pos - prev| (ClothCollision.cpp:95-104)
void launch_max_travel(hipStream_t s, const float4* pos, const float4* prev, int n, SolverCtl* ctl);
// Links in the reference's order for its Jacobi and conjugate-residual
// solvers: ends lk, (rest length, stiffness) lr, per-link displacement and
// search direction scratch, each particle's (link * 2 + end) entries in
// ascending link order (CSR off / inc).
struct JacobiLinks {
    const int2* lk;
    const float2* lr;
    float4* disp;
    float* d;
    const int* off;
    const int* inc;
    int nl, nv;
};
// one SolveLinksJacobi pass (ClothPBD.cpp:616-651), k = h^2 area unit
void launch_jacobi_pass(hipStream_t s, float4* pos, const JacobiLinks& J, float k, float omega);
// SolveCR (ClothPBD.cpp:384-478 as compiled): d = 0 once per solve, then one
// pass per iteration with beta = (iter / (N - 1))^0.6 and alpha
void launch_cr_reset(hipStream_t s, const JacobiLinks& J);
void launch_cr_pass(hipStream_t s, float4* pos, const JacobiLinks& J, float unit, float beta, float alpha);
// Render mesh of the cloth (Physics/ClothPatch.cpp:266-273, Geometry/Mesh.cpp:10-63):
// outV = pos + offset and outN = angle-weighted unit vertex normals (nv x 3
// floats each), triN = unit face normals (nt x 3); idx = 3 nt render indices
// into the particles; vt lists each vertex's (triangle * 4 + corner) in
// ascending order, vtOff its CSR offsets; triA / triB are nt-sized scratch.
// map (null: identity) gives each render vertex's particle.
void launch_render_mesh(hipStream_t s, const float4* pos, const int* map, int nv, const unsigned* idx, int nt,
                        f3 offset, const int* vtOff, const int* vt, float4* triA, float2* triB, float* outV,
                        float* outN, float* triN);

// Batch: one workgroup per cloth (cloths in launch order), one launch per
// substep. blockSize 256, 512 or 1024; cloths with at most stageVerts
// vertices keep their positions in LDS during the iterations.
// wgClock (optional, 3 words per cloth): each workgroup's start and end on the
// device's constant-rate clock and the unit it ran on (XCC id << 32 | HW_ID).
void launch_batch_substep(hipStream_t s, const ClothDesc* descs, int numCloths, int blockSize, int stageVerts,
                          unsigned long long* wgClock = nullptr);
// Allows the batch kernel `bytes` of dynamic LDS (called before capture).
void batch_allow_lds(int bytes);

// Mesh obstacles (mesh.hip). Prediction that also records max |pos - prev|
// and clears the feature-contact groups; detection of obstacle-edge and
// obstacle-vertex contacts; their Gauss-Seidel passes.
void launch_predict_mesh(hipStream_t s, float4* pos, float4* prev, float4* vel, int n, float h, f3 gdv, float drag,
                         SolverCtl* ctl, FeatureContactBuf ec, FeatureContactBuf tc);
void launch_detect_features(hipStream_t s, const float4* pos, const float4* prev, const int2* cEdges,
                            const int4* cTris, const CollSet& cs, FeatureContactBuf ec, FeatureContactBuf tc,
                            float thickness, float tolerance, int collFlags, const SolverCtl* ctl);
// SDF node values of an (nx, ny, nz)-cell grid from the nearest triangle of a
// mesh obstacle (SDF::CreateFromMesh); (nx + 1) (ny + 1) (nz + 1) values,
// x fastest, at most 2^31 - 1.
void launch_sdf_from_mesh(hipStream_t s, const MeshGeo& g, int nx, int ny, int nz, f3 mn, f3 steps, float* out);

// One launch per colour group, in group order; groupBase is the host copy of
// the groups' offsets (each group's capacity sizes its grid).
void launch_edge_contacts(hipStream_t s, float4* pos, const float4* prev, const int2* cEdges, FeatureContactBuf ec,
                          const int* groupBase, const MeshGeo* geo, float thickness, float mu, SolverCtl* ctl);
void launch_tri_contacts(hipStream_t s, float4* pos, const float4* prev, const int4* cTris, FeatureContactBuf tc,
                         const int* groupBase, const MeshGeo* geo, float thickness, float mu, SolverCtl* ctl);

// Self collision (self.hip): swept boxes of triangles and edges, continuous
// vertex-triangle and edge-edge detection over the LBVHs, colouring of the
// contacts, and their Gauss-Seidel passes.
// (with vtxLo non-null also the vertices' swept boxes)
void launch_self_boxes(hipStream_t s, const float4* pos, const float4* prev, const int4* cTris, int nt,
                       const int2* cEdges, int ne, float tol, float4* triLo, float4* triHi, float4* edgeLo,
                       float4* edgeHi, int nv, float4* vtxLo, float4* vtxHi);
// Candidate pairs (triangle, vertex) and (edge1, edge2) in one flat pool per
// kind, reserved in chunks of kCandChunk slots; fill[c] is the number of
// pairs in chunk c (a chunk's tail may be unused). Capacities are multiples
// of kCandChunk. Then the continuous tests over them.
constexpr int kCandChunk = 1024;
struct SelfCandidates {
    int2* vt;
    int2* ee;
    int* fillVT;
    int* fillEE;
    int capVT, capEE;
};
// The triangle and edge trees carry each leaf's vertices as its record's
// payload (Lbvh::set_payload: the triangles' int4, the edges' (x, y, -1, -1)),
// so a walk tests shared vertices without a dependent load.
// vertexOrder: the order in which vertices query the triangle tree (may be null)
void launch_self_detect(hipStream_t s, const float4* pos, const float4* prev, const int4* cTris, const int2* cEdges,
                        int nv, int ne, BvhView triBvh, BvhView edgeBvh, const int* vertexOrder, SelfParams sp,
                        SelfCandidates cand, SelfBuf vt, SelfBuf ee, SolverCtl* ctl);
// vloc: one int per cloth vertex, every byte 0x7f between calls (the
// colouring claims and releases entries while it lists a kind's vertices)
// Colours both kinds' contacts at once (one workgroup each); vmax and vloc
// hold 2 nv entries, the first nv for the vertex-triangle kind.
void launch_self_colour(hipStream_t s, SelfBuf vt, SelfBuf ee, unsigned long long* vmax, int* vloc, int nv,
                        SolverCtl* ctl);
void launch_self_solve(hipStream_t s, float4* pos, SelfBuf b, int kind, SelfParams sp, float h, SolverCtl* ctl);

}  // namespace sc
