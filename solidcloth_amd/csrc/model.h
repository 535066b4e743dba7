// Host-side objects behind the C ABI: collision world and collidables, grid
// SDF, indexed mesh, cloth model and cloth patch. Names and defaults mirror
// the reference's classes (Physics/Common.h, Physics/CollisionWorld.h,
// Physics/ClothModel.h, Physics/ClothPatch.h); the solver state lives in HBM
// and is owned by GpuSolver.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../../include/solidcloth.h"
#include "kernels.h"
#include "lbvh.h"
#include "schedule.h"
#include "warmstart.h"

namespace sc {

struct Collidable {
    int type = CT_INVALID;
    scQuaternion rot{1.f, {0.f, 0.f, 0.f}};
    f3 center{0.f, 0.f, 0.f};
    virtual ~Collidable() {}
};

struct Walls : Collidable {
    scAabb3 box;
};

struct Sphere : Collidable {
    float radius = 0.f;
};

struct Capsule : Collidable {
    float r = 0.f, hh = 0.f;
};
// The ends of a capsule's axis, P = centre + rot (0, hh, 0), Q = centre - it.
void capsule_axis(const Capsule& k, f3& P, f3& Q);

// Node-sampled grid SDF (Geometry/SDF.h:13-251)
struct SDF {
    int nc[3] = {0, 0, 0};
    int nv[3] = {0, 0, 0};
    scAabb3 box;
    f3 steps{0.f, 0.f, 0.f};
    std::vector<float> vals;
    std::vector<f3> grads;
    float4* d_nodes = nullptr;  // (grad, value) per node
    ~SDF();
    void create(int nx, int ny, int nz, const scAabb3& b);
    void compute_gradient();
    void upload();
};

struct CollisionSDF : Collidable {
    SDF* sdf = nullptr;
};

struct MeshEdge {
    uint32_t i1, i2;
    int t1 = -1, t2 = -1;
};

// Device copy of a static obstacle mesh with its three LBVHs.
struct MeshDevice {
    float4 *v = nullptr, *vn = nullptr, *en = nullptr;
    int4* tri = nullptr;
    int2* edge = nullptr;
    float4 *triLo = nullptr, *triHi = nullptr, *vtxLo = nullptr, *edgeLo = nullptr, *edgeHi = nullptr;
    Lbvh triB, vtxB, edgeB;
    MeshGeo geo{};
    ~MeshDevice();
};

// Indexed triangle mesh (Geometry/Mesh.h:12-264)
struct Mesh {
    std::vector<scVector3> vertices;
    std::vector<unsigned> indices;
    std::vector<scVector3> normals;
    std::vector<scVector3> triNormals;   // unit face normals (Mesh.cpp:24-27)
    std::vector<scVector3> edgeNormals;  // edge pseudo-normals (Mesh.cpp:105, 125-126)
    std::vector<scVector2> uvs;
    std::vector<scVector3> colors;  // per vertex; a mesh-based patch pins vertices whose red is non-zero
    std::vector<MeshEdge> edges;
    std::unique_ptr<MeshDevice> dev;
    void clear();
    void construct_edges();
    void compute_normals();
    void compute_edge_normals();
    // Uploads the mesh and builds its LBVHs on first use (synchronous).
    const MeshGeo& device_geo(hipStream_t s);
};

struct CollisionMesh : Collidable {
    Mesh* mesh = nullptr;
};

struct CollisionWorld {
    std::vector<std::shared_ptr<Collidable>> colls;
};

struct Triangle {
    uint32_t i1, i2, i3;
    float du1, du2, dv1, dv2, det, invDet, lu0, lv0, su, sv, dotuv, area, lu2, lv2;
};

struct Bend {
    uint32_t i1, i2, i3, i4;
};

struct Link {
    uint32_t i1, i2;
    float len, stiffness;
};

struct Quad {
    uint32_t i1, i2, i3, i4;
};

struct ContactOut {
    int coll, vertex, slot, feature;
    f3 point, normal;
};

// An obstacle-edge contact of a cloth edge (ids = edge, feature, collidable)
// or an obstacle-vertex contact of a cloth triangle; w = (w1, w2[, w3]).
struct FeatureContactOut {
    int id, feature, coll;
    f3 point, normal;
    float w[3];
    float lambda, lambdaF;
};

class ClothPatch;

// The learned warm-start network (12-64-64-3 MLP) and its optimiser state on
// the device. Shared by the models it is attached to, which must step one
// after another (the training kernels use one set of gradient buffers).
class WarmStart {
public:
    explicit WarmStart(unsigned seed);
    ~WarmStart();
    static std::vector<float> initial_params(unsigned seed);
    std::vector<float> params() const;
    void set_params(const std::vector<float>& p);
    int steps() const;
    float last_loss() const;  // loss sum of the last gradient (times the weight normaliser)
    MlpWeights weights() const { return MlpWeights{d_master, d_packed}; }
    void enqueue_gradient(hipStream_t s, const uint16_t* X, const float4* T, int n, float invCount);
    void enqueue_train(hipStream_t s, const uint16_t* X, const float4* T, int n, float invCount);
    void forward_host(const float* X, int n, float* Y);
    // device features in, device outputs out, staged in static scratch (no allocation)
    void forward_device_noalloc(const float* X, int n, float* Y, hipStream_t s)
    {
        launch_mlp_forward_noalloc(s, weights(), X, n, Y);
    }
    float gradient_host(const float* X, const float* T, const float* w, int n, float* grad);
    float train_host(const float* X, const float* T, const float* w, int n, int count);
    void set_lr(float lr);
    float lr() const { return mLr; }

private:
    float mLr = 1e-3f;
    float* d_lr = nullptr;
    hipStream_t stream = nullptr;
    float *d_master = nullptr, *d_m = nullptr, *d_v = nullptr, *d_grad = nullptr, *d_partial = nullptr;
    uint16_t* d_packed = nullptr;
    int* d_step = nullptr;
    // staging for the host-array entry points
    uint16_t* d_X = nullptr;
    float* d_Xf = nullptr;
    float4* d_T = nullptr;
    float* d_Y = nullptr;
    int cap = 0;
    void stage(const float* X, const float* T, const float* w, int n, float& invCount);
};

// Sweep schedules: 0 = one launch per colour, 1 = slab phases with LDS
// staging, 2 = slab phases of strip interiors and seams, 3 = FEM triangles on
// tiles with halos (one launch per pass), bends and links as in 1
// SWEEP_PACKED: the schedules of SWEEP_TILES, laid out for the preloading
// sweep kernels (k_fem_packed, k_bend_packed) where they fit their limits
enum { SWEEP_COLOURS = 0, SWEEP_SLABS = 1, SWEEP_SEAMS = 2, SWEEP_TILES = 3, SWEEP_PACKED = 4 };
// Physics/ClothModel.h:23-30 (SolverType); Step runs SolveGS for
// Gauss-Seidel, SolveJacobi for Jacobi and SolveCR for every other value
enum { SOLVER_JACOBI = 0, SOLVER_GAUSS_SEIDEL = 1, SOLVER_JACOBI_CR = 2, SOLVER_IMPLICIT = 3, SOLVER_EXPLICIT = 4 };

// Device copy of one constraint schedule's slab structure.
struct DeviceSlabs {
    ushort4* lidx = nullptr;
    int4* info = nullptr;
    int* colOff = nullptr;
    int* first = nullptr;
    int* verts = nullptr;
};

// Device copy of a schedule packed for the preloading sweep kernels
// (PackedFem for FEM tiles, PackedBend for bend slabs); no phases when the
// schedule is not packed.
struct DevicePacked {
    int* verts = nullptr;
    uint4* s0 = nullptr;
    float4* s1 = nullptr;
    uint2* ids = nullptr;
    struct Phase {
        int slabs = 0, kv = 0, nc = 0, ncKernel = 0, width = 0;
        size_t vertOff = 0, slotOff = 0;
    };
    std::vector<Phase> phases;
};

struct KernelTimer {
    bool enabled = false;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
    std::vector<hipEvent_t> pool;
    int launches[5] = {0, 0, 0, 0, 0};
    double totalMs[5] = {0, 0, 0, 0, 0};
    hipEvent_t get();
    void harvest();
    void reset();
    ~KernelTimer();
};

class GpuSolver {
public:
    ~GpuSolver();
    void release();
    hipStream_t stream = nullptr;
    int nv = 0, nt = 0, nb = 0;
    float4 *d_pos = nullptr, *d_prev = nullptr, *d_vel = nullptr;
    int4* d_tidx = nullptr;
    float4 *d_tA = nullptr, *d_tB = nullptr;
    float2* d_tC = nullptr;
    // tile schedule: the triangles' rest data once per tile instance
    float4 *d_ttA = nullptr, *d_ttB = nullptr;
    float2* d_ttC = nullptr;
    int4* d_bidx = nullptr;
    // links (non-FEM model) in schedule order: vertex pair, (rest length, stiffness)
    int2* d_lk = nullptr;
    float2* d_lr = nullptr;
    int nl = 0;
    int *d_triColOff = nullptr, *d_bendColOff = nullptr;  // schedule phase offsets (colours in colour mode)
    int* d_cCount = nullptr;
    float4 *d_cP = nullptr, *d_cN = nullptr;
    int* d_cSrc = nullptr;
    int* d_cFeat = nullptr;
    MeshGeo* d_geo = nullptr;  // MAX_COLLIDABLES entries, by world index
    std::vector<MeshGeo> geoHost;
    // cloth edges (i1, i2) and triangles (i1, i2, i3, 0) in index order
    int2* d_cEdges = nullptr;
    int4* d_cTris = nullptr;
    int ne = 0;
    // obstacle-edge contacts per cloth edge, obstacle-vertex contacts per cloth triangle
    struct Features {
        int* count = nullptr;
        float4 *P = nullptr, *N = nullptr, *W = nullptr;
        int2* F = nullptr;
        int *group = nullptr, *groupBase = nullptr, *groupCount = nullptr, *list = nullptr;
        int n = 0, groups = 0, slots = 0, slotCap = 0;
        std::vector<int> base;  // host copy of groupBase
    } ec, tc;
    // continuous self collision: swept boxes, LBVHs over triangles and edges,
    // contact lists of both kinds, per-vertex colouring bids
    struct Self {
        float4 *triLo = nullptr, *triHi = nullptr, *edgeLo = nullptr, *edgeHi = nullptr;
        float4 *vtxLo = nullptr, *vtxHi = nullptr;
        Lbvh triB, edgeB, vtxB;  // vtxB: Morton order of the vertex queries (built per step)
        SelfBuf buf[2]{};
        unsigned long long* vmax = nullptr;
        int* vloc = nullptr;  // the colouring's per-vertex claims (vmax, vloc: one half per contact kind)
        int4* edgeVerts4 = nullptr;  // each cloth edge's vertices as (i1, i2, -1, -1): the edge tree's leaf payload
        int2 *candVT = nullptr, *candEE = nullptr;  // candidate pools (chunks of kCandChunk pairs)
        int *candFillVT = nullptr, *candFillEE = nullptr;  // pairs per chunk
        int candCapVT = 0, candCapEE = 0;    // pool capacities in pairs
        long candWantVT = 0, candWantEE = 0;  // capacities asked for by earlier steps' peaks
        int cap = 0;
        bool built = false;
    } self;
    DeviceSlabs triSlabs, bendSlabs, linkSlabs;
    DevicePacked triPacked, bendPacked;
    int triOrphans = 0;  // vertices in no triangle slab
    int bendOrphans = 0;  // vertices in no bend slab
    int slots = 0, slotCap = 0;
    int activeSlots = 0;  // vertex-contact slots the world can fill (0: no vertex contacts)
    SolverCtl* d_ctl = nullptr;
    // warm start: inertial positions, corrections, features, targets,
    // triangle strains, one-ring and incident-triangle lists, snapshots
    struct Warm {
        float4 *xhat = nullptr, *corr = nullptr, *T = nullptr, *triStrain = nullptr;
        uint16_t* X = nullptr;
        int *adjOff = nullptr, *adj = nullptr, *vtOff = nullptr, *vt = nullptr;
        float4 *snapPos = nullptr, *snapPrev = nullptr, *snapVel = nullptr, *snapCorr = nullptr;
        float4 *altPos = nullptr, *altPrev = nullptr, *altVel = nullptr, *altCorr = nullptr;  // plain judge's result
        int altNv = -1;
        double *resPartial = nullptr, *resOut = nullptr;
        int resBlocks = 0, nv = -1, nFree = 0;
        float ell = 1.f;
        bool built = false, saved = false;
    } warm;
    // render-mesh sync: indices, vertex -> (triangle, corner) lists, per-
    // triangle normals and angles, packed outputs (built for nv, nt)
    struct Render {
        unsigned* idx = nullptr;
        int* map = nullptr;
        int *vtOff = nullptr, *vt = nullptr;
        float4* triA = nullptr;
        float2* triB = nullptr;
        float *outV = nullptr, *outN = nullptr, *triN = nullptr;
        int nv = -1, nt = -1;
    } render;
    MouseSpringDev* d_mouse = nullptr;
    // the state a self-colliding step starts from, for its rerun after an overflow
    struct StepSnap {
        float4 *pos = nullptr, *prev = nullptr, *vel = nullptr, *corr = nullptr;
        int nv = -1;
        bool hasCorr = false;
    } snap;
    // links in the reference's order for the Jacobi / conjugate-residual
    // solvers (built when one is selected)
    struct Jacobi {
        int2* lk = nullptr;
        float2* lr = nullptr;
        float4* disp = nullptr;
        float* d = nullptr;
        int *off = nullptr, *inc = nullptr;
        int nl = -1, nv = -1;
    } jac;
    // two cached substep graphs (the warm-started and the plain substep)
    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
    std::string graphKey;
    hipGraph_t graph2 = nullptr;
    hipGraphExec_t graphExec2 = nullptr;
    std::string graphKey2;
    void release_graphs();
    hipEvent_t ev0 = nullptr, ev1 = nullptr;  // around the graph launches of a step
    KernelTimer timer;
};

class ClothModel {
public:
    explicit ClothModel(ClothPatch* owner) : mOwner(owner) {}

    // parameters, defaults of Physics/ClothModel.cpp:10-36
    float mUnit = 10.f, mThickness = 0.3f, mTolerance = 0.2f, mShearStiff = 5e8f, mBendStiff = 5.f;
    float mStretchStiff = 2e10f, mPoisson = 0.5f, mFriction = 0.4f, mArea = 10.f, mDragCoeff = 1.f;
    float mSelfShrink = 0.9f, mSelfCollThreshold = 0.003f;
    // continuous self-collision thresholds (SelfCollisionDetector.cpp:235,
    // 296, 313-317, 446 constants by default); capacity 0 = automatic
    float mCcdVolume = 0.001f, mCcdBisect = 0.01f, mCcdVtDistance = 0.1f;
    int mSelfCapacity = 0;
    int mSelfEdgeWeights = 0;  // 0: the reference's edge-edge displacement weights; 1: each end's own
    int mSelfCapacityGrown = 0;  // contact capacity an overflowed step grew to
    int selfReruns = 0;          // steps rerun after a self-collision overflow
    int lastSelfPeak[2] = {0, 0};  // most vertex-triangle / edge-edge contacts in one substep of the last step
    int lastSelfColours[2] = {0, 0};  // colour groups of each kind in the last substep
    int lastSelfUniq[2] = {0, 0};     // distinct vertices of each kind's contacts in the last substep
    int mNumIterations = 50, mCollFlags = 0, mBendFreq = 2;
    bool mDihedral = true, mFEM = false, mFixedIterations = false, mUseGraph = true;

    // host mirror of the particle state
    std::vector<f3> pos, prev, vel;
    std::vector<float> invMass;
    std::vector<scVector2> uv;
    std::vector<Triangle> tris;
    std::vector<Quad> quads;
    std::vector<Link> links;
    std::vector<MeshEdge> edges;
    std::vector<Bend> bends;

    // mouse spring (Physics/ClothTypes.h:85-92)
    bool msActive = false;
    int msI[3] = {0, 0, 0};
    float msA[3] = {0, 0, 0};
    f3 msPoint{0, 0, 0};
    float msLen = 0.f, msStiffness = 100.f;
    bool msDirty = true;  // the device copy of the spring is stale
    int mSolver = SOLVER_GAUSS_SEIDEL;  // ClothModel.cpp:24

    int statIterations = 0;
    float statBendError = 0.f, statContactError = 0.f, statSelfVTError = 0.f, statSelfEEError = 0.f;
    float statLinkError = 0.f;

    std::vector<ContactOut> contactsHost;
    std::vector<FeatureContactOut> edgeContactsHost, triContactsHost;
    struct SelfContactOut {
        unsigned long long key;
        int ids[4];
        float w[3], normal[3], side, lambda;
    };
    std::vector<SelfContactOut> selfContactsHost[2];  // processing order
    bool contactsValid = false;

    void clear();
    void add_triangle(uint32_t i1, uint32_t i2, uint32_t i3);
    // ClothModel.cpp:54-92 with explicit material coordinates
    void add_triangle_uv(uint32_t i1, uint32_t i2, uint32_t i3, scVector2 uv1, scVector2 uv2, scVector2 uv3);
    void add_link(uint32_t i1, uint32_t i2, float stiffness);
    void add_bend(const MeshEdge& e);

    // state synchronisation
    void mark_host_dirty() { mHostDirty = true; }
    void mark_topology_dirty() { mTopoDirty = true; mHostDirty = true; mColourDirty = true; }
    void sync_to_host();
    void sync_to_device();
    // the render mesh (vertices = particles + offset, normals, face normals)
    // from the device state
    void render_mesh(Mesh& mesh, f3 offset, const std::vector<int>& map);
    void ensure_jacobi();

    void step_substeps(float h, int n);
    void download_contacts();
    void fetch_stats();
    std::vector<std::pair<int, int>> self_candidates(int kind);
    void ensure_colouring();
    void set_sweep(int mode, int slabSize);

    Schedule triCol, bendCol, linkCol;
    // colour groups for obstacle-edge (cloth edges) and obstacle-vertex
    // (cloth triangles) contacts: greedy colourings in index order
    Schedule edgeContactCol, triContactCol;
    int mSweepMode = SWEEP_PACKED;
    int mSlabSize = 0;  // 0: chosen per constraint type by schedule_slabs_auto

    GpuSolver gpu;
    ClothPatch* mOwner;

private:
    bool mHostDirty = true, mTopoDirty = true, mDeviceNewer = false, mStatsPending = false, mColourDirty = true;
    void build_device_topology();
    void ensure_contact_capacity(int slots);
    void ensure_feature_capacity(int edgeSlots, int triSlots);
    void ensure_self();
    SelfParams self_params() const;
    void snapshot_step();
    void restore_step();
    void clear_self_counters();
    bool self_step_overflowed(int attempt);

public:
    // batch support: validate and upload everything a substep needs, then
    // describe the substep for the batch kernel (colour schedule required)
    void validate_batch() const;
    CollSet prepare_batch(int& slots);
    ClothDesc batch_desc(float h, const CollSet* d_cs, bool anyContacts) const;
    void after_device_step() { mDeviceNewer = true; mStatsPending = true; contactsValid = false; }

    // learned warm start: mode 0 off, 1 train on the plain solve, 2 warm-started
    // solve, 3 plain solve that only tracks the corrections (features stay current)
    WarmStart* mWarm = nullptr;
    int mWarmMode = 0, mWarmIters = 4;
    float mWarmThreshold = -1.f;  // fallback residual; < 0: the plain step from the same start judges each step
    double mWarmResidual = -1.0, mWarmPlainResidual = -1.0, mWarmWarmResidual = -1.0;
    int mWarmFallbacks = 0, mWarmSteps = 0;
    bool mWarmResetCorr = true, mWarmPlainNow = false, mWarmLastPlain = false;
    void set_warm(WarmStart* ws, int mode, int iterations);
    void ensure_warm();
    WarmBufs warm_bufs() const;
    double mLastStepMs = 0.0;  // device time of the last step's graph launches (warm-started and fallback)
    double residual();
    void residual_terms(double out[3]);  // RMS of warp/weft, of shear, of bend errors
    double residualSum = 0.0;
    void save_state();
    void restore_state();
    void keep_plain_result();
    void restore_plain_result();
    int warm_data(int which, float* out, int count);
    void run_substeps(float h, int n, const CollSet& cs, bool timed);
    CollSet make_collset(int& slots, int& edgeSlots, int& triSlots);
    FeatureContactBuf feature_buf(GpuSolver::Features& f) const;
    void record_substep(float h, const CollSet& cs, bool timed);
    std::string graph_key(float h, const CollSet& cs) const;
};

class ClothPatch {
public:
    explicit ClothPatch(CollisionWorld* world) : mWorld(world), mModel(new ClothModel(this)) {}
    void init_grid(int divX, int divY, float inc, f3 offset, bool horizontal, bool attached);
    // Physics/ClothPatch.cpp:43-133: a patch from a triangle mesh
    void init_mesh(const Mesh& mesh, f3 offset, bool attached);
    void step(float dt);
    void update_mesh();
    void compute_strain_map(scVector3* colors, int count);

    CollisionWorld* mWorld;
    std::unique_ptr<ClothModel> mModel;
    Mesh mMesh;
    int mNumSteps = 1;
    float mIncline = 0.f;
    f3 mPosition{0.f, 0.f, 0.f};
    // mesh-based patches: render-mesh vertex -> particle (duplicate vertices
    // welded); empty for grid patches, whose render vertices are the particles
    std::vector<int> mMap;
};

// Many independent cloths stepped together: one launch per substep, one
// workgroup per cloth (largest first), each patch's own device buffers.
class ClothBatch {
public:
    ~ClothBatch();
    std::vector<ClothPatch*> patches;
    int blockSize = 0;  // 0: autotune over the first steps
    int stage = -1;     // -1: autotune; 0: positions in HBM; 1: in LDS when they fit
    bool tuned = false;
    bool useGraph = true;  // replay the captured substep launches (else launch directly)
    double lastStepMs = 0.0;  // device time of the last step's substep launches
    int lastSubsteps = 0;
    // per-cloth workgroup clocks of the last substep launch (recordClock):
    // start, end (device clock ticks) and the unit each cloth's workgroup
    // ran on, in patch order
    bool recordClock = false;
    std::vector<unsigned long long> clockStart, clockEnd, clockUnit;
    void add(ClothPatch* p);
    void step(float dt);
    void set_launch(int bs, int st);

private:
    hipStream_t stream = nullptr;
    ClothDesc* d_descs = nullptr;
    CollSet* d_cs = nullptr;
    int cap = 0;
    std::vector<ClothDesc> hostDescs;
    std::vector<CollSet> hostCs;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string graphKey;
    int stageLimit = 0;
    unsigned long long* d_clock = nullptr;  // 3 words per cloth, dispatch order
    int clockCap = 0;
    std::vector<std::pair<int, int>> trials;  // (block size, stage) candidates
public:
    std::vector<std::pair<int, int>> trialConfigs;  // timed so far, with trialMs
    std::vector<double> trialMs;
private:
    void release_graph();
};

void set_error(const std::string& msg);

}  // namespace sc
