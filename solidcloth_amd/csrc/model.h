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
#include "schedule.h"

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

// Indexed triangle mesh (Geometry/Mesh.h:12-264)
struct Mesh {
    std::vector<scVector3> vertices;
    std::vector<unsigned> indices;
    std::vector<scVector3> normals;
    std::vector<scVector2> uvs;
    std::vector<MeshEdge> edges;
    void clear();
    void construct_edges();
    void compute_normals();
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
    int coll, vertex, slot;
    f3 point, normal;
};

class ClothPatch;

// Sweep schedules: 0 = one launch per colour, 1 = slab phases with LDS staging
enum { SWEEP_COLOURS = 0, SWEEP_SLABS = 1 };

// Device copy of one constraint schedule's slab structure.
struct DeviceSlabs {
    ushort4* lidx = nullptr;
    int4* info = nullptr;
    int* colOff = nullptr;
    int* verts = nullptr;
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
    int4* d_bidx = nullptr;
    int* d_cCount = nullptr;
    float4 *d_cP = nullptr, *d_cN = nullptr;
    int* d_cSrc = nullptr;
    DeviceSlabs triSlabs, bendSlabs;
    int triOrphans = 0;  // vertices in no triangle slab
    int slots = 0, slotCap = 0;
    SolverCtl* d_ctl = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
    std::string graphKey;
    KernelTimer timer;
};

class ClothModel {
public:
    explicit ClothModel(ClothPatch* owner) : mOwner(owner) {}

    // parameters, defaults of Physics/ClothModel.cpp:10-36
    float mUnit = 10.f, mThickness = 0.3f, mTolerance = 0.2f, mShearStiff = 5e8f, mBendStiff = 5.f;
    float mStretchStiff = 2e10f, mPoisson = 0.5f, mFriction = 0.4f, mArea = 10.f, mDragCoeff = 1.f;
    float mSelfShrink = 0.9f, mSelfCollThreshold = 0.003f;
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

    int statIterations = 0;
    float statBendError = 0.f, statContactError = 0.f;

    std::vector<ContactOut> contactsHost;
    bool contactsValid = false;

    void clear();
    void add_triangle(uint32_t i1, uint32_t i2, uint32_t i3);
    void add_link(uint32_t i1, uint32_t i2, float stiffness);
    void add_bend(const MeshEdge& e);

    // state synchronisation
    void mark_host_dirty() { mHostDirty = true; }
    void mark_topology_dirty() { mTopoDirty = true; mHostDirty = true; mColourDirty = true; }
    void sync_to_host();
    void sync_to_device();

    void step_substeps(float h, int n);
    void download_contacts();
    void fetch_stats();
    void ensure_colouring();
    void set_sweep(int mode, int slabSize);

    Schedule triCol, bendCol;
    int mSweepMode = SWEEP_SLABS;
    int mSlabSize = 0;  // 0: chosen per constraint type by schedule_slabs_auto

    GpuSolver gpu;
    ClothPatch* mOwner;

private:
    bool mHostDirty = true, mTopoDirty = true, mDeviceNewer = false, mStatsPending = false, mColourDirty = true;
    void build_device_topology();
    void ensure_contact_capacity(int slots);
    CollSet make_collset(int& slots) const;
    void record_substep(float h, const CollSet& cs, bool timed);
    std::string graph_key(float h, const CollSet& cs) const;
};

class ClothPatch {
public:
    explicit ClothPatch(CollisionWorld* world) : mWorld(world), mModel(new ClothModel(this)) {}
    void init_grid(int divX, int divY, float inc, f3 offset, bool horizontal, bool attached);
    void step(float dt);
    void update_mesh();
    void compute_strain_map(scVector3* colors, int count);

    CollisionWorld* mWorld;
    std::unique_ptr<ClothModel> mModel;
    Mesh mMesh;
    int mNumSteps = 1;
    float mIncline = 0.f;
    f3 mPosition{0.f, 0.f, 0.f};
};

void set_error(const std::string& msg);

}  // namespace sc
