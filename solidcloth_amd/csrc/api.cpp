// The C ABI (include/solidcloth.h). Each export mirrors one function of the
// reference's Demo/ClothAPI.cpp (cited per function) over the MI355X-native
// objects of model.h. Exceptions never cross the boundary: they become the
// thread's last error and the call returns a neutral value.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <stdexcept>

#include "hiputil.h"
#include "lbvh.h"
#include "model.h"

namespace sc {

static thread_local std::string g_lastError;
static thread_local int g_errorCount = 0;

void set_error(const std::string& msg)
{
    g_lastError = msg;
    g_errorCount++;
}

template <class F>
static void guard(F&& f)
{
    try {
        f();
    } catch (const std::exception& e) {
        set_error(e.what());
    } catch (...) {
        set_error("unknown error");
    }
}

template <class R, class F>
static R guard_r(R fallback, F&& f)
{
    try {
        return f();
    } catch (const std::exception& e) {
        set_error(e.what());
    } catch (...) {
        set_error("unknown error");
    }
    return fallback;
}

static CollisionWorld* W(scCollisionWorld w)
{
    if (!w.ptr) throw std::runtime_error("null collision world");
    return (CollisionWorld*)w.ptr;
}
static ClothPatch* P(scClothPatch p)
{
    if (!p.ptr) throw std::runtime_error("null cloth patch");
    return (ClothPatch*)p.ptr;
}
static ClothModel* M(scModel m)
{
    if (!m.ptr) throw std::runtime_error("null cloth model");
    return (ClothModel*)m.ptr;
}
static Mesh* ME(scMesh m)
{
    if (!m.ptr) throw std::runtime_error("null mesh");
    return (Mesh*)m.ptr;
}
static Collidable* C(scHandle h)
{
    if (!h) throw std::runtime_error("null collidable");
    return (Collidable*)h;
}
static void check_index(int i, size_t n)
{
    if (i < 0 || (size_t)i >= n) throw std::runtime_error("index out of range");
}

}  // namespace sc

using namespace sc;

extern "C" {

// ------------------------------------------------- world (ClothAPI.cpp:13-49)
scCollisionWorld scCreateCollisionWorld(void)
{
    scCollisionWorld w;
    w.ptr = guard_r<void*>(nullptr, [] { return (void*)new CollisionWorld(); });
    return w;
}

void scFreeCollisionWorld(scCollisionWorld collWorld) { guard([&] { delete (CollisionWorld*)collWorld.ptr; }); }

void scClearCollidables(scCollisionWorld collWorld) { guard([&] { W(collWorld)->colls.clear(); }); }

int scGetNumCollidables(scCollisionWorld collWorld)
{
    return guard_r<int>(0, [&] { return (int)W(collWorld)->colls.size(); });
}

void scAddCollidable(scCollisionWorld collWorld, scHandle coll)
{
    guard([&] { W(collWorld)->colls.push_back(std::shared_ptr<Collidable>(C(coll))); });
}

scHandle scGetCollidable(scCollisionWorld collWorld, int i)
{
    return guard_r<scHandle>(nullptr, [&] {
        CollisionWorld* w = W(collWorld);
        check_index(i, w->colls.size());
        return (scHandle)w->colls[i].get();
    });
}

// --------------------------------------------- collidables (ClothAPI.cpp:51-74)
scHandle scCreateWalls(scAabb3 box)
{
    return guard_r<scHandle>(nullptr, [&] {
        Walls* w = new Walls();
        w->type = CT_WALLS;
        w->box = box;
        return (scHandle)w;
    });
}

scHandle scCreateSphere(scVector3 center, float radius)
{
    return guard_r<scHandle>(nullptr, [&] {
        Sphere* s = new Sphere();
        s->type = CT_SPHERE;
        s->center = f3{center.x, center.y, center.z};
        s->radius = radius;
        return (scHandle)s;
    });
}

scHandle scCreateCapsule(scVector3 c, float rad, float halfHeight, scQuaternion quat)
{
    return guard_r<scHandle>(nullptr, [&] {
        Capsule* k = new Capsule();
        k->type = CT_CAPSULE;
        k->center = f3{c.x, c.y, c.z};
        k->r = rad;
        k->hh = halfHeight;
        k->rot = quat;
        return (scHandle)k;
    });
}

scHandle scCreateSDF(scSDF s)
{
    return guard_r<scHandle>(nullptr, [&] {
        if (!s.ptr) throw std::runtime_error("null SDF");
        CollisionSDF* c = new CollisionSDF();
        c->type = CT_SDF;
        c->sdf = (SDF*)s.ptr;
        return (scHandle)c;
    });
}

scHandle scCreateMesh(scMesh m)
{
    return guard_r<scHandle>(nullptr, [&] {
        CollisionMesh* c = new CollisionMesh();
        c->type = CT_MESH;
        c->mesh = ME(m);
        return (scHandle)c;
    });
}

// ------------------------------------------- introspection (ClothAPI.cpp:76-124)
int scGetCollidableType(scHandle collidable)
{
    return guard_r<int>(CT_INVALID, [&] { return C(collidable)->type; });
}

void* scGetMeshTree(scHandle collMesh)
{
    // The obstacle tree is not built by this library; the reference returns
    // nullptr for non-mesh collidables and for a mesh not yet stepped.
    return guard_r<void*>(nullptr, [&] {
        (void)C(collMesh);
        return (void*)nullptr;
    });
}

void scGetWallsInfo(scHandle walls, scAabb3* box)
{
    guard([&] {
        Collidable* c = C(walls);
        if (c->type != CT_WALLS || !box) return;
        *box = ((Walls*)c)->box;
    });
}

void scGetSphereInfo(scHandle sphere, float* radius, scVector3* center)
{
    guard([&] {
        Collidable* c = C(sphere);
        if (c->type != CT_SPHERE) return;
        Sphere* s = (Sphere*)c;
        if (center) *center = scVector3{s->center.x, s->center.y, s->center.z};
        if (radius) *radius = s->radius;
    });
}

void scGetCapsuleInfo(scHandle capsule, float* radius, float* halfHeight, scVector3* center, scQuaternion* rot)
{
    guard([&] {
        Collidable* c = C(capsule);
        if (c->type != CT_CAPSULE) return;
        Capsule* k = (Capsule*)c;
        if (radius) *radius = k->r;
        if (halfHeight) *halfHeight = k->hh;
        if (center) *center = scVector3{k->center.x, k->center.y, k->center.z};
        if (rot) *rot = k->rot;
    });
}

// ------------------------------------------------ patch (ClothAPI.cpp:126-178)
scClothPatch scCreateClothPatch(scCollisionWorld collWorld)
{
    scClothPatch p;
    p.ptr = guard_r<void*>(nullptr, [&] { return (void*)new ClothPatch(W(collWorld)); });
    return p;
}

void scFreeClothPatch(scClothPatch patch) { guard([&] { delete (ClothPatch*)patch.ptr; }); }

scModel scGetClothModel(scClothPatch patch)
{
    scModel m;
    m.ptr = guard_r<void*>(nullptr, [&] { return (void*)P(patch)->mModel.get(); });
    return m;
}

scMesh scGetClothMesh(scClothPatch patch)
{
    scMesh m;
    m.ptr = guard_r<void*>(nullptr, [&] { return (void*)&P(patch)->mMesh; });
    return m;
}

void scInitCloth(scClothPatch patch, int divX, int divY, float inc, scVector3 offset, bool horizontal, bool attached)
{
    guard([&] {
        if (divX < 2 || divY < 2) throw std::runtime_error("scInitCloth needs divX, divY >= 2");
        P(patch)->init_grid(divX, divY, inc, f3{offset.x, offset.y, offset.z}, horizontal, attached);
    });
}

void scStepCloth(scClothPatch patch, float dt)
{
    guard([&] { P(patch)->step(dt); });
}

void scUpdateClothMesh(scClothPatch patch)
{
    guard([&] { P(patch)->update_mesh(); });
}

void scComputeStrainMap(scClothPatch patch, scVector3* colors, int count)
{
    guard([&] {
        if (!colors) throw std::runtime_error("null colour buffer");
        P(patch)->compute_strain_map(colors, count);
    });
}

// ---------------------------------------------- model getters (ClothAPI.cpp:180-300)
int scGetNumParticles(scModel model)
{
    return guard_r<int>(0, [&] { return (int)M(model)->pos.size(); });
}

void scGetParticlePos(scModel model, int i, scVector3* pos)
{
    guard([&] {
        ClothModel* m = M(model);
        check_index(i, m->pos.size());
        m->sync_to_host();
        if (pos) *pos = scVector3{m->pos[i].x, m->pos[i].y, m->pos[i].z};
    });
}

float scGetParticleInvMass(scModel model, int i)
{
    return guard_r<float>(0.f, [&] {
        ClothModel* m = M(model);
        check_index(i, m->invMass.size());
        return m->invMass[i];
    });
}

int scGetNumTris(scModel model)
{
    return guard_r<int>(0, [&] { return (int)M(model)->tris.size(); });
}

void scGetTriangle(scModel model, int i, int* i1, int* i2, int* i3)
{
    guard([&] {
        ClothModel* m = M(model);
        check_index(i, m->tris.size());
        const Triangle& t = m->tris[i];
        if (i1) *i1 = (int)t.i1;
        if (i2) *i2 = (int)t.i2;
        if (i3) *i3 = (int)t.i3;
    });
}

void scGetTriangleInfo(scModel model, int i, float* invDet, scVector2* du, scVector2* dv, scVector2* s)
{
    guard([&] {
        ClothModel* m = M(model);
        check_index(i, m->tris.size());
        const Triangle& t = m->tris[i];
        if (invDet) *invDet = t.invDet;
        if (du) *du = scVector2{t.du1, t.du2};
        if (dv) *dv = scVector2{t.dv1, t.dv2};
        if (s) *s = scVector2{t.su, t.sv};
    });
}

int scGetNumQuads(scModel model)
{
    return guard_r<int>(0, [&] { return (int)M(model)->quads.size(); });
}

void scGetQuad(scModel model, int i, int* i1, int* i2, int* i3, int* i4)
{
    guard([&] {
        ClothModel* m = M(model);
        check_index(i, m->quads.size());
        const Quad& q = m->quads[i];
        if (i1) *i1 = (int)q.i1;
        if (i2) *i2 = (int)q.i2;
        if (i3) *i3 = (int)q.i3;
        if (i4) *i4 = (int)q.i4;
    });
}

int scGetNumLinks(scModel model)
{
    return guard_r<int>(0, [&] { return (int)M(model)->links.size(); });
}

void scGetLink(scModel model, int i, int* i1, int* i2)
{
    guard([&] {
        ClothModel* m = M(model);
        check_index(i, m->links.size());
        if (i1) *i1 = (int)m->links[i].i1;
        if (i2) *i2 = (int)m->links[i].i2;
    });
}

int scGetNumContacts(scModel model)
{
    return guard_r<int>(0, [&] {
        ClothModel* m = M(model);
        m->download_contacts();
        return (int)m->contactsHost.size();
    });
}

void scGetContact(scModel model, int i, scVector3* point, scVector3* normal)
{
    guard([&] {
        ClothModel* m = M(model);
        m->download_contacts();
        check_index(i, m->contactsHost.size());
        const ContactOut& c = m->contactsHost[i];
        if (point) *point = scVector3{c.point.x, c.point.y, c.point.z};
        if (normal) *normal = scVector3{c.normal.x, c.normal.y, c.normal.z};
    });
}

int scGetNumTriContacts(scModel model)
{
    return guard_r<int>(0, [&] {
        ClothModel* m = M(model);
        m->download_contacts();
        return (int)m->triContactsHost.size();
    });
}

void scGetTriContact(scModel model, int i, scVector3* point, scVector3* normal)
{
    guard([&] {
        ClothModel* m = M(model);
        m->download_contacts();
        check_index(i, m->triContactsHost.size());
        const FeatureContactOut& c = m->triContactsHost[i];
        if (point) *point = scVector3{c.point.x, c.point.y, c.point.z};
        if (normal) *normal = scVector3{c.normal.x, c.normal.y, c.normal.z};
    });
}

// ------------------------------------------------- parameters (ClothAPI.cpp:302-380)
int scGetCollisionFlags(scModel model) { return guard_r<int>(0, [&] { return M(model)->mCollFlags; }); }
void scSetCollisionFlags(scModel model, int val) { guard([&] { M(model)->mCollFlags = val; }); }
float scGetThickness(scModel model) { return guard_r<float>(0.f, [&] { return M(model)->mThickness; }); }
void scSetThickness(scModel model, float val) { guard([&] { M(model)->mThickness = val; }); }
float scGetTolerance(scModel model) { return guard_r<float>(0.f, [&] { return M(model)->mTolerance; }); }
void scSetTolerance(scModel model, float val) { guard([&] { M(model)->mTolerance = val; }); }
float scGetFriction(scModel model) { return guard_r<float>(0.f, [&] { return M(model)->mFriction; }); }
void scSetFriction(scModel model, float val) { guard([&] { M(model)->mFriction = val; }); }
int scGetNumIterations(scModel model) { return guard_r<int>(0, [&] { return M(model)->mNumIterations; }); }
void scSetNumIterations(scModel model, int val) { guard([&] { M(model)->mNumIterations = val; }); }
bool scGetUseFEM(scModel model) { return guard_r<bool>(false, [&] { return M(model)->mFEM; }); }
void scSetUseFEM(scModel model, bool val) { guard([&] { M(model)->mFEM = val; }); }

// ------------------------------------------ tree and mouse spring (ClothAPI.cpp:382-410)
void* scGetClothTree(scModel model)
{
    // No cloth AABB tree exists on this path: the reference rebuilds one per
    // substep but nothing on the external-contact path consumes it.
    return guard_r<void*>(nullptr, [&] {
        (void)M(model);
        return (void*)nullptr;
    });
}

// Physics/ClothModel.inl:86-98
void scAddMouseSpring(scModel model, int i1, int i2, int i3, scVector3 coords, scVector3 p)
{
    guard([&] {
        ClothModel* m = M(model);
        m->msActive = true;
        m->msLen = 0.f;
        m->msStiffness = 100.f;
        m->msI[0] = i1; m->msI[1] = i2; m->msI[2] = i3;
        m->msA[0] = coords.x; m->msA[1] = coords.y; m->msA[2] = coords.z;
        m->msPoint = f3{p.x, p.y, p.z};
        m->msDirty = true;
    });
}

void scUpdateMouseSpring(scModel model, scVector3 p)
{
    guard([&] {
        ClothModel* m = M(model);
        if (m->msActive) {
            m->msPoint = f3{p.x, p.y, p.z};
            m->msDirty = true;
        }
    });
}

void scRemoveMouseSpring(scModel model) { guard([&] { M(model)->msActive = false; }); }

// -------------------------------------------------- mesh buffers (ClothAPI.cpp:412-440)
int scGetMeshVertexCount(scMesh mesh) { return guard_r<int>(0, [&] { return (int)ME(mesh)->vertices.size(); }); }
int scGetMeshIndexCount(scMesh mesh) { return guard_r<int>(0, [&] { return (int)ME(mesh)->indices.size(); }); }
scVector3* scGetMeshNormalBuffer(scMesh mesh)
{
    return guard_r<scVector3*>(nullptr, [&] { return ME(mesh)->normals.empty() ? nullptr : ME(mesh)->normals.data(); });
}
scVector3* scGetMeshVertexBuffer(scMesh mesh)
{
    return guard_r<scVector3*>(nullptr,
                               [&] { return ME(mesh)->vertices.empty() ? nullptr : ME(mesh)->vertices.data(); });
}
unsigned int* scGetMeshIndexBuffer(scMesh mesh)
{
    return guard_r<unsigned int*>(nullptr,
                                  [&] { return ME(mesh)->indices.empty() ? nullptr : ME(mesh)->indices.data(); });
}

// ================================================================ extensions
void scSetUnit(scModel model, float val) { guard([&] { M(model)->mUnit = val; }); }
void scInitClothFromMesh(scClothPatch patch, scMesh mesh, scVector3 offset, bool attached)
{
    guard([&] {
        if (!mesh.ptr) throw std::runtime_error("null mesh");
        P(patch)->init_mesh(*ME(mesh), f3{offset.x, offset.y, offset.z}, attached);
    });
}

int scSetMeshUVs(scMesh mesh, const scVector2* uvs, int count)
{
    return guard_r<int>(-1, [&] {
        Mesh* me = ME(mesh);
        if (count != (int)me->vertices.size()) throw std::runtime_error("one uv per mesh vertex expected");
        me->uvs.assign(uvs, uvs + count);
        return count;
    });
}

int scSetMeshColors(scMesh mesh, const scVector3* colors, int count)
{
    return guard_r<int>(-1, [&] {
        Mesh* me = ME(mesh);
        if (count != (int)me->vertices.size()) throw std::runtime_error("one colour per mesh vertex expected");
        me->colors.assign(colors, colors + count);
        return count;
    });
}

void scSetSolver(scModel model, int solver)
{
    guard([&] {
        if (solver < SOLVER_JACOBI || solver > SOLVER_EXPLICIT) throw std::runtime_error("unknown solver type");
        M(model)->mSolver = solver;
    });
}
int scGetSolver(scModel model) { return guard_r<int>(-1, [&] { return M(model)->mSolver; }); }
float scGetUnit(scModel model) { return guard_r<float>(0.f, [&] { return M(model)->mUnit; }); }
void scSetShearStiffness(scModel model, float val) { guard([&] { M(model)->mShearStiff = val; }); }
void scSetStretchStiffness(scModel model, float val) { guard([&] { M(model)->mStretchStiff = val; }); }
void scSetBendStiffness(scModel model, float val) { guard([&] { M(model)->mBendStiff = val; }); }
void scSetPoisson(scModel model, float val) { guard([&] { M(model)->mPoisson = val; }); }
void scSetDihedral(scModel model, bool val) { guard([&] { M(model)->mDihedral = val; }); }
void scSetDragCoeff(scModel model, float val) { guard([&] { M(model)->mDragCoeff = val; }); }
float scGetShearStiffness(scModel model) { return guard_r<float>(0.f, [&] { return M(model)->mShearStiff; }); }
float scGetStretchStiffness(scModel model) { return guard_r<float>(0.f, [&] { return M(model)->mStretchStiff; }); }
float scGetBendStiffness(scModel model) { return guard_r<float>(0.f, [&] { return M(model)->mBendStiff; }); }
float scGetPoisson(scModel model) { return guard_r<float>(0.f, [&] { return M(model)->mPoisson; }); }
bool scGetDihedral(scModel model) { return guard_r<bool>(false, [&] { return M(model)->mDihedral; }); }
float scGetDragCoeff(scModel model) { return guard_r<float>(0.f, [&] { return M(model)->mDragCoeff; }); }
void scGetSelfParams(scModel model, float* shrink, float* threshold)
{
    guard([&] {
        ClothModel* m = M(model);
        if (shrink) *shrink = m->mSelfShrink;
        if (threshold) *threshold = m->mSelfCollThreshold;
    });
}
void scSetNumSteps(scClothPatch patch, int val)
{
    guard([&] {
        if (val < 1) throw std::runtime_error("NumSteps must be >= 1");
        P(patch)->mNumSteps = val;
    });
}
int scGetNumSteps(scClothPatch patch) { return guard_r<int>(0, [&] { return P(patch)->mNumSteps; }); }
void scSetFixedIterations(scModel model, bool fixed) { guard([&] { M(model)->mFixedIterations = fixed; }); }
void scSetInitIncline(scClothPatch patch, float incline) { guard([&] { P(patch)->mIncline = incline; }); }

scSDF scCreateGridSDF(const float* values, int nx, int ny, int nz, scAabb3 box)
{
    scSDF s;
    s.ptr = guard_r<void*>(nullptr, [&] {
        if (!values || nx < 1 || ny < 1 || nz < 1) throw std::runtime_error("bad grid SDF arguments");
        SDF* g = new SDF();
        try {
            g->create(nx, ny, nz, box);
            memcpy(g->vals.data(), values, g->vals.size() * sizeof(float));
            g->compute_gradient();
        } catch (...) {
            delete g;
            throw;
        }
        return (void*)g;
    });
    return s;
}

void scFreeSDF(scSDF sdf) { guard([&] { delete (SDF*)sdf.ptr; }); }

scSDF scCreateSDFFromMesh(scMesh mesh, int nx, int ny, int nz, scAabb3 box)
{
    scSDF s;
    s.ptr = guard_r<void*>(nullptr, [&] {
        if (!mesh.ptr || nx < 1 || ny < 1 || nz < 1) throw std::runtime_error("bad SDF-from-mesh arguments");
        Mesh* me = ME(mesh);
        if (me->indices.empty()) throw std::runtime_error("the mesh has no triangles");
        std::unique_ptr<SDF> g(new SDF());
        g->create(nx, ny, nz, box);
        hipStream_t st;
        HIP_OK(hipStreamCreate(&st));
        float* d = nullptr;
        try {
            const MeshGeo& geo = me->device_geo(st);
            dev_alloc(d, g->vals.size());
            launch_sdf_from_mesh(st, geo, nx, ny, nz, f3{box.min.x, box.min.y, box.min.z}, g->steps, d);
            HIP_OK(hipGetLastError());
            HIP_OK(hipMemcpyAsync(g->vals.data(), d, g->vals.size() * sizeof(float), hipMemcpyDeviceToHost, st));
            HIP_OK(hipStreamSynchronize(st));
        } catch (...) {
            dev_free(d);
            hipStreamDestroy(st);
            throw;
        }
        dev_free(d);
        HIP_OK(hipStreamDestroy(st));
        g->compute_gradient();
        return (void*)g.release();
    });
    return s;
}

scMesh scCreateMeshFromArrays(const scVector3* vertices, int numVertices, const unsigned int* indices, int numIndices)
{
    scMesh m;
    m.ptr = guard_r<void*>(nullptr, [&] {
        if (!vertices || !indices || numVertices < 0 || numIndices < 0 || numIndices % 3)
            throw std::runtime_error("bad mesh arrays");
        for (int i = 0; i < numIndices; i++)
            if ((int)indices[i] >= numVertices) throw std::runtime_error("mesh index out of range");
        Mesh* me = new Mesh();
        me->vertices.assign(vertices, vertices + numVertices);
        me->indices.assign(indices, indices + numIndices);
        me->compute_normals();
        me->construct_edges();
        me->compute_edge_normals();
        return (void*)me;
    });
    return m;
}

void scFreeMesh(scMesh mesh) { guard([&] { delete (Mesh*)mesh.ptr; }); }

// ------------------------------------------------------------ file formats

namespace {

using FileCloser = std::unique_ptr<FILE, int (*)(FILE*)>;

FileCloser open_file(const char* path, const char* mode, const char* what)
{
    if (!path) throw std::runtime_error(std::string("no ") + what + " path");
    FileCloser f(fopen(path, mode), fclose);
    if (!f) throw std::runtime_error(std::string("cannot open ") + what + " file " + path);
    return f;
}

// Whitespace-separated words of one line.
std::vector<std::string> words(const std::string& line)
{
    std::vector<std::string> w;
    size_t i = 0;
    while (i < line.size()) {
        while (i < line.size() && isspace((unsigned char)line[i])) i++;
        size_t j = i;
        while (j < line.size() && !isspace((unsigned char)line[j])) j++;
        if (j > i) w.push_back(line.substr(i, j - i));
        i = j;
    }
    return w;
}

unsigned obj_index(const std::string& tok, int lineNo)
{
    char* end = nullptr;
    long v = strtol(tok.c_str(), &end, 10);
    if (end == tok.c_str() || v < 1)
        throw std::runtime_error("OBJ line " + std::to_string(lineNo) + ": unsupported index '" + tok + "'");
    return (unsigned)(v - 1);
}

}  // namespace

// Geometry/SDF.cpp:115-151: 3 int cell counts, the box as 6 doubles
// (min.x, max.x, min.y, max.y, min.z, max.z), then the node values as doubles,
// x fastest; values and box are narrowed to float as the reference stores them.
scSDF scLoadSDFFile(const char* path)
{
    scSDF s;
    s.ptr = guard_r<void*>(nullptr, [&] {
        FileCloser f = open_file(path, "rb", "SDF");
        int32_t n[3];
        double b[6];
        if (fread(n, sizeof(int32_t), 3, f.get()) != 3 || fread(b, sizeof(double), 6, f.get()) != 6)
            throw std::runtime_error(std::string("truncated SDF header in ") + path);
        for (int a = 0; a < 3; a++)
            if (n[a] < 1 || n[a] > 4095) throw std::runtime_error("SDF cell count out of range (1..4095)");
        if ((double)(n[0] + 1.0) * (n[1] + 1.0) * (n[2] + 1.0) > (double)INT_MAX)
            throw std::runtime_error("SDF grid out of range: more than 2^31 - 1 nodes");
        size_t count = (size_t)(n[0] + 1) * (n[1] + 1) * (n[2] + 1);
        // the values must all be there before they are allocated
        const long here = ftell(f.get());
        if (here < 0 || fseek(f.get(), 0, SEEK_END) != 0)
            throw std::runtime_error(std::string("cannot seek in SDF file ") + path);
        const long end = ftell(f.get());
        if (end < 0 || (unsigned long)(end - here) < count * sizeof(double))
            throw std::runtime_error(std::string("truncated SDF values in ") + path);
        if (fseek(f.get(), here, SEEK_SET) != 0) throw std::runtime_error(std::string("cannot seek in SDF file ") + path);
        std::vector<double> v(count);
        if (fread(v.data(), sizeof(double), count, f.get()) != count)
            throw std::runtime_error(std::string("truncated SDF values in ") + path);
        scAabb3 box{{(float)b[0], (float)b[2], (float)b[4]}, {(float)b[1], (float)b[3], (float)b[5]}};
        std::unique_ptr<SDF> g(new SDF());
        g->create(n[0], n[1], n[2], box);
        for (size_t i = 0; i < count; i++) g->vals[i] = (float)v[i];
        g->compute_gradient();
        return (void*)g.release();
    });
    return s;
}

// Geometry/SDF.cpp:75-113 (WriteToFile): the same layout as scLoadSDFFile reads.
int scWriteSDFFile(scSDF sdf, const char* path)
{
    return guard_r<int>(0, [&] {
        const SDF* g = (const SDF*)sdf.ptr;
        if (!g) throw std::runtime_error("null SDF");
        FileCloser f = open_file(path, "wb", "SDF");
        int32_t n[3] = {g->nc[0], g->nc[1], g->nc[2]};
        double b[6] = {g->box.min.x, g->box.max.x, g->box.min.y, g->box.max.y, g->box.min.z, g->box.max.z};
        std::vector<double> v(g->vals.begin(), g->vals.end());
        if (fwrite(n, sizeof(int32_t), 3, f.get()) != 3 || fwrite(b, sizeof(double), 6, f.get()) != 6 ||
            fwrite(v.data(), sizeof(double), v.size(), f.get()) != v.size())
            throw std::runtime_error(std::string("cannot write SDF file ") + path);
        return 1;
    });
}

int scGetSDFGrid(scSDF sdf, int* cells, scAabb3* box, float* values, int capacity)
{
    return guard_r<int>(0, [&] {
        const SDF* g = (const SDF*)sdf.ptr;
        if (!g) throw std::runtime_error("null SDF");
        if (cells) for (int a = 0; a < 3; a++) cells[a] = g->nc[a];
        if (box) *box = g->box;
        if (values && capacity > 0)
            memcpy(values, g->vals.data(), std::min((size_t)capacity, g->vals.size()) * sizeof(float));
        return (int)g->vals.size();
    });
}

// Geometry/Assets.cpp:309-451 (LoadMeshFromObj). Reads v, vt, f lines (vn are
// skipped: the mesh's normals are computed from its triangles, as for every
// mesh built here); each face contributes its first three corners. naive:
// one mesh vertex per v line and the faces' position indices. Otherwise one
// mesh vertex per distinct (position, normal, uv) corner in order of first
// appearance, carrying its position and uv.
scMesh scLoadMeshFromObj(const char* path, bool naive)
{
    scMesh m;
    m.ptr = guard_r<void*>(nullptr, [&] {
        if (!path) throw std::runtime_error("no OBJ path");
        std::ifstream in(path);
        if (!in) throw std::runtime_error(std::string("cannot open OBJ file ") + path);
        std::vector<scVector3> pos;
        std::vector<scVector2> tex;
        std::vector<unsigned> idxV, idxU;
        std::vector<std::tuple<unsigned, unsigned, unsigned>> corners;  // (position, normal, uv)
        std::map<std::tuple<unsigned, unsigned, unsigned>, unsigned> first;
        std::string line;
        int lineNo = 0;
        while (std::getline(in, line)) {
            lineNo++;
            if (line.size() < 2) continue;
            std::vector<std::string> w = words(line);
            if (w.empty()) continue;
            if (w[0] == "v") {
                if (w.size() < 4) throw std::runtime_error("OBJ line " + std::to_string(lineNo) + ": short vertex");
                pos.push_back({strtof(w[1].c_str(), nullptr), strtof(w[2].c_str(), nullptr),
                               strtof(w[3].c_str(), nullptr)});
            } else if (w[0] == "vt") {
                // u from the first number, v from the last (atof, then narrowed)
                if (w.size() < 3) throw std::runtime_error("OBJ line " + std::to_string(lineNo) + ": short uv");
                tex.push_back({(float)strtod(w[1].c_str(), nullptr), (float)strtod(w.back().c_str(), nullptr)});
            } else if (w[0] == "f") {
                if (w.size() < 4) throw std::runtime_error("OBJ line " + std::to_string(lineNo) + ": face with < 3 corners");
                for (int c = 1; c <= 3; c++) {
                    const std::string& t = w[c];
                    size_t s1 = t.find('/');
                    unsigned v = obj_index(t.substr(0, s1), lineNo), uv = 0, nn = 0;
                    if (s1 != std::string::npos) {
                        size_t s2 = t.find('/', s1 + 1);
                        std::string ts = t.substr(s1 + 1, s2 == std::string::npos ? std::string::npos : s2 - s1 - 1);
                        if (!ts.empty()) uv = obj_index(ts, lineNo);
                        if (s2 != std::string::npos && s2 + 1 < t.size()) nn = obj_index(t.substr(s2 + 1), lineNo);
                    }
                    idxV.push_back(v);
                    auto key = std::make_tuple(v, nn, uv);
                    auto it = first.find(key);
                    if (it == first.end()) {
                        it = first.emplace(key, (unsigned)corners.size()).first;
                        corners.push_back(key);
                    }
                    idxU.push_back(it->second);
                }
            }
        }
        std::unique_ptr<Mesh> me(new Mesh());
        for (unsigned v : idxV)
            if (v >= pos.size()) throw std::runtime_error("OBJ face refers to a missing vertex");
        if (naive) {
            me->vertices = pos;
            me->indices = idxV;
        } else {
            me->vertices.resize(corners.size());
            if (!tex.empty()) me->uvs.resize(corners.size());
            for (size_t k = 0; k < corners.size(); k++) {
                me->vertices[k] = pos[std::get<0>(corners[k])];
                if (!tex.empty()) {
                    unsigned uv = std::get<2>(corners[k]);
                    if (uv >= tex.size()) throw std::runtime_error("OBJ face refers to a missing uv");
                    me->uvs[k] = tex[uv];
                }
            }
            me->indices = idxU;
        }
        me->compute_normals();
        me->construct_edges();
        me->compute_edge_normals();
        return (void*)me.release();
    });
    return m;
}

scVector2* scGetMeshUVBuffer(scMesh mesh)
{
    return guard_r<scVector2*>(nullptr, [&] {
        Mesh* me = (Mesh*)mesh.ptr;
        if (!me) throw std::runtime_error("null mesh");
        return me->uvs.empty() ? (scVector2*)nullptr : me->uvs.data();
    });
}

int scGetParticlePositions(scModel model, scVector3* out, int count)
{
    return guard_r<int>(-1, [&] {
        ClothModel* m = M(model);
        m->sync_to_host();
        int n = std::min(count, (int)m->pos.size());
        for (int i = 0; i < n; i++) out[i] = scVector3{m->pos[i].x, m->pos[i].y, m->pos[i].z};
        return n;
    });
}

int scSetParticlePositions(scModel model, const scVector3* in, int count)
{
    return guard_r<int>(-1, [&] {
        ClothModel* m = M(model);
        m->sync_to_host();
        int n = std::min(count, (int)m->pos.size());
        for (int i = 0; i < n; i++) m->pos[i] = f3{in[i].x, in[i].y, in[i].z};
        m->mark_host_dirty();
        return n;
    });
}

int scSetParticleInvMasses(scModel model, const float* in, int count)
{
    return guard_r<int>(-1, [&] {
        ClothModel* m = M(model);
        m->sync_to_host();
        int n = std::min(count, (int)m->invMass.size());
        for (int i = 0; i < n; i++) {
            if (!(in[i] >= 0.f)) throw std::runtime_error("inverse masses must be non-negative");
            m->invMass[i] = in[i];
        }
        m->mark_host_dirty();
        return n;
    });
}

static ClothBatch* B(scClothBatch b)
{
    if (!b.ptr) throw std::runtime_error("null cloth batch");
    return (ClothBatch*)b.ptr;
}

scClothBatch scCreateClothBatch(void)
{
    scClothBatch b{nullptr};
    guard([&] { b.ptr = new ClothBatch(); });
    return b;
}

void scFreeClothBatch(scClothBatch batch) { guard([&] { delete (ClothBatch*)batch.ptr; }); }

void scBatchAddPatch(scClothBatch batch, scClothPatch patch) { guard([&] { B(batch)->add(P(patch)); }); }

int scBatchGetNumPatches(scClothBatch batch)
{
    return guard_r<int>(0, [&] { return (int)B(batch)->patches.size(); });
}

void scStepClothBatch(scClothBatch batch, float dt) { guard([&] { B(batch)->step(dt); }); }

void scBatchSetLaunch(scClothBatch batch, int blockSize, int stage)
{
    guard([&] {
        if (blockSize != 0 && blockSize != 256 && blockSize != 512 && blockSize != 1024)
            throw std::runtime_error("block size must be 0, 256, 512 or 1024");
        if (stage < -1 || stage > 1) throw std::runtime_error("stage must be -1, 0 or 1");
        B(batch)->set_launch(blockSize, stage);
    });
}

void scBatchGetLaunch(scClothBatch batch, int* blockSize, int* stage, int* tuned)
{
    guard([&] {
        ClothBatch* b = B(batch);
        if (blockSize) *blockSize = b->blockSize;
        if (stage) *stage = b->stage;
        if (tuned) *tuned = b->tuned ? 1 : 0;
    });
}

int scBatchGetTuning(scClothBatch batch, int* configs, double* ms, int cap)
{
    return guard_r<int>(-1, [&] {
        ClothBatch* b = B(batch);
        const int n = (int)b->trialMs.size();
        for (int i = 0; i < n && i < cap; i++) {
            if (configs) {
                configs[2 * i] = b->trialConfigs[i].first;
                configs[2 * i + 1] = b->trialConfigs[i].second;
            }
            if (ms) ms[i] = b->trialMs[i];
        }
        return n;
    });
}

static WarmStart* WS(scWarmStart w)
{
    if (!w.ptr) throw std::runtime_error("null warm-start network");
    return (WarmStart*)w.ptr;
}

scWarmStart scCreateWarmStart(unsigned int seed)
{
    scWarmStart w{nullptr};
    guard([&] { w.ptr = new WarmStart(seed); });
    return w;
}

void scFreeWarmStart(scWarmStart ws) { guard([&] { delete (WarmStart*)ws.ptr; }); }

int scWarmStartGetParams(scWarmStart ws, float* out, int count)
{
    return guard_r<int>(-1, [&] {
        const std::vector<float> p = WS(ws)->params();
        if (out) memcpy(out, p.data(), (size_t)std::min(count, (int)p.size()) * sizeof(float));
        return (int)p.size();
    });
}

int scWarmStartSetParams(scWarmStart ws, const float* in, int count)
{
    return guard_r<int>(-1, [&] {
        if (count != kMlpParams || !in) throw std::runtime_error("the warm-start network has 5187 parameters");
        WS(ws)->set_params(std::vector<float>(in, in + count));
        return count;
    });
}

void scWarmStartSetLearningRate(scWarmStart ws, float lr)
{
    guard([&] {
        if (!(lr > 0.f) || !std::isfinite(lr)) throw std::runtime_error("learning rate must be positive");
        WS(ws)->set_lr(lr);
    });
}

int scWarmStartGetSteps(scWarmStart ws) { return guard_r<int>(-1, [&] { return WS(ws)->steps(); }); }

int scWarmStartForward(scWarmStart ws, const float* features, int n, float* out)
{
    return guard_r<int>(-1, [&] {
        if (n <= 0 || !features || !out) throw std::runtime_error("no samples");
        WS(ws)->forward_host(features, n, out);
        return n;
    });
}

int scWarmStartForwardNoAlloc(scWarmStart ws, const float* features, int n, float* out, void* stream)
{
    return guard_r<int>(-1, [&] {
        if (!features || !out) throw std::runtime_error("null feature or output pointer");
        WS(ws)->forward_device_noalloc(features, n, out, (hipStream_t)stream);
        return n;
    });
}

float scWarmStartGradient(scWarmStart ws, const float* features, const float* targets, const float* weights, int n,
                          float* grad)
{
    return guard_r<float>(-1.f, [&] {
        if (n <= 0 || !features || !targets) throw std::runtime_error("no samples");
        return WS(ws)->gradient_host(features, targets, weights, n, grad);
    });
}

float scWarmStartTrain(scWarmStart ws, const float* features, const float* targets, const float* weights, int n,
                       int steps)
{
    return guard_r<float>(-1.f, [&] {
        if (n <= 0 || !features || !targets) throw std::runtime_error("no samples");
        if (steps < 1) throw std::runtime_error("steps must be >= 1");
        return WS(ws)->train_host(features, targets, weights, n, steps);
    });
}

void scSetWarmStart(scModel model, scWarmStart ws, int mode, int iterations)
{
    guard([&] { M(model)->set_warm((WarmStart*)ws.ptr, mode, iterations); });
}

void scSetWarmStartFallback(scModel model, float threshold)
{
    guard([&] {
        if (std::isnan(threshold)) throw std::runtime_error("threshold must be a number");
        ClothModel* m = M(model);
        m->mWarmThreshold = threshold;
    });
}

double scGetResidual(scModel model) { return guard_r<double>(-1.0, [&] { return M(model)->residual(); }); }

void scGetResidualTerms(scModel model, double* out) { guard([&] { M(model)->residual_terms(out); }); }

void scGetWarmStartStats(scModel model, double* out)
{
    guard([&] {
        ClothModel* m = M(model);
        out[0] = m->mWarmResidual;
        out[1] = m->mWarmThreshold >= 0.f ? (double)m->mWarmThreshold : m->mWarmPlainResidual;
        out[2] = m->mWarmFallbacks;
        out[3] = m->mWarmPlainResidual;
        out[4] = m->mWarmSteps;
        out[5] = m->mWarmLastPlain ? 1.0 : 0.0;
        out[6] = m->mWarmWarmResidual;
    });
}

int scGetWarmStartData(scModel model, int which, float* out, int count)
{
    return guard_r<int>(-1, [&] { return M(model)->warm_data(which, out, count); });
}

double scGetLastStepTime(scModel model) { return guard_r<double>(-1.0, [&] { return M(model)->mLastStepMs; }); }

void scSaveState(scModel model) { guard([&] { M(model)->save_state(); }); }

void scRestoreState(scModel model) { guard([&] { M(model)->restore_state(); }); }

void scBatchSetUseGraph(scClothBatch batch, bool enable) { guard([&] { B(batch)->useGraph = enable; }); }

void scBatchRecordWorkgroupTimes(scClothBatch batch, bool enable)
{
    guard([&] { B(batch)->recordClock = enable; });
}

int scBatchGetWorkgroupTimes(scClothBatch batch, double* startUs, double* endUs, long long* unit, int cap)
{
    return guard_r<int>(-1, [&] {
        ClothBatch* b = B(batch);
        const int n = (int)b->clockStart.size();
        if (n == 0) return 0;
        int dev = 0, kHz = 0;
        HIP_OK(hipGetDevice(&dev));
        HIP_OK(hipDeviceGetAttribute(&kHz, hipDeviceAttributeWallClockRate, dev));
        if (kHz <= 0) throw std::runtime_error("device reports no wall clock rate");
        const unsigned long long t0 = *std::min_element(b->clockStart.begin(), b->clockStart.end());
        const double usPerTick = 1000.0 / kHz;
        for (int k = 0; k < std::min(n, cap); k++) {
            if (startUs) startUs[k] = (double)(b->clockStart[k] - t0) * usPerTick;
            if (endUs) endUs[k] = (double)(b->clockEnd[k] - t0) * usPerTick;
            if (unit) unit[k] = (long long)b->clockUnit[k];
        }
        return n;
    });
}

double scBatchGetStepTime(scClothBatch batch, int* launches)
{
    return guard_r<double>(-1.0, [&] {
        ClothBatch* b = B(batch);
        if (launches) *launches = b->lastSubsteps;
        return b->lastStepMs;
    });
}

int scGetParticleVelocities(scModel model, scVector3* out, int count)
{
    return guard_r<int>(-1, [&] {
        ClothModel* m = M(model);
        m->sync_to_host();
        int n = std::min(count, (int)m->vel.size());
        for (int i = 0; i < n; i++) out[i] = scVector3{m->vel[i].x, m->vel[i].y, m->vel[i].z};
        return n;
    });
}

int scSetParticleVelocities(scModel model, const scVector3* in, int count)
{
    return guard_r<int>(-1, [&] {
        ClothModel* m = M(model);
        m->sync_to_host();
        int n = std::min(count, (int)m->vel.size());
        for (int i = 0; i < n; i++) m->vel[i] = f3{in[i].x, in[i].y, in[i].z};
        m->mark_host_dirty();
        return n;
    });
}

int scGetNumBends(scModel model) { return guard_r<int>(0, [&] { return (int)M(model)->bends.size(); }); }

int scGetBends(scModel model, int* out, int count)
{
    return guard_r<int>(-1, [&] {
        ClothModel* m = M(model);
        int n = std::min(count, (int)m->bends.size());
        for (int i = 0; i < n; i++) {
            out[4 * i] = (int)m->bends[i].i1;
            out[4 * i + 1] = (int)m->bends[i].i2;
            out[4 * i + 2] = (int)m->bends[i].i3;
            out[4 * i + 3] = (int)m->bends[i].i4;
        }
        return n;
    });
}

// kind 0: FEM triangles, 1: bends, 2: obstacle-edge contacts (by cloth
// edge), 3: obstacle-vertex contacts (by cloth triangle), 4: links
static const Schedule& schedule_of(ClothModel* m, int kind)
{
    switch (kind) {
    case 0: return m->triCol;
    case 1: return m->bendCol;
    case 2: return m->edgeContactCol;
    case 3: return m->triContactCol;
    case 4: return m->linkCol;
    default: throw std::runtime_error("unknown constraint kind " + std::to_string(kind));
    }
}

int scGetConstraintOrder(scModel model, int kind, int* order, int count)
{
    return guard_r<int>(-1, [&] {
        ClothModel* m = M(model);
        m->ensure_colouring();
        const Schedule& c = schedule_of(m, kind);
        int n = std::min(count, (int)c.order.size());
        if (order) memcpy(order, c.order.data(), n * sizeof(int));
        return (int)c.order.size();
    });
}

int scGetColouring(scModel model, int kind, int* colourOffsets, int maxColours)
{
    return guard_r<int>(-1, [&] {
        ClothModel* m = M(model);
        m->ensure_colouring();
        const Schedule& c = schedule_of(m, kind);
        int nc = c.numPhases();
        if (nc < 0) nc = 0;
        if (colourOffsets)
            for (int k = 0; k <= nc && k <= maxColours; k++) colourOffsets[k] = c.phaseOffsets[k];
        return nc;
    });
}

int scGetSchedule(scModel model, int kind, int* phase, int* slab, int* colour, int count)
{
    return guard_r<int>(-1, [&] {
        ClothModel* m = M(model);
        m->ensure_colouring();
        const Schedule& c = schedule_of(m, kind);
        int n = std::min(count, (int)c.order.size());
        if (phase) memcpy(phase, c.phaseOf.data(), n * sizeof(int));
        if (slab) memcpy(slab, c.slabOf.data(), n * sizeof(int));
        if (colour) memcpy(colour, c.colourOf.data(), n * sizeof(int));
        return (int)c.order.size();
    });
}

int scGetTileSchedule(scModel model, int* counts, int* tileVertBegin, int* tileVerts, int* owned,
                      int* tileColOffsets, int* instances)
{
    return guard_r<int>(-1, [&] {
        ClothModel* m = M(model);
        m->ensure_colouring();
        const Schedule& c = m->triCol;
        if (!c.tiled) {
            if (counts) counts[0] = counts[1] = counts[2] = counts[3] = 0;
            return 0;
        }
        const int nt = (int)c.slabVertBegin.size() - 1;
        if (counts) {
            counts[0] = nt;
            counts[1] = (int)c.slabVerts.size();
            counts[2] = c.maxSlabColours;
            counts[3] = (int)c.instCons.size();
        }
        if (tileVertBegin) memcpy(tileVertBegin, c.slabVertBegin.data(), c.slabVertBegin.size() * sizeof(int));
        for (size_t i = 0; i < c.slabVerts.size(); i++) {
            if (tileVerts) tileVerts[i] = (int)c.slabVerts[i];
            if (owned) owned[i] = c.slabOwned[i];
        }
        if (tileColOffsets) memcpy(tileColOffsets, c.colOffsets.data(), c.colOffsets.size() * sizeof(int));
        if (instances) memcpy(instances, c.instCons.data(), c.instCons.size() * sizeof(int));
        return nt;
    });
}

int scGetContactFeatures(scModel model, int* out, int count)
{
    return guard_r<int>(-1, [&] {
        ClothModel* m = M(model);
        m->download_contacts();
        const int n = std::min(count, (int)m->contactsHost.size());
        for (int i = 0; i < n && out; i++) out[i] = m->contactsHost[i].feature;
        return (int)m->contactsHost.size();
    });
}

static int feature_contacts(const std::vector<FeatureContactOut>& v, int nw, int* ids, float* data, int cap)
{
    const int n = std::min(cap, (int)v.size());
    const int stride = 6 + nw + 2;
    for (int i = 0; i < n; i++) {
        const FeatureContactOut& c = v[i];
        if (ids) {
            ids[3 * i] = c.id;
            ids[3 * i + 1] = c.feature;
            ids[3 * i + 2] = c.coll;
        }
        if (data) {
            float* d = data + (size_t)stride * i;
            d[0] = c.point.x; d[1] = c.point.y; d[2] = c.point.z;
            d[3] = c.normal.x; d[4] = c.normal.y; d[5] = c.normal.z;
            for (int k = 0; k < nw; k++) d[6 + k] = c.w[k];
            d[6 + nw] = c.lambda;
            d[7 + nw] = c.lambdaF;
        }
    }
    return (int)v.size();
}

int scGetSelfContacts(scModel model, int kind, unsigned long long* keys, int* ids, float* data, int cap)
{
    return guard_r<int>(-1, [&] {
        if (kind != 0 && kind != 1) throw std::runtime_error("self contact kind must be 0 or 1");
        ClothModel* m = M(model);
        m->download_contacts();
        const auto& v = m->selfContactsHost[kind];
        const int n = std::min(cap, (int)v.size());
        for (int i = 0; i < n; i++) {
            if (keys) keys[i] = v[i].key;
            if (ids)
                for (int k = 0; k < 4; k++) ids[4 * i + k] = v[i].ids[k];
            if (data) {
                float* d = data + 8 * (size_t)i;
                d[0] = v[i].w[0]; d[1] = v[i].w[1]; d[2] = v[i].w[2];
                d[3] = v[i].normal[0]; d[4] = v[i].normal[1]; d[5] = v[i].normal[2];
                d[6] = v[i].side; d[7] = v[i].lambda;
            }
        }
        return (int)v.size();
    });
}

void scSetSelfParams(scModel model, float shrink, float threshold)
{
    guard([&] {
        ClothModel* m = M(model);
        m->mSelfShrink = shrink;
        m->mSelfCollThreshold = threshold;
    });
}

void scSetSelfCCD(scModel model, float volume, float bisect, float vtDistance)
{
    guard([&] {
        if (!(volume >= 0.f) || !(bisect >= 0.f) || !(vtDistance >= 0.f))
            throw std::runtime_error("self-collision thresholds must be non-negative");
        ClothModel* m = M(model);
        m->mCcdVolume = volume;
        m->mCcdBisect = bisect;
        m->mCcdVtDistance = vtDistance;
    });
}

void scSetSelfCapacity(scModel model, int capacity)
{
    guard([&] {
        if (capacity < 0) throw std::runtime_error("negative self-contact capacity");
        M(model)->mSelfCapacity = capacity;
    });
}

long long scGetSelfCandidates(scModel model, int kind, int* pairs, long long cap)
{
    return guard_r<long long>(-1, [&] {
        if (kind != 0 && kind != 1) throw std::runtime_error("candidate kind must be 0 (vertex-triangle) or 1 (edge-edge)");
        const auto v = M(model)->self_candidates(kind);
        for (long long k = 0; k < (long long)v.size() && k < cap; k++) {
            pairs[2 * k] = v[k].first;
            pairs[2 * k + 1] = v[k].second;
        }
        return (long long)v.size();
    });
}

void scSetSelfEdgeWeights(scModel model, int mode)
{
    guard([&] {
        if (mode != 0 && mode != 1) throw std::runtime_error("self edge-edge weight mode must be 0 or 1");
        M(model)->mSelfEdgeWeights = mode;
    });
}

int scGetSelfStats(scModel model, int* out, int count)
{
    return guard_r<int>(-1, [&] {
        ClothModel* m = M(model);
        const int v[9] = {m->gpu.self.cap, m->selfReruns, m->lastSelfPeak[0], m->lastSelfPeak[1],
                          m->gpu.self.candCapVT + m->gpu.self.candCapEE, m->lastSelfColours[0],
                          m->lastSelfColours[1], m->lastSelfUniq[0], m->lastSelfUniq[1]};
        for (int k = 0; k < std::min(count, 9); k++) out[k] = v[k];
        return 9;
    });
}

int scGetEdgeContactData(scModel model, int* ids, float* data, int cap)
{
    return guard_r<int>(-1, [&] {
        ClothModel* m = M(model);
        m->download_contacts();
        return feature_contacts(m->edgeContactsHost, 2, ids, data, cap);
    });
}

int scGetTriContactData(scModel model, int* ids, float* data, int cap)
{
    return guard_r<int>(-1, [&] {
        ClothModel* m = M(model);
        m->download_contacts();
        return feature_contacts(m->triContactsHost, 3, ids, data, cap);
    });
}

void scSetSweep(scModel model, int mode, int slabSize) { guard([&] { M(model)->set_sweep(mode, slabSize); }); }

void scGetSolverStats(scModel model, float* out)
{
    guard([&] {
        ClothModel* m = M(model);
        m->fetch_stats();
        out[0] = (float)m->statIterations;
        out[1] = m->statBendError;
        out[2] = m->statContactError;
        out[3] = m->statSelfVTError;
        out[4] = m->statSelfEEError;
        out[5] = m->statLinkError;
    });
}

int scGetLinks(scModel model, int* ids, float* data, int count)
{
    return guard_r<int>(-1, [&] {
        ClothModel* m = M(model);
        const int n = (int)m->links.size();
        for (int k = 0; k < n && k < count; k++) {
            if (ids) { ids[2 * k] = (int)m->links[k].i1; ids[2 * k + 1] = (int)m->links[k].i2; }
            if (data) { data[2 * k] = m->links[k].len; data[2 * k + 1] = m->links[k].stiffness; }
        }
        return n;
    });
}

void scEnableKernelTiming(scModel model, bool enable)
{
    guard([&] {
        ClothModel* m = M(model);
        m->gpu.timer.enabled = enable;
        m->gpu.timer.reset();
    });
}

int scGetKernelTiming(scModel model, int which, int* launches, double* totalMs)
{
    return guard_r<int>(-1, [&] {
        ClothModel* m = M(model);
        if (which < 0 || which > 4) throw std::runtime_error("kernel id out of range");
        if (launches) *launches = m->gpu.timer.launches[which];
        if (totalMs) *totalMs = m->gpu.timer.totalMs[which];
        return 0;
    });
}

void scSetUseGraph(scModel model, bool enable) { guard([&] { M(model)->mUseGraph = enable; }); }

double scBandwidthProbe(long long bytes, int reps)
{
    return guard_r<double>(-1.0, [&] {
        if (bytes < 16 || reps < 1) throw std::runtime_error("bad probe size");
        const long n = (long)(bytes / 16);
        float4 *a = nullptr, *b = nullptr;
        hipStream_t s = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        double ms = -1.0;
        auto cleanup = [&] {
            if (a) (void)hipFree(a);
            if (b) (void)hipFree(b);
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            if (s) (void)hipStreamDestroy(s);
        };
        try {
            auto ok = [](hipError_t e) {
                if (e != hipSuccess) throw std::runtime_error(hipGetErrorString(e));
            };
            ok(hipMalloc((void**)&a, n * sizeof(float4)));
            ok(hipMalloc((void**)&b, n * sizeof(float4)));
            ok(hipMemset(a, 0, n * sizeof(float4)));
            ok(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
            ok(hipEventCreate(&e0));
            ok(hipEventCreate(&e1));
            launch_copy(s, a, b, n);  // warm-up
            ok(hipEventRecord(e0, s));
            for (int r = 0; r < reps; r++) launch_copy(s, a, b, n);
            ok(hipEventRecord(e1, s));
            ok(hipEventSynchronize(e1));
            ok(hipGetLastError());
            float t = 0.f;
            ok(hipEventElapsedTime(&t, e0, e1));
            ms = t / reps;
        } catch (...) {
            cleanup();
            throw;
        }
        cleanup();
        return ms;
    });
}

int scLbvhSelfCheck(int n, unsigned seed, int refits, double* msOut)
{
    return guard_r<int>(-1, [&] {
        if (n < 1 || refits < 0) throw std::runtime_error("bad LBVH check size");
        // random boxes from a 32-bit LCG: clustered centres, varied extents
        uint32_t st = seed * 2654435761u + 1u;
        auto rnd = [&] {
            st = st * 1664525u + 1013904223u;
            return (float)(st >> 8) * (1.0f / 16777216.0f);
        };
        std::vector<float4> lo(n), hi(n);
        auto fill = [&](float jitter) {
            for (int i = 0; i < n; i++) {
                const float cx = rnd() * rnd(), cy = rnd(), cz = (i % 7) * 0.1f + jitter * rnd();
                const float e = 0.001f + 0.01f * rnd();
                lo[i] = make_float4(cx - e, cy - e, cz - e, 0.f);
                hi[i] = make_float4(cx + e, cy + e, cz + e, 0.f);
            }
        };
        float4 *dlo = nullptr, *dhi = nullptr;
        hipStream_t s = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        int bad = 0;
        auto cleanup = [&] {
            if (dlo) (void)hipFree(dlo);
            if (dhi) (void)hipFree(dhi);
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            if (s) (void)hipStreamDestroy(s);
        };
        try {
            auto ok = [](hipError_t e) {
                if (e != hipSuccess) throw std::runtime_error(hipGetErrorString(e));
            };
            ok(hipMalloc((void**)&dlo, n * sizeof(float4)));
            ok(hipMalloc((void**)&dhi, n * sizeof(float4)));
            ok(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
            ok(hipEventCreate(&e0));
            ok(hipEventCreate(&e1));
            Lbvh bvh;
            fill(0.f);
            ok(hipMemcpy(dlo, lo.data(), n * sizeof(float4), hipMemcpyHostToDevice));
            ok(hipMemcpy(dhi, hi.data(), n * sizeof(float4), hipMemcpyHostToDevice));
            bvh.build(s, dlo, dhi, n);  // allocates; the timed build below reuses the buffers
            ok(hipEventRecord(e0, s));
            bvh.build(s, dlo, dhi, n);
            ok(hipEventRecord(e1, s));
            ok(hipEventSynchronize(e1));
            float t = 0.f;
            ok(hipEventElapsedTime(&t, e0, e1));
            if (msOut) msOut[0] = t;
            bad += bvh.verify(s, dlo, dhi);
            double refitMs = 0.0;
            for (int r = 0; r < refits; r++) {
                fill(0.05f);
                ok(hipMemcpy(dlo, lo.data(), n * sizeof(float4), hipMemcpyHostToDevice));
                ok(hipMemcpy(dhi, hi.data(), n * sizeof(float4), hipMemcpyHostToDevice));
                ok(hipEventRecord(e0, s));
                bvh.refit(s, dlo, dhi);
                ok(hipEventRecord(e1, s));
                ok(hipEventSynchronize(e1));
                ok(hipGetLastError());
                ok(hipEventElapsedTime(&t, e0, e1));
                refitMs += t;
                bad += bvh.verify(s, dlo, dhi);
            }
            if (msOut) msOut[1] = refits ? refitMs / refits : 0.0;
        } catch (...) {
            cleanup();
            throw;
        }
        cleanup();
        return bad;
    });
}

const char* scGetLastError(void) { return g_lastError.c_str(); }

int scGetErrorCount(void) { return g_errorCount; }

const char* scGetBuildInfo(void) { return "solidcloth-mi355x gfx950 fp32 coloured-GS"; }

}  // extern "C"
