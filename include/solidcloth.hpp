/* C++ classes over the solidcloth C ABI, with the surface the reference's
 * own programs use: Physics::ClothPatch (Physics/ClothPatch.h:24-72),
 * Physics::ClothModel (Physics/ClothModel.h:42-146), Physics::ClothModelPBD
 * (Physics/ClothPBD.h:6-22), Physics::CollisionWorld and the collidable
 * descriptions of Physics/Common.h:27-139. Header only; link against
 * libsolidcloth.so.
 *
 * Usage follows the reference demo (Demo/ClothDemo.cpp:100-112, 173-176,
 * 424-427):
 *
 *     Physics::CollisionWorld world;
 *     world.AddCollidable(std::make_shared<Physics::Sphere>(Vector3(0, 0, 0), 0.3f));
 *     Physics::ClothPatch cloth(world);
 *     cloth.GetModel().SetCollisionFlags(Physics::CF_VERTICES);
 *     cloth.Init(64, 64, 2.f / 63, Vector3(0, 0.5f, 0), true, false);
 *     cloth.Step(1.f / 60);
 *
 * Differences from the reference:
 *  - Particles, triangles and contacts are returned by value (the state
 *    lives on the GPU); GetPositions / SetPositions move all positions in
 *    one copy.
 *  - Calls the library refuses (scGetLastError) throw solidcloth::Error; the
 *    reference reports no errors.
 *  - A collidable added to a world is created in the library from its
 *    description at that moment; a CollisionMesh or CollisionSDF refers to
 *    its Mesh or SDF, which must outlive the world, as in the reference.
 *  - ClothModelPBD::Step(h) advances one substep through its patch (the
 *    solver state belongs to the patch on the GPU).
 * Define SOLIDCLOTH_NO_PHYSICS_NAMESPACE to keep the classes out of
 * namespace Physics (they always live in namespace solidcloth). */
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "solidcloth.h"

namespace solidcloth {

class Error : public std::runtime_error {
public:
    using std::runtime_error::runtime_error;
};

namespace detail {
// Runs a C call; throws Error when it recorded an error on this thread.
template <class F>
auto checked(F&& f) -> decltype(f())
{
    const int before = scGetErrorCount();
    if constexpr (std::is_void_v<decltype(f())>) {
        f();
        if (scGetErrorCount() != before) throw Error(scGetLastError());
    } else {
        auto r = f();
        if (scGetErrorCount() != before) throw Error(scGetLastError());
        return r;
    }
}
}  // namespace detail

struct Vector3 {
    float x = 0.f, y = 0.f, z = 0.f;
    Vector3() = default;
    Vector3(float x_, float y_, float z_) : x(x_), y(y_), z(z_) {}
    Vector3(scVector3 v) : x(v.x), y(v.y), z(v.z) {}
    operator scVector3() const { return scVector3{x, y, z}; }
};

struct Vector2 {
    float x = 0.f, y = 0.f;
    Vector2() = default;
    Vector2(float x_, float y_) : x(x_), y(y_) {}
    Vector2(scVector2 v) : x(v.x), y(v.y) {}
    operator scVector2() const { return scVector2{x, y}; }
};

struct Aabb3 {
    Vector3 min, max;
    Aabb3() = default;
    Aabb3(Vector3 a, Vector3 b) : min(a), max(b) {}
    operator scAabb3() const { return scAabb3{min, max}; }
};

struct Quaternion {
    float s = 1.f;
    Vector3 v;
    operator scQuaternion() const { return scQuaternion{s, v}; }
};

// Geometry/Collision3D.h:12-19
enum CollisionFlags { CF_VERTICES = 1, CF_TRIANGLES = 2, CF_SELF = 4, CF_WALLS = 8, CF_EDGES = 16 };
// Physics/ClothModel.h:24-31
enum SolverType { SOLVER_JACOBI, SOLVER_GAUSS_SEIDEL, SOLVER_JACOBI_CR, SOLVER_IMPLICIT, SOLVER_EXPLICIT };
// Physics/ClothPatch.h:19-22
enum MethodType { METHOD_PBD };
// Physics/Common.h:17-27
enum CollidableType { CT_INVALID, CT_WALLS, CT_SLAB, CT_SPHERE, CT_MESH, CT_BOX, CT_CAPSULE, CT_SDF };

// A triangle mesh owned by this object (or, for a cloth's render mesh, a view
// that its patch owns).
class Mesh {
public:
    Mesh() = default;
    static Mesh FromArrays(const std::vector<Vector3>& vertices, const std::vector<unsigned>& indices)
    {
        std::vector<scVector3> v(vertices.begin(), vertices.end());
        return Mesh(detail::checked([&] {
                        return scCreateMeshFromArrays(v.data(), (int)v.size(), indices.data(), (int)indices.size());
                    }),
                    true);
    }
    static Mesh LoadObj(const std::string& path, bool naive = false)
    {
        return Mesh(detail::checked([&] { return scLoadMeshFromObj(path.c_str(), naive); }), true);
    }
    ~Mesh() { release(); }
    Mesh(Mesh&& o) noexcept : h_(o.h_), owned_(o.owned_) { o.h_.ptr = nullptr; }
    Mesh& operator=(Mesh&& o) noexcept
    {
        if (this != &o) {
            release();
            h_ = o.h_;
            owned_ = o.owned_;
            o.h_.ptr = nullptr;
        }
        return *this;
    }
    Mesh(const Mesh&) = delete;
    Mesh& operator=(const Mesh&) = delete;

    int GetNumVertices() const { return detail::checked([&] { return scGetMeshVertexCount(h_); }); }
    int GetNumIndices() const { return detail::checked([&] { return scGetMeshIndexCount(h_); }); }
    const scVector3* Vertices() const { return detail::checked([&] { return scGetMeshVertexBuffer(h_); }); }
    const scVector3* Normals() const { return detail::checked([&] { return scGetMeshNormalBuffer(h_); }); }
    const unsigned* Indices() const { return detail::checked([&] { return scGetMeshIndexBuffer(h_); }); }
    const scVector2* UVs() const { return detail::checked([&] { return scGetMeshUVBuffer(h_); }); }
    void SetUVs(const std::vector<Vector2>& uv)
    {
        std::vector<scVector2> a(uv.begin(), uv.end());
        detail::checked([&] { return scSetMeshUVs(h_, a.data(), (int)a.size()); });
    }
    scMesh handle() const { return h_; }

private:
    friend class ClothPatch;
    Mesh(scMesh h, bool owned) : h_(h), owned_(owned) {}
    void release()
    {
        if (h_.ptr && owned_) scFreeMesh(h_);
        h_.ptr = nullptr;
    }
    scMesh h_{nullptr};
    bool owned_ = false;
};

// A signed distance grid (Geometry/SDF.h), owned by this object.
class SDF {
public:
    // values: (nx + 1) (ny + 1) (nz + 1) node values, x fastest
    SDF(const std::vector<float>& values, int nx, int ny, int nz, Aabb3 box)
        : h_(detail::checked([&] { return scCreateGridSDF(values.data(), nx, ny, nz, box); }))
    {
    }
    static SDF Load(const std::string& path) { return SDF(detail::checked([&] { return scLoadSDFFile(path.c_str()); })); }
    void Save(const std::string& path) const { detail::checked([&] { return scWriteSDFFile(h_, path.c_str()); }); }
    ~SDF()
    {
        if (h_.ptr) scFreeSDF(h_);
    }
    SDF(SDF&& o) noexcept : h_(o.h_) { o.h_.ptr = nullptr; }
    SDF& operator=(SDF&& o) noexcept
    {
        if (this != &o) {
            if (h_.ptr) scFreeSDF(h_);
            h_ = o.h_;
            o.h_.ptr = nullptr;
        }
        return *this;
    }
    SDF(const SDF&) = delete;
    SDF& operator=(const SDF&) = delete;
    scSDF handle() const { return h_; }

private:
    explicit SDF(scSDF h) : h_(h) {}
    scSDF h_{nullptr};
};

// Collidable descriptions (Physics/Common.h:27-139): plain data, turned into
// library objects when added to a world.
struct Collidable {
    int mType = CT_INVALID;
    Vector3 center;
    Quaternion rot;
    virtual ~Collidable() = default;
    // the library object for this description (the world takes it over)
    virtual scHandle Create() const = 0;
};

struct Walls : Collidable {
    Aabb3 mBox;
    Walls() { mType = CT_WALLS; }
    explicit Walls(const Aabb3& box) : mBox(box) { mType = CT_WALLS; }
    scHandle Create() const override { return scCreateWalls(mBox); }
};

struct Sphere : Collidable {
    float radius = 0.f;
    Sphere() { mType = CT_SPHERE; }
    Sphere(Vector3 c, float r) : radius(r)
    {
        mType = CT_SPHERE;
        center = c;
    }
    scHandle Create() const override { return scCreateSphere(center, radius); }
};

struct Capsule : Collidable {
    float hh = 0.f;  // half height along the local y axis
    float r = 0.f;   // cap radius
    Capsule() { mType = CT_CAPSULE; }
    Capsule(Vector3 c, float rad, float halfHeight, Quaternion q = Quaternion()) : hh(halfHeight), r(rad)
    {
        mType = CT_CAPSULE;
        center = c;
        rot = q;
    }
    scHandle Create() const override { return scCreateCapsule(center, r, hh, rot); }
};

struct CollisionMesh : Collidable {
    const Mesh* mesh;
    explicit CollisionMesh(const Mesh* m) : mesh(m) { mType = CT_MESH; }
    scHandle Create() const override { return scCreateMesh(mesh->handle()); }
};

struct CollisionSDF : Collidable {
    const SDF* sdf;
    explicit CollisionSDF(const SDF* s) : sdf(s) { mType = CT_SDF; }
    scHandle Create() const override { return scCreateSDF(sdf->handle()); }
};

// Physics/CollisionWorld: the collidables a patch is tested against, in order.
class CollisionWorld {
public:
    CollisionWorld() : h_(detail::checked([] { return scCreateCollisionWorld(); })) {}
    ~CollisionWorld()
    {
        if (h_.ptr) scFreeCollisionWorld(h_);
    }
    CollisionWorld(const CollisionWorld&) = delete;
    CollisionWorld& operator=(const CollisionWorld&) = delete;

    void AddCollidable(const std::shared_ptr<Collidable>& c)
    {
        const scHandle o = detail::checked([&] { return c->Create(); });
        detail::checked([&] { scAddCollidable(h_, o); });
        descs_.push_back(c);
    }
    size_t GetNumCollidables() const { return (size_t)detail::checked([&] { return scGetNumCollidables(h_); }); }
    // the description added i-th (null for collidables added through the C ABI)
    const Collidable* GetCollidable(size_t i) const { return i < descs_.size() ? descs_[i].get() : nullptr; }
    void ClearCollidables()
    {
        detail::checked([&] { scClearCollidables(h_); });
        descs_.clear();
    }
    scCollisionWorld handle() const { return h_; }

private:
    scCollisionWorld h_;
    std::vector<std::shared_ptr<Collidable>> descs_;
};

// Physics/ClothTypes.h: the parts of a particle and a triangle the model
// exposes, by value.
struct Particle {
    Vector3 pos;
    float invMass = 0.f;
};

struct Triangle {
    int i1 = 0, i2 = 0, i3 = 0;
    float invDet = 0.f, du1 = 0.f, du2 = 0.f, dv1 = 0.f, dv2 = 0.f;
};

struct Contact {
    Vector3 point, normal;
};

// Physics/ClothModel.h:32-40, with the library's vertex-contact error
struct SolverStats {
    int iterations = 0;
    float linksError = 0.f;
    float bendError = 0.f;
    float contactError = 0.f;
    float selfVTError = 0.f;
    float selfEEError = 0.f;
};

class ClothPatch;

// Physics/ClothModel.h:42-146: a view of a patch's model.
class ClothModel {
public:
    virtual ~ClothModel() = default;
    ClothModel(const ClothModel&) = delete;
    ClothModel& operator=(const ClothModel&) = delete;

    size_t GetNumParticles() const { return (size_t)call([&] { return scGetNumParticles(h_); }); }
    Particle GetParticle(size_t i) const
    {
        Particle p;
        scVector3 v;
        call([&] { scGetParticlePos(h_, (int)i, &v); });
        p.pos = v;
        p.invMass = call([&] { return scGetParticleInvMass(h_, (int)i); });
        return p;
    }
    std::vector<Vector3> GetPositions() const
    {
        std::vector<scVector3> a(GetNumParticles());
        call([&] { return scGetParticlePositions(h_, a.data(), (int)a.size()); });
        return std::vector<Vector3>(a.begin(), a.end());
    }
    // Physics/ClothModel.h:98
    void SetPositions(const Vector3* positions)
    {
        std::vector<scVector3> a(positions, positions + GetNumParticles());
        call([&] { return scSetParticlePositions(h_, a.data(), (int)a.size()); });
    }
    // inverse masses (0 pins a particle)
    void SetInvMasses(const float* invMasses)
    {
        call([&] { return scSetParticleInvMasses(h_, invMasses, (int)GetNumParticles()); });
    }

    size_t GetNumTris() const { return (size_t)call([&] { return scGetNumTris(h_); }); }
    Triangle GetTriangle(size_t i) const
    {
        Triangle t;
        scVector2 du, dv, s;
        call([&] { scGetTriangle(h_, (int)i, &t.i1, &t.i2, &t.i3); });
        call([&] { scGetTriangleInfo(h_, (int)i, &t.invDet, &du, &dv, &s); });
        t.du1 = du.x;
        t.du2 = du.y;
        t.dv1 = dv.x;
        t.dv2 = dv.y;
        return t;
    }
    size_t GetNumQuads() const { return (size_t)call([&] { return scGetNumQuads(h_); }); }
    size_t GetNumLinks() const { return (size_t)call([&] { return scGetNumLinks(h_); }); }
    size_t GetNumBends() const { return (size_t)call([&] { return scGetNumBends(h_); }); }

    size_t GetNumContacts() const { return (size_t)call([&] { return scGetNumContacts(h_); }); }
    Contact GetContact(size_t i) const
    {
        scVector3 p, n;
        call([&] { scGetContact(h_, (int)i, &p, &n); });
        return Contact{p, n};
    }
    size_t GetNumTriContacts() const { return (size_t)call([&] { return scGetNumTriContacts(h_); }); }
    Contact GetTriContact(size_t i) const
    {
        scVector3 p, n;
        call([&] { scGetTriContact(h_, (int)i, &p, &n); });
        return Contact{p, n};
    }

    float GetStretchStiffness() const { return call([&] { return scGetStretchStiffness(h_); }); }
    void SetStretchStiffness(float v) { call([&] { scSetStretchStiffness(h_, v); }); }
    float GetShearStiffness() const { return call([&] { return scGetShearStiffness(h_); }); }
    void SetShearStiffness(float v) { call([&] { scSetShearStiffness(h_, v); }); }
    float GetBendStiffness() const { return call([&] { return scGetBendStiffness(h_); }); }
    void SetBendStiffness(float v) { call([&] { scSetBendStiffness(h_, v); }); }
    void SetThickness(float v) { call([&] { scSetThickness(h_, v); }); }
    float GetThickness() const { return call([&] { return scGetThickness(h_); }); }
    void SetTolerance(float v) { call([&] { scSetTolerance(h_, v); }); }
    float GetTolerance() const { return call([&] { return scGetTolerance(h_); }); }
    void SetNumIterations(int v) { call([&] { scSetNumIterations(h_, v); }); }
    int GetNumIterations() const { return call([&] { return scGetNumIterations(h_); }); }
    bool GetDihedral() const { return call([&] { return scGetDihedral(h_); }); }
    void SetDihedral(bool v) { call([&] { scSetDihedral(h_, v); }); }
    bool GetFEM() const { return call([&] { return scGetUseFEM(h_); }); }
    void SetFEM(bool v) { call([&] { scSetUseFEM(h_, v); }); }
    float GetFriction() const { return call([&] { return scGetFriction(h_); }); }
    void SetFriction(float v) { call([&] { scSetFriction(h_, v); }); }
    float GetPoisson() const { return call([&] { return scGetPoisson(h_); }); }
    void SetPoisson(float v) { call([&] { scSetPoisson(h_, v); }); }
    float GetUnit() const { return call([&] { return scGetUnit(h_); }); }
    void SetUnit(float v) { call([&] { scSetUnit(h_, v); }); }
    SolverType GetSolver() const { return (SolverType)call([&] { return scGetSolver(h_); }); }
    void SetSolver(SolverType v) { call([&] { scSetSolver(h_, (int)v); }); }
    void SetCollisionFlags(int v) { call([&] { scSetCollisionFlags(h_, v); }); }
    int GetCollisionFlags() const { return call([&] { return scGetCollisionFlags(h_); }); }
    float GetSelfShrink() const { return self_params().first; }
    void SetSelfShrink(float v) { call([&] { scSetSelfParams(h_, v, GetSelfThreshold()); }); }
    float GetSelfThreshold() const { return self_params().second; }
    float GetDragCoeff() const { return call([&] { return scGetDragCoeff(h_); }); }
    void SetDragCoeff(float v) { call([&] { scSetDragCoeff(h_, v); }); }

    // Physics/ClothModel.h:150-152
    void AddMouseSpring(int i1, int i2, int i3, Vector3 coords, Vector3 p)
    {
        call([&] { scAddMouseSpring(h_, i1, i2, i3, coords, p); });
    }
    void RemoveMouseSpring() { call([&] { scRemoveMouseSpring(h_); }); }
    void UpdateMouseSpring(Vector3 p) { call([&] { scUpdateMouseSpring(h_, p); }); }

    SolverStats GetStats() const
    {
        float o[6] = {0, 0, 0, 0, 0, 0};
        call([&] { scGetSolverStats(h_, o); });
        SolverStats st;
        st.iterations = (int)o[0];
        st.bendError = o[1];
        st.contactError = o[2];
        st.selfVTError = o[3];
        st.selfEEError = o[4];
        st.linksError = o[5];
        return st;
    }

    // one substep of length h (Physics/ClothModel.h:107)
    virtual void Step(float h) = 0;
    scModel handle() const { return h_; }

protected:
    ClothModel(scModel h, ClothPatch* owner) : h_(h), owner_(owner) {}
    template <class F>
    static auto call(F&& f) -> decltype(f())
    {
        return detail::checked(std::forward<F>(f));
    }
    std::pair<float, float> self_params() const
    {
        float s = 0.f, t = 0.f;
        call([&] { scGetSelfParams(h_, &s, &t); });
        return {s, t};
    }
    scModel h_;
    ClothPatch* owner_;
};

// Physics/ClothPBD.h:6-22: the position-based model every patch uses.
class ClothModelPBD : public ClothModel {
public:
    void Step(float h) override;

private:
    friend class ClothPatch;
    ClothModelPBD(scModel h, ClothPatch* owner) : ClothModel(h, owner) {}
};

// Physics/ClothPatch.h:24-72
class ClothPatch {
public:
    explicit ClothPatch(const CollisionWorld& world)
        : h_(detail::checked([&] { return scCreateClothPatch(world.handle()); })), world_(&world),
          model_(detail::checked([&] { return scGetClothModel(h_); }), this)
    {
    }
    ~ClothPatch()
    {
        if (h_.ptr) scFreeClothPatch(h_);
    }
    ClothPatch(const ClothPatch&) = delete;
    ClothPatch& operator=(const ClothPatch&) = delete;

    // quad sim-mesh (Physics/ClothPatch.cpp:125-245)
    void Init(int divisionsX, int divisionsY, float inc, Vector3 offset, bool horizontal, bool attached)
    {
        detail::checked([&] { scInitCloth(h_, divisionsX, divisionsY, inc, offset, horizontal, attached); });
    }
    // triangle sim-mesh (Physics/ClothPatch.cpp:40-123)
    void Init(const Mesh& mesh, const Vector3& offset, bool attached)
    {
        detail::checked([&] { scInitClothFromMesh(h_, mesh.handle(), offset, attached); });
    }
    // NumSteps substeps of dt / NumSteps; false as the reference (no signal)
    bool Step(float dt)
    {
        detail::checked([&] { scStepCloth(h_, dt); });
        return false;
    }
    // the render mesh: positions, angle-weighted normals (Physics/ClothPatch.cpp:266-273)
    bool UpdateMesh()
    {
        detail::checked([&] { scUpdateClothMesh(h_); });
        return false;
    }
    Mesh& GetMesh()
    {
        if (!mesh_.h_.ptr) mesh_ = Mesh(detail::checked([&] { return scGetClothMesh(h_); }), false);
        return mesh_;
    }
    const ClothModel& GetModel() const { return model_; }
    ClothModel& GetModel() { return model_; }
    ClothModel* GetModelPtr() { return &model_; }
    void SetMethod(MethodType) {}
    MethodType GetMethod() const { return METHOD_PBD; }
    int GetNumSteps() const { return detail::checked([&] { return scGetNumSteps(h_); }); }
    void SetNumSteps(int v) { detail::checked([&] { scSetNumSteps(h_, v); }); }
    const CollisionWorld& GetCollWorld() const { return *world_; }
    // per-vertex strain colours (Physics/ClothPatch.cpp:337-389)
    void ComputeStrainMap(std::vector<Vector3>& colors)
    {
        std::vector<scVector3> c(model_.GetNumParticles());
        detail::checked([&] { scComputeStrainMap(h_, c.data(), (int)c.size()); });
        colors.assign(c.begin(), c.end());
    }
    scClothPatch handle() const { return h_; }

private:
    scClothPatch h_;
    const CollisionWorld* world_;
    ClothModelPBD model_;
    Mesh mesh_;
};

inline void ClothModelPBD::Step(float h)
{
    const int n = owner_->GetNumSteps();
    owner_->SetNumSteps(1);
    try {
        owner_->Step(h);
    } catch (...) {
        owner_->SetNumSteps(n);
        throw;
    }
    owner_->SetNumSteps(n);
}

}  // namespace solidcloth

#ifndef SOLIDCLOTH_NO_PHYSICS_NAMESPACE
namespace Physics {
using solidcloth::Capsule;
using solidcloth::ClothModel;
using solidcloth::ClothModelPBD;
using solidcloth::ClothPatch;
using solidcloth::Collidable;
using solidcloth::CollisionMesh;
using solidcloth::CollisionSDF;
using solidcloth::CollisionWorld;
using solidcloth::Contact;
using solidcloth::Particle;
using solidcloth::SolverStats;
using solidcloth::SolverType;
using solidcloth::Sphere;
using solidcloth::Triangle;
using solidcloth::Walls;
using solidcloth::CF_EDGES;
using solidcloth::CF_SELF;
using solidcloth::CF_TRIANGLES;
using solidcloth::CF_VERTICES;
using solidcloth::CF_WALLS;
using solidcloth::METHOD_PBD;
using solidcloth::MethodType;
}  // namespace Physics
#endif
