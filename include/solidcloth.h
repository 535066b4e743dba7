/* solidcloth.h — C ABI of the MI355X-native SolidCloth substep library
 * (libsolidcloth.so).
 *
 * The exported names, argument order and argument meaning are those of the
 * reference's ClothLib ABI (Demo/ClothAPI.h:65-146). The reference passes C++
 * value types across that boundary; here they are layout-identical C structs
 * (scVector3 = Math::Vector3, Math/Vector3.h:12-16; scVector2 = Math::Vector2;
 * scAabb3 = Geometry::Aabb3, Geometry/Aabb3.h:10; scQuaternion =
 * Math::Quaternion {float s; Vector3 v}, Math/Quaternion.h:13-14) and C++
 * references become pointers, which is the same machine ABI on x86-64.
 *
 * Conventions kept from the reference: handles are {void* ptr} structs passed
 * by value (Demo/ClothAPI.h:11-32); scCreate* allocates and scFree* releases;
 * scAddCollidable transfers ownership of the collidable to the world; a patch
 * keeps a reference to its world, which must outlive it; type-mismatched
 * getters return without writing their outputs (Demo/ClothAPI.cpp:92-124).
 *
 * Difference: the reference reports no errors. This library additionally
 * records the last error (scGetLastError) and refuses calls it cannot honour
 * (for example a capsule or a mouse spring on a batched patch, which the
 * batch kernel does not handle) instead of computing something else. All device work started by a call has
 * completed before any getter returns data.
 */
#ifndef SOLIDCLOTH_H
#define SOLIDCLOTH_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SCEXP __attribute__((visibility("default")))

typedef struct { float x, y, z; } scVector3;
typedef struct { float x, y; } scVector2;
typedef struct { scVector3 min, max; } scAabb3;
typedef struct { float s; scVector3 v; } scQuaternion;

typedef void* scHandle; /* a collidable (Demo/ClothAPI.h:9) */
typedef struct { void* ptr; } scClothPatch;
typedef struct { void* ptr; } scCollisionWorld;
typedef struct { void* ptr; } scMesh;
typedef struct { void* ptr; } scModel;
typedef struct { void* ptr; } scSDF; /* replaces Geometry::SDF* (Demo/ClothAPI.h:80) */
typedef struct { void* ptr; } scClothBatch; /* many patches stepped together (no counterpart) */
typedef struct { void* ptr; } scWarmStart;  /* learned warm-start network (no counterpart) */

enum scCollFlags { /* Demo/ClothAPI.h:34-41 */
    scCollisionFlagVertices = 1,
    scCollisionFlagTriangles = 2,
    scCollisionFlagSelf = 4,
    scCollisionFlagWalls = 8,
    scCollisionFlagEdges = 16,
};

enum scCollidableType { /* Demo/ClothAPI.h:43-53 */
    scCollidableTypeInvalid,
    scCollidableTypeWalls,
    scCollidableTypeSlab,
    scCollidableTypeSphere,
    scCollidableTypeMesh,
    scCollidableTypeBox,
    scCollidableTypeCapsule,
    scCollidableTypeSDF,
};

/* ---- collision world (Demo/ClothAPI.h:69-74) ---- */
SCEXP scCollisionWorld scCreateCollisionWorld(void);                    /* :69 */
SCEXP void scFreeCollisionWorld(scCollisionWorld collWorld);             /* :70 */
SCEXP void scClearCollidables(scCollisionWorld collWorld);               /* :71 */
SCEXP int scGetNumCollidables(scCollisionWorld collWorld);               /* :72 */
SCEXP void scAddCollidable(scCollisionWorld collWorld, scHandle coll);   /* :73 */
SCEXP scHandle scGetCollidable(scCollisionWorld collWorld, int i);       /* :74 */

/* ---- collidables (Demo/ClothAPI.h:77-81) ---- */
SCEXP scHandle scCreateWalls(scAabb3 box);                                                     /* :77 */
SCEXP scHandle scCreateSphere(scVector3 center, float radius);                                 /* :78 */
SCEXP scHandle scCreateCapsule(scVector3 c, float rad, float halfHeight, scQuaternion quat);    /* :79 */
SCEXP scHandle scCreateSDF(scSDF s);                                                           /* :80 */
SCEXP scHandle scCreateMesh(scMesh m);                                                         /* :81 */

/* ---- introspection (Demo/ClothAPI.h:83-87) ---- */
SCEXP int scGetCollidableType(scHandle collidable);                                            /* :83 */
SCEXP void* scGetMeshTree(scHandle collMesh);                                                  /* :84 */
SCEXP void scGetWallsInfo(scHandle walls, scAabb3* box);                                       /* :85 */
SCEXP void scGetSphereInfo(scHandle sphere, float* radius, scVector3* center);                 /* :86 */
SCEXP void scGetCapsuleInfo(scHandle capsule, float* radius, float* halfHeight, scVector3* center,
                            scQuaternion* rot);                                                /* :87 */

/* ---- cloth patch (Demo/ClothAPI.h:91-99) ---- */
SCEXP scClothPatch scCreateClothPatch(scCollisionWorld collWorld);                             /* :91 */
SCEXP void scFreeClothPatch(scClothPatch patch);                                               /* :92 */
SCEXP scModel scGetClothModel(scClothPatch patch);                                             /* :93 */
SCEXP scMesh scGetClothMesh(scClothPatch patch);                                               /* :94 */
SCEXP void scInitCloth(scClothPatch patch, int divX, int divY, float inc, scVector3 offset, bool horizontal,
                       bool attached);                                                         /* :95 */
SCEXP void scStepCloth(scClothPatch patch, float dt);                                          /* :97 */
SCEXP void scUpdateClothMesh(scClothPatch patch);                                              /* :98 */
SCEXP void scComputeStrainMap(scClothPatch patch, scVector3* colors, int count);               /* :99 */

/* ---- cloth model getters (Demo/ClothAPI.h:102-119) ---- */
SCEXP int scGetNumParticles(scModel model);                                                    /* :102 */
SCEXP void scGetParticlePos(scModel model, int i, scVector3* pos);                             /* :103 */
SCEXP float scGetParticleInvMass(scModel model, int i);                                        /* :104 */
SCEXP int scGetNumTris(scModel model);                                                         /* :106 */
SCEXP void scGetTriangle(scModel model, int i, int* i1, int* i2, int* i3);                     /* :107 */
SCEXP void scGetTriangleInfo(scModel model, int i, float* invDet, scVector2* du, scVector2* dv,
                             scVector2* s);                                                    /* :108 */
SCEXP int scGetNumQuads(scModel model);                                                        /* :110 */
SCEXP void scGetQuad(scModel model, int i, int* i1, int* i2, int* i3, int* i4);                /* :111 */
SCEXP int scGetNumLinks(scModel model);                                                        /* :113 */
SCEXP void scGetLink(scModel model, int i, int* i1, int* i2);                                  /* :114 */
SCEXP int scGetNumContacts(scModel model);                                                     /* :116 */
SCEXP void scGetContact(scModel model, int i, scVector3* point, scVector3* normal);            /* :117 */
SCEXP int scGetNumTriContacts(scModel model);                                                  /* :118 */
SCEXP void scGetTriContact(scModel model, int i, scVector3* point, scVector3* normal);         /* :119 */

/* ---- parameters (Demo/ClothAPI.h:121-132) ---- */
SCEXP int scGetCollisionFlags(scModel model);                                                  /* :121 */
SCEXP void scSetCollisionFlags(scModel model, int val);                                        /* :122 */
SCEXP float scGetThickness(scModel model);                                                     /* :123 */
SCEXP void scSetThickness(scModel model, float val);                                           /* :124 */
SCEXP float scGetTolerance(scModel model);                                                     /* :125 */
SCEXP void scSetTolerance(scModel model, float val);                                           /* :126 */
SCEXP float scGetFriction(scModel model);                                                      /* :127 */
SCEXP void scSetFriction(scModel model, float val);                                            /* :128 */
SCEXP int scGetNumIterations(scModel model);                                                   /* :129 */
SCEXP void scSetNumIterations(scModel model, int val);                                         /* :130 */
SCEXP bool scGetUseFEM(scModel model);                                                         /* :131 */
SCEXP void scSetUseFEM(scModel model, bool val);                                               /* :132 */

/* ---- tree and mouse spring (Demo/ClothAPI.h:134-138) ----
 * The spring (stiffness 100, rest length 0, ClothModel.inl:86-98) acts in
 * every substep of scStepCloth right after the prediction
 * (Physics/ClothPBD.cpp:267-282); scStepClothBatch refuses a patch with one. */
SCEXP void* scGetClothTree(scModel model);                                                     /* :134 */
SCEXP void scAddMouseSpring(scModel model, int i1, int i2, int i3, scVector3 coords, scVector3 p); /* :136 */
SCEXP void scUpdateMouseSpring(scModel model, scVector3 p);                                    /* :137 */
SCEXP void scRemoveMouseSpring(scModel model);                                                 /* :138 */

/* ---- mesh buffers (Demo/ClothAPI.h:141-145) ---- */
SCEXP int scGetMeshVertexCount(scMesh mesh);                                                   /* :141 */
SCEXP int scGetMeshIndexCount(scMesh mesh);                                                    /* :142 */
SCEXP scVector3* scGetMeshNormalBuffer(scMesh mesh);                                           /* :143 */
SCEXP scVector3* scGetMeshVertexBuffer(scMesh mesh);                                           /* :144 */
SCEXP unsigned int* scGetMeshIndexBuffer(scMesh mesh);                                         /* :145 */

/* =====================================================================
 * Extensions. The reference sets these through C++ only
 * (Physics/ClothModel.h:110-146, Physics/ClothPatch.h:50) or has no
 * counterpart; the benchmark configurations need them through the ABI.
 * ===================================================================== */

/* ClothModel setters/getters, Physics/ClothModel.h:110-146 */
SCEXP void scSetUnit(scModel model, float val);               /* SetUnit :133 */
SCEXP float scGetUnit(scModel model);
SCEXP void scSetShearStiffness(scModel model, float val);     /* SetShearStiffness :113 (FEM Young modulus) */
SCEXP void scSetStretchStiffness(scModel model, float val);   /* SetStretchStiffness :111 */
SCEXP void scSetBendStiffness(scModel model, float val);      /* SetBendStiffness :115 */
SCEXP void scSetPoisson(scModel model, float val);            /* SetPoisson :129 */
SCEXP void scSetDihedral(scModel model, bool val);            /* SetDihedral :123 */
SCEXP void scSetDragCoeff(scModel model, float val);          /* SetDragCoeff :146 */
SCEXP float scGetShearStiffness(scModel model);               /* GetShearStiffness :112 */
SCEXP float scGetStretchStiffness(scModel model);             /* GetStretchStiffness :110 */
SCEXP float scGetBendStiffness(scModel model);                /* GetBendStiffness :114 */
SCEXP float scGetPoisson(scModel model);                      /* GetPoisson :128 */
SCEXP bool scGetDihedral(scModel model);                      /* GetDihedral :122 */
SCEXP float scGetDragCoeff(scModel model);                    /* GetDragCoeff :145 */
/* GetSelfShrink :141 and GetSelfThreshold :134 (set with scSetSelfParams) */
SCEXP void scGetSelfParams(scModel model, float* shrink, float* threshold);
SCEXP void scSetNumSteps(scClothPatch patch, int val);        /* ClothPatch::SetNumSteps, ClothPatch.h:50 */
SCEXP int scGetNumSteps(scClothPatch patch);

/* Iteration control. fixed = true removes the early exit of
 * Physics/ClothPBD.cpp:356-358 so exactly NumIterations iterations run. */
SCEXP void scSetFixedIterations(scModel model, bool fixed);

/* Vertical-mode incline of scInitCloth (Physics/ClothPatch.cpp:153,183).
 * The reference draws it from rand(); this sets it explicitly. Default 0. */
SCEXP void scSetInitIncline(scClothPatch patch, float incline);

/* Grid SDF with (nx+1)(ny+1)(nz+1) node samples, x fastest (Geometry/SDF.h:
 * 49-52), over box: Create + CreateFromImplicit's stored samples +
 * ComputeGradient (Geometry/SDF.cpp:10-56,210-256). */
SCEXP scSDF scCreateGridSDF(const float* values, int nx, int ny, int nz, scAabb3 box);
SCEXP void scFreeSDF(scSDF sdf);

/* SDF file (Geometry/SDF.cpp:75-151, WriteToFile / LoadFromFile): 3 int32 cell
 * counts, the box as 6 doubles (min.x, max.x, min.y, max.y, min.z, max.z),
 * then (nx+1)(ny+1)(nz+1) double node values, x fastest, native byte order.
 * Loading narrows box and values to float and computes the gradients; a file
 * that cannot be opened or is short records an error and returns a null
 * handle (the reference leaves the SDF unchanged). Writing returns 1, or 0
 * with an error recorded. Grids of more than 2^31 - 1 nodes are refused.
 * scGetSDFGrid copies the cell counts (3 ints), the box and at most
 * `capacity` node values (any pointer may be null) and returns the node
 * count. */
/* SDF::CreateFromMesh (Geometry/SDF.cpp:58-76) after SDF::Create(nx, ny, nz,
 * box), computed on the GPU over the mesh's triangle LBVH: node values are
 * the distance to the nearest triangle (distance 0 skipped; FLT_MAX when no
 * triangle is left), unsigned, because the reference's query returns the
 * unit delta as its normal and so its sign test never flips a value. The
 * gradients follow as for scCreateGridSDF. The mesh is uploaded to the GPU
 * on first use and stays there. Not in the reference's ABI (SDF creation
 * there is C++ only). */
SCEXP scSDF scCreateSDFFromMesh(scMesh mesh, int nx, int ny, int nz, scAabb3 box);
SCEXP scSDF scLoadSDFFile(const char* path);
SCEXP int scWriteSDFFile(scSDF sdf, const char* path);
SCEXP int scGetSDFGrid(scSDF sdf, int* cells, scAabb3* box, float* values, int capacity);

/* Indexed triangle mesh from arrays (replaces Geometry::Mesh* in scCreateMesh). */
SCEXP scMesh scCreateMeshFromArrays(const scVector3* vertices, int numVertices, const unsigned int* indices,
                                    int numIndices);
SCEXP void scFreeMesh(scMesh mesh);

/* Wavefront OBJ (Geometry/Assets.cpp:309-451, LoadMeshFromObj): v, vt and f
 * lines, the first three corners of each face, 1-based indices. naive: one
 * vertex per v line. Otherwise one vertex per distinct (position, normal, uv)
 * corner in order of first appearance, with its uv (scGetMeshUVBuffer). Normals
 * are computed from the triangles, not read. An unreadable file or a bad
 * index records an error and returns a null handle. */
SCEXP scMesh scLoadMeshFromObj(const char* path, bool naive);
/* Per-vertex material coordinates of a mesh, or null when it has none. */
SCEXP scVector2* scGetMeshUVBuffer(scMesh mesh);

/* Bulk state access (host arrays of count elements). */
SCEXP int scGetParticlePositions(scModel model, scVector3* out, int count);
SCEXP int scSetParticlePositions(scModel model, const scVector3* in, int count);
SCEXP int scGetParticleVelocities(scModel model, scVector3* out, int count);
SCEXP int scSetParticleVelocities(scModel model, const scVector3* in, int count);

/* Links in index order (Physics/ClothTypes.h:94-99): ids = (i1, i2),
 * data = (rest length, stiffness) per link; returns the link count. */
SCEXP int scGetLinks(scModel model, int* ids, float* data, int count);

/* Number of bend constraints and their vertex quadruples (i1,i2,i3,i4). */
SCEXP int scGetNumBends(scModel model);
SCEXP int scGetBends(scModel model, int* out, int count);

/* Gauss-Seidel order the GPU applies: kind 0 = triangles, 1 = bends,
 * 2 = obstacle-edge contacts (given as cloth edge indices), 3 = obstacle-
 * vertex contacts (given as cloth triangle indices), 4 = links (the
 * non-FEM stretch model, ClothPBD.cpp:9-50). Writes the constraint
 * indices in processing order and returns their count.
 * scGetColouring returns the number of launch phases and writes phase start
 * offsets into that order (numPhases + 1 values, colourOffsets may be NULL).
 * scGetSchedule writes, per position of the order, the phase, the slab (-1
 * when the sweep launches one colour per phase) and the colour inside the slab. */
SCEXP int scGetConstraintOrder(scModel model, int kind, int* order, int count);
SCEXP int scGetColouring(scModel model, int kind, int* colourOffsets, int maxColours);
SCEXP int scGetSchedule(scModel model, int kind, int* phase, int* slab, int* colour, int count);
/* The triangles' tile schedule (sweep mode 3; no counterpart in the
 * reference): counts = (tiles, tile vertex entries, colours, instances);
 * tileVertBegin (tiles + 1) delimits each tile's vertices in tileVerts
 * (ascending global ids; owned[i] = 1 for the tile's own vertices, 0 for its
 * halo); tileColOffsets (tiles * colours + 1) delimits, per tile and colour,
 * the tile's constraint instances, whose triangle indices are in instances.
 * Any array may be NULL. Returns the tile count (0 when the triangles are
 * not on tiles). */
SCEXP int scGetTileSchedule(scModel model, int* counts, int* tileVertBegin, int* tileVerts, int* owned,
                            int* tileColOffsets, int* instances);

/* A patch from a triangle mesh (ClothPatch::Init(const Mesh&, offset,
 * attached), Physics/ClothPatch.cpp:43-133, which the reference's demo calls
 * from C++): vertices at equal positions are welded into one particle,
 * triangles take the mesh's material coordinates (scSetMeshUVs) and, with
 * attached, vertices whose colour (scSetMeshColors) has a non-zero red are
 * pinned; links, edges and bends come from the mesh's edges. The render mesh
 * (scGetClothMesh) keeps the unwelded vertices. */
SCEXP void scInitClothFromMesh(scClothPatch patch, scMesh mesh, scVector3 offset, bool attached);
SCEXP int scSetMeshUVs(scMesh mesh, const scVector2* uvs, int count);
SCEXP int scSetMeshColors(scMesh mesh, const scVector3* colors, int count);

/* Solver (ClothModel::SetSolver / GetSolver, Physics/ClothModel.h:137-138;
 * values of SolverType, :23-30): 1 Gauss-Seidel (default), 0 Jacobi
 * (SolveJacobi), 2-4 the conjugate-residual solver (SolveCR, which the
 * reference runs for every other value). Jacobi and CR project the links only,
 * as in the reference (no triangles, bends or contacts). */
SCEXP void scSetSolver(scModel model, int solver);
SCEXP int scGetSolver(scModel model);

/* Sweep schedule: mode 0 = one launch per colour (positions in HBM),
 * mode 1 = slab phases with the slab's vertices staged in LDS,
 * slabs being strips of slabSize constraints; mode 2 = slab phases with
 * strip interiors of about slabSize constraints in one phase and the seams
 * between strips in the next. slabSize 0 (modes 1 and 2) lets the library
 * choose among both kinds by estimated sweep cost. Mode 3 = the FEM
 * triangles on tiles (one global colouring; each workgroup sweeps its own
 * vertices and a halo as many rings wide as there are colours, in one
 * launch per pass, writing back its own vertices), bends and links as in
 * mode 1 (a mesh whose tiles would not fit the workgroup's LDS keeps mode 1
 * for its triangles). Mode 4 (the default) = the schedules of mode 3, each
 * laid out so that a launch requests all of its data before its first colour
 * (where a schedule fits those kernels' limits, else as in mode 3); same
 * order and results as mode 3. In modes 3 and 4 a slabSize above 0 sets
 * the interiors of the bend and link slabs (interiors and seams, as in mode
 * 2), 0 lets the library choose. Changes the Gauss-Seidel order. */
SCEXP void scSetSweep(scModel model, int mode, int slabSize);

/* Contacts of the last substep against mesh obstacles, in the reference's
 * list order (collidable, then cloth vertex / edge / triangle).
 * scGetContactFeatures: the obstacle triangle of each vertex contact of
 * scGetContact (-1 for walls, spheres and SDFs).
 * scGetEdgeContactData (Physics/ClothTypes.h:110-120 EdgeContact): ids =
 * (cloth edge, obstacle edge, collidable), data = (point[3], normal[3], w1,
 * w2, lambda, lambdaF) per contact. scGetTriContactData (TriContact): ids =
 * (cloth triangle, obstacle vertex, collidable), data = (point[3], normal[3],
 * w1, w2, w3, lambda, lambdaF). All return the contact count. */
SCEXP int scGetContactFeatures(scModel model, int* out, int count);

/* Diagnostic: builds the device LBVH over n random boxes (seeded), refits it
 * `refits` times to new random boxes and checks every result on the host.
 * Returns the number of violations (0 = correct); msOut[0] = build time,
 * msOut[1] = average refit time (ms, device events). No counterpart in the
 * reference (its tree is a host AabbTree). */
SCEXP int scLbvhSelfCheck(int n, unsigned int seed, int refits, double* msOut);
SCEXP int scGetEdgeContactData(scModel model, int* ids, float* data, int cap);
SCEXP int scGetTriContactData(scModel model, int* ids, float* data, int cap);

/* Continuous self collision (CF_SELF): the detector of
 * Physics/SelfCollisionDetector.cpp (vertex-triangle and edge-edge CCD over
 * swept boxes; the coplanarity cubic in FP64) and the projections of
 * ClothCollisionHandler.cpp:364-541.
 * scGetSelfContacts: kind 0 = vertex-triangle (ids i1 i2 i3 = triangle,
 * i4 = vertex), 1 = edge-edge (i1 i2 against i3 i4), in the order they are
 * solved (colour group by colour group, then key); keys = (triangle << 32 |
 * vertex) or (edge1 << 32 | edge2); data = (w1, w2, w3, normal[3], side,
 * lambda) per contact. Returns the count.
 * scSetSelfParams: ClothModel's self shrink and exit threshold
 * (Physics/ClothModel.cpp:10-36 defaults 0.9, 0.003).
 * scSetSelfCCD: the detector's absolute thresholds (coplanarity volume 0.001,
 * bisection 0.01, vertex-triangle distance 0.1 in the reference) for scenes
 * at other length scales.
 * scSetSelfCapacity: initial contacts per kind per substep (0 = automatic).
 * A step whose contacts or candidate pairs outgrow their buffers is rerun
 * from its starting state with the buffers grown to the step's peaks; a
 * step that still cannot complete (more than 2^30 contacts or candidate
 * pairs of a kind) restores its starting state and records an error.
 * scGetSelfStats: out[0] contact capacity in use, out[1] steps rerun after
 * an overflow so far, out[2] / out[3] most vertex-triangle / edge-edge
 * contacts in one substep of the last step, out[4] candidate-pair pool
 * capacity (both kinds, pairs), out[5] / out[6] vertex-triangle / edge-edge colour
 * groups of the last substep, out[7] / out[8] distinct vertices of the
 * vertex-triangle / edge-edge contacts of the last substep. Returns the
 * number of values (9), writing at most count. No counterpart in the
 * reference (its lists are std::vectors). */
SCEXP int scGetSelfContacts(scModel model, int kind, unsigned long long* keys, int* ids, float* data, int cap);
SCEXP void scSetSelfParams(scModel model, float shrink, float threshold);
SCEXP void scSetSelfCCD(scModel model, float volume, float bisect, float vtDistance);
SCEXP void scSetSelfCapacity(scModel model, int capacity);
/* scSetSelfEdgeWeights: how a self edge-edge projection
 * (Physics/ClothCollisionHandler.cpp:452-540) distributes its displacement.
 * 0 (default) as the reference writes it: with the closest points
 * p = (1 - t1) x1 + t1 x2 and q = (1 - t2) x3 + t2 x4, it moves x1 by t1 and
 * x2 by 1 - t1 (likewise x3, x4; :531-534), the other end's weight, while its
 * effective mass uses each end's own; a correction near x1 then lands mostly
 * on x2 and does not close the gap, and dense self contact collapses the
 * cloth. 1: each end moves by its own weight (x1 by 1 - t1, x2 by t1, ...).
 * No counterpart in the reference (an extension). */
SCEXP void scSetSelfEdgeWeights(scModel model, int mode);
/* scGetSelfCandidates: the broad-phase candidate pairs of the last substep
 * (SelfCollisionDetector.cpp:60-135, Midphase, as one set: every pair whose
 * swept boxes overlap as closed intervals, Geometry/Aabb3.h:72-87, after the
 * filters of :92 and :122), kind 0 = (triangle, vertex), 1 = (edge1, edge2)
 * with edge1 < edge2, sorted, each pair once (the reference's tree may list
 * a pair twice). Writes at most cap pairs (2 ints each) and returns the
 * count. No counterpart in the reference's API. */
SCEXP long long scGetSelfCandidates(scModel model, int kind, int* pairs, long long cap);
SCEXP int scGetSelfStats(scModel model, int* out, int count);

/* Solver statistics of the last substep (Physics/ClothModel.h:32-40):
 * out[0] iterations, out[1] bend error, out[2] vertex-contact error,
 * out[3] self vertex-triangle error, out[4] self edge-edge error,
 * out[5] link error (link model; ClothModel.h:32-40 linksError). out holds
 * 6 floats. */
SCEXP void scGetSolverStats(scModel model, float* out);

/* Kernel timing with HIP events on the solver's stream. which: 0 FEM sweep,
 * 1 bend sweep, 2 contacts, 3 detection, 4 predict+velocity. While enabled,
 * steps run without graph replay so each launch is bracketed. */
SCEXP void scEnableKernelTiming(scModel model, bool enable);
SCEXP int scGetKernelTiming(scModel model, int which, int* launches, double* totalMs);

/* Use a captured hipGraph for the substep (default true). */
SCEXP void scSetUseGraph(scModel model, bool enable);

/* Device-to-device copy of `bytes` bytes (rounded down to 16) with a plain
 * streaming kernel, repeated `reps` times; returns the average kernel time in
 * milliseconds (HIP events), or a negative value on error. Used to measure
 * attainable HBM bandwidth and to calibrate traffic counters. */
SCEXP double scBandwidthProbe(long long bytes, int reps);

/* Inverse masses of all particles (0 = pinned), in the order of
 * scGetParticleInvMass (Physics/ClothModel.h:102 GetParticle(i).invMass). */
SCEXP int scSetParticleInvMasses(scModel model, const float* in, int count);

/* Batched stepping (no counterpart in the reference, whose demo steps one
 * patch per scStepCloth, Demo/ClothAPI.cpp:139-142): patches of any size,
 * each with its own collision world, stepped by one launch per substep in
 * which one workgroup runs one cloth's whole substep. The result of each
 * patch equals scStepCloth with the colour sweep (scSetSweep mode 0), which
 * scStepClothBatch selects for its patches. Patches support vertex and wall
 * contacts against walls, spheres and grid SDFs; all need the same NumSteps.
 * Free the batch before its patches.
 * scBatchSetLaunch: block size 256, 512 or 1024 (0 = autotune) and position
 * staging in LDS (0 off, 1 on, -1 = autotune). Autotuning times each
 * candidate on one real step (every candidate computes identical results).
 * scBatchGetLaunch: the chosen values and whether tuning is complete.
 * scBatchGetTuning: the configurations timed so far (block size, staging)
 * and their step times in ms; returns their count.
 * scBatchGetStepTime: device time in ms of the last step's substep launches
 * (HIP events on the batch's stream) and the number of launches. */
SCEXP scClothBatch scCreateClothBatch(void);
SCEXP void scFreeClothBatch(scClothBatch batch);
SCEXP void scBatchAddPatch(scClothBatch batch, scClothPatch patch);
SCEXP int scBatchGetNumPatches(scClothBatch batch);
SCEXP void scStepClothBatch(scClothBatch batch, float dt);
SCEXP void scBatchSetLaunch(scClothBatch batch, int blockSize, int stage);
SCEXP void scBatchGetLaunch(scClothBatch batch, int* blockSize, int* stage, int* tuned);
SCEXP int scBatchGetTuning(scClothBatch batch, int* configs, double* ms, int cap);
SCEXP double scBatchGetStepTime(scClothBatch batch, int* launches);
/* Replay the captured substep launches (default) or launch them directly. */
SCEXP void scBatchSetUseGraph(scClothBatch batch, bool enable);
/* Per-cloth load balance of the batch kernel (no counterpart in the
 * reference): with enable, every step records, for its last substep launch,
 * when each cloth's workgroup started and ended on the device's
 * constant-rate clock and the unit it ran on. scBatchGetWorkgroupTimes
 * writes, per patch in the order added (cap entries at most; any pointer may
 * be null), start and end in microseconds from the launch's earliest start
 * and the unit as XCC_ID << 32 | HW_ID (the shader engine, array and compute
 * unit are HW_ID bits 13-15, 12 and 8-11); returns the number of patches
 * recorded (0 before a recorded step), or -1 on error. */
SCEXP void scBatchRecordWorkgroupTimes(scClothBatch batch, bool enable);
SCEXP int scBatchGetWorkgroupTimes(scClothBatch batch, double* startUs, double* endUs, long long* unit, int cap);

/* Learned warm start (no counterpart in the reference). A per-vertex MLP
 * (12 one-ring features -> 64 -> 64 -> 3, GELU, BF16 MFMA with FP32
 * accumulation, FP32 Adam masters) predicts the correction the Gauss-Seidel
 * sweeps of a substep will apply (ClothPBD.cpp:324-369); the prediction is
 * added to the inertial positions before the sweeps.
 * scCreateWarmStart: seeded Glorot-uniform weights, zero biases.
 * scWarmStartGetParams / SetParams: the 5187 FP32 parameters (W1 [12][64],
 * b1, W2 [64][64], b2, W3 [64][3], b3); setting them restarts Adam.
 * Returns the parameter count. scWarmStartSetLearningRate: Adam's step size
 * (default 1e-3). scWarmStartGetSteps: Adam steps taken.
 * Host-array entry points (features n x 12, targets n x 3, weights n or
 * NULL for all ones): scWarmStartForward writes y (n x 3);
 * scWarmStartGradient writes the gradient of the weighted mean squared error
 * sum_v w_v |y_v - t_v|^2 / sum_v w_v and returns that loss;
 * scWarmStartTrain takes `steps` Adam steps on the data and returns the
 * loss before the last one.
 * scSetWarmStart: mode 0 off, 1 train (the plain solve runs; after every
 * substep the network takes one Adam step towards the correction the sweeps
 * applied), 2 solve (prediction, then `iterations` sweeps), 3 track (the
 * plain solve; only the corrections are recorded, so that a later switch to
 * mode 2 starts with current features). The network must
 * outlive the attachment; models sharing a network step one at a time.
 * In mode 2 each scStepCloth is judged by the plain solver. With
 * scSetWarmStartFallback(t < 0) (the default) the plain step is run from the
 * same starting state first and judges the warm-started step: whenever the
 * warm-started residual (scGetResidual) exceeds the plain one, the plain
 * result is kept (a verification mode: every step costs a plain and a
 * warm-started step). With t >= 0 the warm-started step is rerun plain
 * whenever its residual exceeds t; t = +infinity turns the judge off (the
 * production mode: no snapshot, no residual, no plain step).
 * scGetResidual: RMS constraint residual of the current positions, FP64:
 * sqrt((sum_tri warp^2 + weft^2 + sum_bend (s theta)^2) / (2T + B)) with
 * warp, weft as in ClothPBD.cpp:78,89 and s theta the bend error of
 * ClothPBD.cpp:179-224. The compliant shear constraint (ClothPBD.cpp:96-101),
 * whose error after a substep is set by its compliance, is left out;
 * scGetResidualTerms writes the RMS of warp/weft, of shear and of the bend
 * errors separately.
 * scGetWarmStartStats (out holds 7 doubles): out[0] last step's residual,
 * [1] bound in use (t, or the last plain step's residual), [2] fallbacks,
 * [3] the last step's plain residual (-1 if none ran), [4] judged steps,
 * [5] 1 if the last step's result is the plain one, [6] the last step's
 * warm-started residual.
 * scGetWarmStartData: which 0 features (n x 12) of the last substep, 1 its
 * inertial positions (n x 3), 2 the last correction (n x 3), 3 training
 * targets and weights (n x 4), 4 (ell, free vertices); returns n (2 for 4).
 * scSaveState / scRestoreState: one device-side snapshot of positions,
 * velocities and the warm start's last correction. */
SCEXP scWarmStart scCreateWarmStart(unsigned int seed);
SCEXP void scFreeWarmStart(scWarmStart ws);
SCEXP int scWarmStartGetParams(scWarmStart ws, float* out, int count);
SCEXP int scWarmStartSetParams(scWarmStart ws, const float* in, int count);
SCEXP void scWarmStartSetLearningRate(scWarmStart ws, float lr);
SCEXP int scWarmStartGetSteps(scWarmStart ws);
SCEXP int scWarmStartForward(scWarmStart ws, const float* features, int n, float* out);
/* scWarmStartForwardNoAlloc: no counterpart in the reference. For callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc): scWarmStartForward with
 * DEVICE pointers, features (n x 12 FP32) in and out (n x 3 FP32) out, the
 * BF16 staging of the rows taken from a static per-GPU array of the library
 * (8 MiB, held until the process ends) instead of allocated memory. At most
 * 262144 samples (larger n records an error and returns -1); no hipMalloc,
 * hipFree or other allocation. Enqueued on `stream` (a hipStream_t; NULL =
 * the null stream) and returns n without waiting; the caller synchronizes.
 * Calls that overlap on one GPU share the staging array and race. */
SCEXP int scWarmStartForwardNoAlloc(scWarmStart ws, const float* features, int n, float* out, void* stream);
SCEXP float scWarmStartGradient(scWarmStart ws, const float* features, const float* targets, const float* weights,
                                int n, float* grad);
SCEXP float scWarmStartTrain(scWarmStart ws, const float* features, const float* targets, const float* weights,
                             int n, int steps);
SCEXP void scSetWarmStart(scModel model, scWarmStart ws, int mode, int iterations);
SCEXP void scSetWarmStartFallback(scModel model, float threshold);
SCEXP double scGetResidual(scModel model);
SCEXP void scGetResidualTerms(scModel model, double* out);
SCEXP void scGetWarmStartStats(scModel model, double* out);
SCEXP int scGetWarmStartData(scModel model, int which, float* out, int count);
/* Device time (ms, HIP events) of the last scStepCloth's substep graph
 * launches, a fallback rerun included; 0 when the step launched directly
 * (kernel timing on or graphs off). */
SCEXP double scGetLastStepTime(scModel model);
SCEXP void scSaveState(scModel model);
SCEXP void scRestoreState(scModel model);

/* Last error message of the calling thread, or "" (not in the reference),
 * and the number of errors recorded on the calling thread so far, so a
 * caller can tell whether its own call failed. */
SCEXP const char* scGetLastError(void);
SCEXP int scGetErrorCount(void);

/* Library identification. */
SCEXP const char* scGetBuildInfo(void);

#ifdef __cplusplus
}
#endif

#endif /* SOLIDCLOTH_H */
