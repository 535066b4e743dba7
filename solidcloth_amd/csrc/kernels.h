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

#include "scmath.h"

namespace sc {

enum { CF_VERTICES = 1, CF_TRIANGLES = 2, CF_SELF = 4, CF_WALLS = 8, CF_EDGES = 16 };
enum { CT_INVALID = 0, CT_WALLS = 1, CT_SLAB = 2, CT_SPHERE = 3, CT_MESH = 4, CT_BOX = 5, CT_CAPSULE = 6, CT_SDF = 7 };

// Per-substep solver control block, written by the kernels. Errors are
// non-negative floats stored by their bit pattern so atomicMax orders them.
struct SolverCtl {
    unsigned done;
    int k;
    unsigned cb;  // bend error of the last bend pass
    unsigned cv;  // vertex-contact error of the current iteration
    int iterations;
    float bendError;
    float contactError;
    int pad;
};

enum { MAX_COLLIDABLES = 8 };

struct CollDev {
    int type;
    int pad;
    float ax, ay, az;  // walls: box min; sphere: centre; sdf: box min
    float bx, by, bz;  // walls: box max; sdf: cell steps
    float r;           // sphere radius
    const float4* nodes;
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
    int* src;  // producing collidable's world index, per slot
    int nv;
    int slots;
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
    const int4* info;
    const int* colOff;
    const int* verts;
    int base;  // first slab of the launched phase
};

enum { kSlabBlock = 256, kMaxSlabColours = 63 };

// Vertex contacts applied by the FEM slab kernel while it stages vertices
// (the previous iteration's contacts, by each vertex's owner slab).
struct SlabContacts {
    ContactBuf cb;
    const float4* prev;
    float4* vel;
    float thickness, mu, invH;
    int enabled;
};

// Launchers (kernels.hip). All are asynchronous on the given stream.
void launch_predict(hipStream_t s, float4* pos, float4* prev, float4* vel, int n, float h, f3 gdv, float drag,
                    SolverCtl* ctl);
void launch_detect(hipStream_t s, const float4* pos, const float4* prev, int n, const CollSet& cs, ContactBuf cb,
                   float thickness, float tolerance, int collFlags);
void launch_fem(hipStream_t s, float4* pos, TriStreams tr, int start, int count, float h2, float epsUV,
                const SolverCtl* ctl);
void launch_bend(hipStream_t s, float4* pos, const int4* bidx, int start, int count, float stiff, SolverCtl* ctl);
void launch_fem_slab(hipStream_t s, float4* pos, TriStreams tr, const ushort4* lidx, SlabSet ss, int numSlabs,
                     int maxVerts, float h2, float epsUV, const SolverCtl* ctl, const SlabContacts& kc);
void launch_predict_detect(hipStream_t s, float4* pos, float4* prev, float4* vel, int n, float h, f3 gdv, float drag,
                           SolverCtl* ctl, const CollSet& cs, ContactBuf cb, float thickness, float tolerance,
                           int collFlags);
void launch_contacts_velocity(hipStream_t s, float4* pos, const float4* prev, float4* vel, ContactBuf cb,
                              float thickness, float mu, float invH);
void launch_bend_slab(hipStream_t s, float4* pos, const ushort4* lidx, SlabSet ss, int numSlabs, int maxVerts,
                      float stiff, SolverCtl* ctl);
void launch_contacts(hipStream_t s, float4* pos, const float4* prev, float4* vel, ContactBuf cb, float thickness,
                     float mu, float invH, SolverCtl* ctl);
void launch_iter_end(hipStream_t s, SolverCtl* ctl, int maxIter, int fixedIter, int bendFreq, int dihedral,
                     float selfThreshold);
void launch_velocity(hipStream_t s, float4* pos, const float4* prev, float4* vel, int n, float invH);
void launch_copy(hipStream_t s, const float4* src, float4* dst, long n);

}  // namespace sc
