// Render-mesh sync on the device: the cloth's render mesh (Physics/
// ClothPatch.cpp:266-273 UpdateMesh) takes the particle positions plus the
// patch offset (Physics/ClothModel.inl:123-130) and angle-weighted vertex
// normals (Geometry/Mesh.cpp:10-63). Two passes: one thread per triangle
// computes its unit normal and corner angles, then one thread per vertex
// sums its triangles' contributions in ascending triangle order — the order
// in which the reference's triangle loop adds them — and normalises. The
// arithmetic follows the reference's operation order in FP32; the corner
// angles use the device's acosf / atan2f, whose last bit may differ from the
// host C library's.
#include <hip/hip_runtime.h>

#include "devutil.h"
#include "hiputil.h"
#include "kernels.h"

namespace sc {

namespace {

constexpr int kBlock = 256;
inline int blocks(int n) { return (n + kBlock - 1) / kBlock; }

// Geometry/Mesh.cpp:18-47 for triangle t: unit normal (xyz) and the corner
// angle at its first vertex (w); the other two corner angles in B.
__global__ void __launch_bounds__(kBlock) k_render_tris(const float4* __restrict__ pos, const int* __restrict__ map,
                                                        const unsigned* __restrict__ idx, int nt, f3 offset,
                                                        float4* __restrict__ triA, float2* __restrict__ triB,
                                                        float* __restrict__ triN)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    unsigned i1 = idx[3 * t], i2 = idx[3 * t + 1], i3 = idx[3 * t + 2];
    if (map) {
        i1 = map[i1];
        i2 = map[i2];
        i3 = map[i3];
    }
    const f3 v1 = add(xyz(pos[i1]), offset), v2 = add(xyz(pos[i2]), offset), v3 = add(xyz(pos[i3]), offset);
    const f3 n = normalized(cross(sub(v2, v1), sub(v3, v1)));
    const f3 e1 = normalized(sub(v2, v1)), e2 = normalized(sub(v3, v2)), e3 = normalized(sub(v1, v3));
    const float c1 = -dot(e1, e3), c2 = -dot(e1, e2), c3 = -dot(e2, e3);
    const float t1 = fabsf(c1) > 1.f ? fabsf(atan2f(len(cross(e1, neg(e3))), c1)) : acosf(c1);
    const float t2 = fabsf(c2) > 1.f ? fabsf(atan2f(len(cross(e2, neg(e1))), c2)) : acosf(c2);
    const float t3 = fabsf(c3) > 1.f ? fabsf(atan2f(len(cross(e3, neg(e1))), c3)) : acosf(c3);
    triA[t] = make_float4(n.x, n.y, n.z, t1);
    triB[t] = make_float2(t2, t3);
    triN[3 * t] = n.x;
    triN[3 * t + 1] = n.y;
    triN[3 * t + 2] = n.z;
}

// Geometry/Mesh.cpp:44-62 for vertex v: the sum over its (triangle, corner)
// entries vt[vtOff[v] .. vtOff[v+1]) (ascending triangle, then corner), then
// normalised unless zero; and the render position pos + offset.
__global__ void __launch_bounds__(kBlock) k_render_verts(const float4* __restrict__ pos, const int* __restrict__ map,
                                                         int nv, f3 offset,
                                                         const int* __restrict__ vtOff, const int* __restrict__ vt,
                                                         const float4* __restrict__ triA,
                                                         const float2* __restrict__ triB, float* __restrict__ outV,
                                                         float* __restrict__ outN)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    f3 acc{0.f, 0.f, 0.f};
    for (int k = vtOff[v], e = vtOff[v + 1]; k < e; k++) {
        const int code = vt[k];  // triangle * 4 + corner
        const int t = code >> 2, corner = code & 3;
        const float4 A = triA[t];
        const float2 B = triB[t];
        const float w = corner == 0 ? A.w : corner == 1 ? B.x : B.y;
        acc = add(acc, mul(w, xyz(A)));
    }
    const f3 n = len(acc) == 0.f ? f3{0.f, 0.f, 0.f} : normalized(acc);
    const f3 p = add(xyz(pos[map ? map[v] : v]), offset);
    outV[3 * v] = p.x;
    outV[3 * v + 1] = p.y;
    outV[3 * v + 2] = p.z;
    outN[3 * v] = n.x;
    outN[3 * v + 1] = n.y;
    outN[3 * v + 2] = n.z;
}

}  // namespace

void launch_render_mesh(hipStream_t s, const float4* pos, const int* map, int nv, const unsigned* idx, int nt,
                        f3 offset, const int* vtOff, const int* vt, float4* triA, float2* triB, float* outV,
                        float* outN, float* triN)
{
    if (nt > 0)
        hipLaunchKernelGGL(k_render_tris, dim3(blocks(nt)), dim3(kBlock), 0, s, pos, map, idx, nt, offset, triA, triB,
                           triN);
    if (nv > 0)
        hipLaunchKernelGGL(k_render_verts, dim3(blocks(nv)), dim3(kBlock), 0, s, pos, map, nv, offset, vtOff, vt, triA,
                           triB, outV, outN);
}

}  // namespace sc
