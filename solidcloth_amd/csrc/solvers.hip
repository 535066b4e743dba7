// The reference's link solvers other than Gauss-Seidel (Physics/ClothPBD.cpp:
// 290-298): Jacobi (SolveJacobi / SolveLinksJacobi, :616-662) and the
// nonlinear conjugate-residual solver as compiled (SolveCR with SKIP_ALPHA and
// SKIP_BETA defined, :371-478). Both compute one displacement per link from
// the positions at the start of the pass, then apply all of them; on the
// device one thread per link computes the displacements and one thread per
// particle applies its links' in ascending link order — the order in which
// the reference's loop reaches that particle — so the sums round alike.
#include <hip/hip_runtime.h>

#include "devutil.h"
#include "hiputil.h"
#include "kernels.h"

namespace sc {

namespace {

constexpr int kBlock = 256;
inline int blocks(int n) { return (n + kBlock - 1) / kBlock; }

// ClothPBD.cpp:616-640 (first loop of SolveLinksJacobi)
__global__ void __launch_bounds__(kBlock) k_jacobi_links(const float4* __restrict__ pos, const int2* __restrict__ lk,
                                                         const float2* __restrict__ lr, int nl, float k, float omega,
                                                         float4* __restrict__ disp)
{
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nl) return;
    const int2 e = lk[l];
    const float4 P1 = pos[e.x], P2 = pos[e.y];
    const float mu = 1.f / (P1.w + P2.w);
    f3 delta = sub(xyz(P1), xyz(P2));
    const float len0 = lr[l].x;
    const float ln = len(delta);
    delta = mul(1.f / ln, delta);
    const float err = ln - len0;
    const float epsilon = len0 / (k * lr[l].y);
    const float s = 1.f / (1.f + mu * epsilon);
    const float lambda = err * mu * s;
    disp[l] = f4(mul(omega * lambda, delta), 0.f);
}

// ClothPBD.cpp:371-379 (ComputeErrorAndNormal), :433-439 (d = r + beta d
// without the Fletcher-Reeves test) and :470-477 (dLambda = alpha d)
__global__ void __launch_bounds__(kBlock) k_cr_links(const float4* __restrict__ pos, const int2* __restrict__ lk,
                                                     const float2* __restrict__ lr, int nl, float unit, float beta,
                                                     float alpha, float* __restrict__ d, float4* __restrict__ disp)
{
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nl) return;
    const int2 e = lk[l];
    const float2 L = lr[l];
    f3 delta = sub(xyz(pos[e.y]), xyz(pos[e.x]));
    const float c = (len(delta) - L.x) * L.y * unit;
    delta = mul(1.f / len(delta), delta);  // Vector3::Normalize
    const float r = -c;
    const float dn = r + beta * d[l];
    d[l] = dn;
    const float dLambda = alpha * dn;
    disp[l] = f4(mul(dLambda, delta), 0.f);
}

// the displacements of a particle's links in ascending link order: its
// first end subtracts m disp, its second end adds it
__global__ void __launch_bounds__(kBlock) k_link_apply(float4* __restrict__ pos, const int* __restrict__ off,
                                                       const int* __restrict__ inc, const float4* __restrict__ disp,
                                                       int nv)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    const int b = off[v], e = off[v + 1];
    if (b == e) return;
    const float4 P = pos[v];
    f3 p = xyz(P);
    for (int k = b; k < e; k++) {
        const int code = inc[k];  // link * 2 + end
        const f3 dd = mul(P.w, xyz(disp[code >> 1]));
        p = (code & 1) ? add(p, dd) : sub(p, dd);
    }
    pos[v] = f4(p, P.w);
}

__global__ void __launch_bounds__(kBlock) k_fill_zero(float* __restrict__ x, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0.f;
}

}  // namespace

void launch_jacobi_pass(hipStream_t s, float4* pos, const JacobiLinks& J, float k, float omega)
{
    if (J.nl <= 0) return;
    hipLaunchKernelGGL(k_jacobi_links, dim3(blocks(J.nl)), dim3(kBlock), 0, s, pos, J.lk, J.lr, J.nl, k, omega, J.disp);
    hipLaunchKernelGGL(k_link_apply, dim3(blocks(J.nv)), dim3(kBlock), 0, s, pos, J.off, J.inc, J.disp, J.nv);
}

void launch_cr_reset(hipStream_t s, const JacobiLinks& J)
{
    if (J.nl <= 0) return;
    hipLaunchKernelGGL(k_fill_zero, dim3(blocks(J.nl)), dim3(kBlock), 0, s, J.d, J.nl);
}

void launch_cr_pass(hipStream_t s, float4* pos, const JacobiLinks& J, float unit, float beta, float alpha)
{
    if (J.nl <= 0) return;
    hipLaunchKernelGGL(k_cr_links, dim3(blocks(J.nl)), dim3(kBlock), 0, s, pos, J.lk, J.lr, J.nl, unit, beta, alpha,
                       J.d, J.disp);
    hipLaunchKernelGGL(k_link_apply, dim3(blocks(J.nv)), dim3(kBlock), 0, s, pos, J.off, J.inc, J.disp, J.nv);
}

}  // namespace sc
