// Learned warm-start kernels for gfx950: features, the BF16-MFMA MLP forward
// and backward passes, gradient reduction, Adam, operand packing, and the
// FP64 constraint residual that judges the warm-started solve.
//
// MFMA layout (v_mfma_f32_16x16x16_bf16, one wavefront): for A (16 x 16,
// M x K) lane l holds A[l % 16][4 (l / 16) + j], j = 0..3; for B (K x N) lane
// l holds B[4 (l / 16) + j][l % 16]; the result D (M x N) gives lane l
// D[4 (l / 16) + j][l % 16]. The network runs transposed, Z^T = W^T H^T, with
// weights as A and activations as B (N = 16 vertices of a tile): a layer's D
// is then, element for element, the next layer's B. Weight gradients reduce
// over vertices (K = vertices), so their operands pass through LDS once,
// transposed.
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <string>

#include "devutil.h"
#include "hiputil.h"
#include "warmstart.h"

namespace sc {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

enum { kWsBlock = 256 };

// FP32 -> BF16, round to nearest even (NaN kept a quiet NaN).
__device__ __forceinline__ uint16_t to_bf16(float x)
{
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// GELU, tanh form (Hendrycks & Gimpel 2016), and its derivative.
__device__ __forceinline__ float gelu(float z)
{
    const float c = 0.7978845608f, a = 0.044715f;
    const float t = tanhf(c * (z + a * z * z * z));
    return 0.5f * z * (1.f + t);
}

__device__ __forceinline__ float gelu_grad(float z)
{
    const float c = 0.7978845608f, a = 0.044715f;
    const float t = tanhf(c * (z + a * z * z * z));
    return 0.5f * (1.f + t) + 0.5f * z * (1.f - t * t) * c * (1.f + 3.f * a * z * z);
}

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x4 pack4(float a, float b, float c, float d)
{
    bf16x4 r;
    r[0] = (short)to_bf16(a);
    r[1] = (short)to_bf16(b);
    r[2] = (short)to_bf16(c);
    r[3] = (short)to_bf16(d);
    return r;
}

// ------------------------------------------------------------- features ----

__device__ __forceinline__ float4 tri_strain(const f3 x1, const f3 x2, const f3 x3, const float4 A, const float4 B)
{
    const float du1 = A.x, du2 = A.y, dv1 = A.z, dv2 = A.w;
    const float invDet = B.x, lu2 = B.y, lv2 = B.z;
    const f3 dx1 = sub(x2, x1);
    const f3 dx2 = sub(x3, x1);
    const f3 wu = mul(invDet, sub(mul(dv2, dx1), mul(dv1, dx2)));
    const f3 wv = mul(invDet, add(mul(-du2, dx1), mul(du1, dx2)));
    return make_float4(0.5f * (dot(wu, wu) - lu2), 0.5f * (dot(wv, wv) - lv2), dot(wu, wv), 0.f);
}

__global__ void __launch_bounds__(kWsBlock) k_ws_tri_strain(const float4* __restrict__ pos, TriStreams tr, int nt,
                                                            float4* __restrict__ out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const int4 id = tr.idx[t];
    out[t] = tri_strain(xyz(pos[id.x]), xyz(pos[id.y]), xyz(pos[id.z]), tr.a[t], tr.b[t]);
}

__global__ void __launch_bounds__(kWsBlock) k_ws_features(const float4* __restrict__ pos,
                                                          const float4* __restrict__ prev, WarmBufs b, int nv,
                                                          float invEll)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    const float4 P = pos[v];
    b.xhat[v] = P;
    const f3 c = xyz(b.corr[v]);
    const f3 d = mul(invEll, sub(xyz(P), xyz(prev[v])));
    const f3 cs = mul(kWsScaleCorr * invEll, c);
    f3 sum{0.f, 0.f, 0.f};
    const int a0 = b.adjOff[v], a1 = b.adjOff[v + 1];
    for (int k = a0; k < a1; k++) sum = add(sum, xyz(b.corr[b.adj[k]]));
    f3 lap{0.f, 0.f, 0.f};
    if (a1 > a0) lap = mul(kWsScaleCorr * invEll, sub(mul(1.f / (float)(a1 - a0), sum), c));
    float s0 = 0.f, s1 = 0.f, s2 = 0.f;
    const int t0 = b.vtOff[v], t1 = b.vtOff[v + 1];
    for (int k = t0; k < t1; k++) {
        const float4 S = b.triStrain[b.vt[k]];
        s0 += S.x;
        s1 += S.y;
        s2 += S.z;
    }
    if (t1 > t0) {
        const float f = kWsScaleStrain / (float)(t1 - t0);
        s0 *= f;
        s1 *= f;
        s2 *= f;
    }
    uint16_t* x = b.X + (size_t)v * 16;
    const float vals[12] = {d.x, d.y, d.z, cs.x, cs.y, cs.z, lap.x, lap.y, lap.z, s0, s1, s2};
    uint4 w0, w1;
    w0.x = to_bf16(vals[0]) | ((uint32_t)to_bf16(vals[1]) << 16);
    w0.y = to_bf16(vals[2]) | ((uint32_t)to_bf16(vals[3]) << 16);
    w0.z = to_bf16(vals[4]) | ((uint32_t)to_bf16(vals[5]) << 16);
    w0.w = to_bf16(vals[6]) | ((uint32_t)to_bf16(vals[7]) << 16);
    w1.x = to_bf16(vals[8]) | ((uint32_t)to_bf16(vals[9]) << 16);
    w1.y = to_bf16(vals[10]) | ((uint32_t)to_bf16(vals[11]) << 16);
    w1.z = 0u;
    w1.w = 0u;
    ((uint4*)x)[0] = w0;
    ((uint4*)x)[1] = w1;
}

__global__ void __launch_bounds__(kWsBlock) k_ws_after(const float4* __restrict__ pos, WarmBufs b, int nv, int train,
                                                       float targetScale)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    const float4 P = pos[v];
    const f3 c = sub(xyz(P), xyz(b.xhat[v]));
    b.corr[v] = f4(c, 0.f);
    if (train) b.T[v] = f4(mul(targetScale, c), P.w != 0.f ? 1.f : 0.f);
}

__global__ void k_pack_features(const float* __restrict__ Xf, int n, uint16_t* __restrict__ X)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * 16) return;
    const int v = i >> 4, k = i & 15;
    X[i] = k < kMlpIn ? to_bf16(Xf[v * kMlpIn + k]) : (uint16_t)0;
}

__global__ void k_unpack_features(const uint16_t* __restrict__ X, int n, float* __restrict__ Xf)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * kMlpIn) return;
    const int v = i / kMlpIn, k = i - v * kMlpIn;
    Xf[i] = from_bf16(X[v * 16 + k]);
}

// ------------------------------------------------------------ operands ----
// Operand o, lane l, element j (i = l % 16, k = 4 (l / 16) + j):
//   0-3    F1[mt]      W1^T tile:   W1[k][16 mt + i]           (k < 12)
//   4-19   F2[mt][kb]  W2^T tile:   W2[16 kb + k][16 mt + i]
//   20-23  F3[kb]      W3^T tile:   W3[16 kb + k][i]           (i < 3)
//   24-27  G3[mt]      W3 tile:     W3[16 mt + i][k]           (k < 3)
//   28-43  G2[mt][kb]  W2 tile:     W2[16 mt + i][16 kb + k]
__global__ void k_mlp_pack(const float* __restrict__ W, uint16_t* __restrict__ packed, int* step)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx == 0 && step) *step += 1;
    if (idx >= kMlpOperands * 256) return;
    const int o = idx >> 8, l = (idx >> 2) & 63, j = idx & 3;
    const int i = l & 15, k = 4 * (l >> 4) + j;
    float v = 0.f;
    if (o < 4) {
        if (k < kMlpIn) v = W[kOffW1 + k * 64 + 16 * o + i];
    } else if (o < 20) {
        const int mt = (o - 4) >> 2, kb = (o - 4) & 3;
        v = W[kOffW2 + (16 * kb + k) * 64 + 16 * mt + i];
    } else if (o < 24) {
        const int kb = o - 20;
        if (i < kMlpOut) v = W[kOffW3 + (16 * kb + k) * 3 + i];
    } else if (o < 28) {
        const int mt = o - 24;
        if (k < kMlpOut) v = W[kOffW3 + (16 * mt + i) * 3 + k];
    } else {
        const int mt = (o - 28) >> 2, kb = (o - 28) & 3;
        v = W[kOffW2 + (16 * mt + i) * 64 + 16 * kb + k];
    }
    packed[idx] = to_bf16(v);
}

// ------------------------------------------------------------- forward ----

struct MlpBias {
    float b1[4][4], b2[4][4], b3[4];
};

__device__ __forceinline__ void load_bias(const float* __restrict__ W, int lane, MlpBias& b)
{
    const int r0 = 4 * (lane >> 4);
    for (int mt = 0; mt < 4; mt++)
        for (int j = 0; j < 4; j++) {
            b.b1[mt][j] = W[kOffB1 + 16 * mt + r0 + j];
            b.b2[mt][j] = W[kOffB2 + 16 * mt + r0 + j];
        }
    for (int j = 0; j < 4; j++) b.b3[j] = (r0 + j < kMlpOut) ? W[kOffB3 + r0 + j] : 0.f;
}

struct MlpAct {
    f32x4 z1[4], z2[4];
    bf16x4 h1[4], h2[4];
    f32x4 y;
};

// One tile of 16 vertices (xb: lane l holds features 4 (l/16)..+3 of vertex l % 16).
__device__ __forceinline__ void tile_forward(const bf16x4* __restrict__ ops, int lane, const MlpBias& b, bf16x4 xb,
                                             MlpAct& a)
{
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
    for (int mt = 0; mt < 4; mt++) {
        const f32x4 acc = mfma(ops[mt * 64 + lane], xb, zero);
        float h[4];
        for (int j = 0; j < 4; j++) {
            const float z = acc[j] + b.b1[mt][j];
            a.z1[mt][j] = z;
            h[j] = gelu(z);
        }
        a.h1[mt] = pack4(h[0], h[1], h[2], h[3]);
    }
    for (int mt = 0; mt < 4; mt++) {
        f32x4 acc = zero;
        for (int kb = 0; kb < 4; kb++) acc = mfma(ops[(4 + mt * 4 + kb) * 64 + lane], a.h1[kb], acc);
        float h[4];
        for (int j = 0; j < 4; j++) {
            const float z = acc[j] + b.b2[mt][j];
            a.z2[mt][j] = z;
            h[j] = gelu(z);
        }
        a.h2[mt] = pack4(h[0], h[1], h[2], h[3]);
    }
    f32x4 acc = zero;
    for (int kb = 0; kb < 4; kb++) acc = mfma(ops[(20 + kb) * 64 + lane], a.h2[kb], acc);
    for (int j = 0; j < 4; j++) a.y[j] = acc[j] + b.b3[j];
}

__device__ __forceinline__ bf16x4 load_x(const uint16_t* __restrict__ X, int v, int n, int lane)
{
    bf16x4 xb = {0, 0, 0, 0};
    if (v < n) xb = *(const bf16x4*)(X + (size_t)v * 16 + 4 * (lane >> 4));
    return xb;
}

// kMode 0: Y[v] = y; kMode 1: pos[v] += outScale y (free vertices only).
template <int kMode>
__global__ void __launch_bounds__(kWsBlock) k_mlp_forward(MlpWeights w, const uint16_t* __restrict__ X, int n,
                                                          float* __restrict__ Y, float4* __restrict__ pos,
                                                          float outScale)
{
    __shared__ bf16x4 ops[24 * 64];
    const bf16x4* src = (const bf16x4*)w.packed;
    for (int i = threadIdx.x; i < 24 * 64; i += kWsBlock) ops[i] = src[i];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    MlpBias b;
    load_bias(w.master, lane, b);
    const int tiles = (n + 15) >> 4;
    const int waves = gridDim.x * (kWsBlock / 64);
    // the loop bound depends on the wave only: every MFMA runs with all lanes
    for (int t = (blockIdx.x * kWsBlock + threadIdx.x) >> 6; t < tiles; t += waves) {
        const int v = t * 16 + (lane & 15);
        MlpAct a;
        tile_forward(ops, lane, b, load_x(X, v, n, lane), a);
        if (lane < 16 && v < n) {
            if (kMode == 0) {
                Y[3 * v] = a.y[0];
                Y[3 * v + 1] = a.y[1];
                Y[3 * v + 2] = a.y[2];
            } else {
                const float4 P = pos[v];
                if (P.w != 0.f) {
                    const f3 d{outScale * a.y[0], outScale * a.y[1], outScale * a.y[2]};
                    pos[v] = f4(add(xyz(P), d), P.w);
                }
            }
        }
    }
}

// ------------------------------------------------------------ backward ----

enum { kTrUnits = 16 + 64 + 64 + 64 + 64 + 16 };  // x, h1, dz2, dz1, h2, dy
enum { kTrX = 0, kTrH1 = 16, kTrDz2 = 80, kTrDz1 = 144, kTrH2 = 208, kTrDy = 272 };

// Stores a D-layout tile (lane: vertex l % 16, units 4 (l/16) + j) as [unit][vertex].
__device__ __forceinline__ void tr_store(uint16_t* tr, int unitBase, int lane, bf16x4 v)
{
    const int u = unitBase + 4 * (lane >> 4), vx = lane & 15;
    for (int j = 0; j < 4; j++) tr[(u + j) * 16 + vx] = (uint16_t)v[j];
}

// Reads the operand with lane l: unit unitBase + l % 16, vertices 4 (l/16)..+3.
__device__ __forceinline__ bf16x4 tr_load(const uint16_t* tr, int unitBase, int lane)
{
    return *(const bf16x4*)(tr + (unitBase + (lane & 15)) * 16 + 4 * (lane >> 4));
}

// Sum over the 16 lanes that share l / 16 (fixed butterfly: deterministic).
__device__ __forceinline__ float sum16(float v)
{
    for (int o = 1; o < 16; o <<= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__global__ void __launch_bounds__(kWsBlock) k_mlp_backward(MlpWeights w, const uint16_t* __restrict__ X,
                                                           const float4* __restrict__ T, int n, float invCount,
                                                           float* __restrict__ partial)
{
    __shared__ bf16x4 ops[kMlpOperands * 64];
    __shared__ __align__(16) uint16_t trs[kWsBlock / 64][kTrUnits * 16];
    __shared__ float acc[kMlpStride];
    const bf16x4* src = (const bf16x4*)w.packed;
    for (int i = threadIdx.x; i < kMlpOperands * 64; i += kWsBlock) ops[i] = src[i];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint16_t* tr = trs[wave];
    MlpBias b;
    load_bias(w.master, lane, b);
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
    f32x4 gW1[4], gW2[16], gW3[4];
    for (int k = 0; k < 4; k++) gW1[k] = gW3[k] = zero;
    for (int k = 0; k < 16; k++) gW2[k] = zero;
    float gB1[4][4], gB2[4][4], gB3[4], loss = 0.f;
    for (int mt = 0; mt < 4; mt++)
        for (int j = 0; j < 4; j++) gB1[mt][j] = gB2[mt][j] = 0.f;
    for (int j = 0; j < 4; j++) gB3[j] = 0.f;

    const int groups = (n + 63) >> 6;  // four tiles per block pass, one per wave
    for (int g = blockIdx.x; g < groups; g += gridDim.x) {
        const int v = (g * 4 + wave) * 16 + (lane & 15);
        const bf16x4 xb = load_x(X, v, n, lane);
        MlpAct a;
        tile_forward(ops, lane, b, xb, a);
        // dL/dy for the lanes that hold outputs (lanes 0-15, rows 0-2)
        float dy[4] = {0.f, 0.f, 0.f, 0.f};
        if (lane < 16 && v < n) {
            const float4 tg = T[v];
            const float t3[3] = {tg.x, tg.y, tg.z};
            for (int j = 0; j < 3; j++) {
                const float e = a.y[j] - t3[j];
                dy[j] = tg.w * 2.f * e * invCount;
                loss += tg.w * e * e;
                gB3[j] += dy[j];
            }
        }
        const bf16x4 dyb = pack4(dy[0], dy[1], dy[2], dy[3]);
        bf16x4 dz2b[4], dz1b[4];
        for (int mt = 0; mt < 4; mt++) {
            const f32x4 d = mfma(ops[(24 + mt) * 64 + lane], dyb, zero);
            float z[4];
            for (int j = 0; j < 4; j++) {
                z[j] = d[j] * gelu_grad(a.z2[mt][j]);
                gB2[mt][j] += z[j];
            }
            dz2b[mt] = pack4(z[0], z[1], z[2], z[3]);
        }
        for (int mt = 0; mt < 4; mt++) {
            f32x4 d = zero;
            for (int kb = 0; kb < 4; kb++) d = mfma(ops[(28 + mt * 4 + kb) * 64 + lane], dz2b[kb], d);
            float z[4];
            for (int j = 0; j < 4; j++) {
                z[j] = d[j] * gelu_grad(a.z1[mt][j]);
                gB1[mt][j] += z[j];
            }
            dz1b[mt] = pack4(z[0], z[1], z[2], z[3]);
        }
        // transpose the operands of the weight gradients through LDS
        tr_store(tr, kTrX, lane, xb);
        tr_store(tr, kTrDy, lane, dyb);
        for (int mt = 0; mt < 4; mt++) {
            tr_store(tr, kTrH1 + 16 * mt, lane, a.h1[mt]);
            tr_store(tr, kTrH2 + 16 * mt, lane, a.h2[mt]);
            tr_store(tr, kTrDz2 + 16 * mt, lane, dz2b[mt]);
            tr_store(tr, kTrDz1 + 16 * mt, lane, dz1b[mt]);
        }
        __syncthreads();
        const bf16x4 xa = tr_load(tr, kTrX, lane);
        for (int nt = 0; nt < 4; nt++) gW1[nt] = mfma(xa, tr_load(tr, kTrDz1 + 16 * nt, lane), gW1[nt]);
        for (int mt = 0; mt < 4; mt++) {
            const bf16x4 ha = tr_load(tr, kTrH1 + 16 * mt, lane);
            for (int nt = 0; nt < 4; nt++) gW2[mt * 4 + nt] = mfma(ha, tr_load(tr, kTrDz2 + 16 * nt, lane), gW2[mt * 4 + nt]);
        }
        const bf16x4 dyB = tr_load(tr, kTrDy, lane);
        for (int mt = 0; mt < 4; mt++) gW3[mt] = mfma(tr_load(tr, kTrH2 + 16 * mt, lane), dyB, gW3[mt]);
        __syncthreads();
    }

    // block sum in wave order (deterministic), then one partial per block
    const int r0 = 4 * (lane >> 4), c0 = lane & 15;
    for (int mt = 0; mt < 4; mt++)
        for (int j = 0; j < 4; j++) {
            gB1[mt][j] = sum16(gB1[mt][j]);
            gB2[mt][j] = sum16(gB2[mt][j]);
        }
    for (int j = 0; j < 4; j++) gB3[j] = sum16(gB3[j]);
    for (int o = 1; o < 64; o <<= 1) loss += __shfl_xor(loss, o, 64);
    for (int wv = 0; wv < kWsBlock / 64; wv++) {
        if (wave == wv) {
            const bool first = wv == 0;
            for (int nt = 0; nt < 4; nt++)
                for (int j = 0; j < 4; j++) {
                    const int in = r0 + j;
                    if (in < kMlpIn) {
                        float& d = acc[kOffW1 + in * 64 + 16 * nt + c0];
                        d = first ? gW1[nt][j] : d + gW1[nt][j];
                    }
                }
            for (int mt = 0; mt < 4; mt++)
                for (int nt = 0; nt < 4; nt++)
                    for (int j = 0; j < 4; j++) {
                        float& d = acc[kOffW2 + (16 * mt + r0 + j) * 64 + 16 * nt + c0];
                        d = first ? gW2[mt * 4 + nt][j] : d + gW2[mt * 4 + nt][j];
                    }
            if (c0 < kMlpOut)
                for (int mt = 0; mt < 4; mt++)
                    for (int j = 0; j < 4; j++) {
                        float& d = acc[kOffW3 + (16 * mt + r0 + j) * 3 + c0];
                        d = first ? gW3[mt][j] : d + gW3[mt][j];
                    }
            if (c0 == 0) {
                for (int mt = 0; mt < 4; mt++)
                    for (int j = 0; j < 4; j++) {
                        float& d1 = acc[kOffB1 + 16 * mt + r0 + j];
                        d1 = first ? gB1[mt][j] : d1 + gB1[mt][j];
                        float& d2 = acc[kOffB2 + 16 * mt + r0 + j];
                        d2 = first ? gB2[mt][j] : d2 + gB2[mt][j];
                    }
                if (lane == 0) {
                    for (int j = 0; j < kMlpOut; j++) {
                        float& d = acc[kOffB3 + j];
                        d = first ? gB3[j] : d + gB3[j];
                    }
                    float& d = acc[kMlpParams];
                    d = first ? loss : d + loss;
                }
            }
        }
        __syncthreads();
    }
    float* out = partial + (size_t)blockIdx.x * kMlpStride;
    for (int i = threadIdx.x; i < kMlpStride; i += kWsBlock) out[i] = acc[i];
}

// 64 parameters per workgroup, 16 slices of the blocks per parameter: each
// slice sums its blocks in order, then the slices are summed in order.
__global__ void __launch_bounds__(1024) k_mlp_reduce(const float* __restrict__ partial, int blocks,
                                                     float* __restrict__ grad)
{
    __shared__ float red[16][64];
    const int pl = threadIdx.x & 63, sl = threadIdx.x >> 6;
    const int p = blockIdx.x * 64 + pl;
    float s = 0.f;
    if (p < kMlpStride)
        for (int k = sl; k < blocks; k += 16) s += partial[(size_t)k * kMlpStride + p];
    red[sl][pl] = s;
    __syncthreads();
    if (sl == 0 && p < kMlpStride) {
        float t = red[0][pl];
        for (int k = 1; k < 16; k++) t += red[k][pl];
        grad[p] = t;
    }
}

// Adam (Kingma & Ba 2015), beta1 0.9, beta2 0.999, eps 1e-8; step t = *step + 1.
__global__ void k_mlp_adam(float* __restrict__ W, float* __restrict__ m, float* __restrict__ v,
                           const float* __restrict__ grad, const int* __restrict__ step, const float* __restrict__ lrp)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    const float lr = *lrp;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float t = (float)(*step + 1);
    const float g = grad[p];
    const float mn = b1 * m[p] + (1.f - b1) * g;
    const float vn = b2 * v[p] + (1.f - b2) * g * g;
    m[p] = mn;
    v[p] = vn;
    const float mh = mn / (1.f - powf(b1, t));
    const float vh = vn / (1.f - powf(b2, t));
    W[p] = W[p] - lr * mh / (sqrtf(vh) + eps);
}

// ------------------------------------------------------------ residual ----

// warp^2 + weft^2 (x) and shear^2 (y) of one triangle in FP64
__device__ __forceinline__ double2 tri_residual2(const float4* __restrict__ pos, TriStreams tr, int t)
{
    const int4 id = tr.idx[t];
    const float4 A = tr.a[t], B = tr.b[t];
    const double du1 = A.x, du2 = A.y, dv1 = A.z, dv2 = A.w, invDet = B.x, lu2 = B.y, lv2 = B.z;
    const float4 P1 = pos[id.x], P2 = pos[id.y], P3 = pos[id.z];
    const double a1 = (double)P2.x - P1.x, a2 = (double)P2.y - P1.y, a3 = (double)P2.z - P1.z;
    const double e1 = (double)P3.x - P1.x, e2 = (double)P3.y - P1.y, e3 = (double)P3.z - P1.z;
    const double u1 = invDet * (dv2 * a1 - dv1 * e1), u2 = invDet * (dv2 * a2 - dv1 * e2),
                 u3 = invDet * (dv2 * a3 - dv1 * e3);
    const double w1 = invDet * (-du2 * a1 + du1 * e1), w2 = invDet * (-du2 * a2 + du1 * e2),
                 w3 = invDet * (-du2 * a3 + du1 * e3);
    const double warp = 0.5 * ((u1 * u1 + u2 * u2 + u3 * u3) - lu2);
    const double weft = 0.5 * ((w1 * w1 + w2 * w2 + w3 * w3) - lv2);
    const double shear = u1 * w1 + u2 * w2 + u3 * w3;
    return make_double2(warp * warp + weft * weft, shear * shear);
}

__device__ __forceinline__ double bend_residual2(const float4* __restrict__ pos, const int4 b, double sStiff)
{
    const float4 P1 = pos[b.x], P2 = pos[b.y], P3 = pos[b.z], P4 = pos[b.w];
    const double x12[3] = {(double)P2.x - P1.x, (double)P2.y - P1.y, (double)P2.z - P1.z};
    const double x13[3] = {(double)P3.x - P1.x, (double)P3.y - P1.y, (double)P3.z - P1.z};
    const double x14[3] = {(double)P4.x - P1.x, (double)P4.y - P1.y, (double)P4.z - P1.z};
    const double n1[3] = {x12[1] * x13[2] - x12[2] * x13[1], x12[2] * x13[0] - x12[0] * x13[2],
                          x12[0] * x13[1] - x12[1] * x13[0]};
    const double n2[3] = {x14[1] * x12[2] - x14[2] * x12[1], x14[2] * x12[0] - x14[0] * x12[2],
                          x14[0] * x12[1] - x14[1] * x12[0]};
    const double l1 = sqrt(n1[0] * n1[0] + n1[1] * n1[1] + n1[2] * n1[2]);
    const double l2 = sqrt(n2[0] * n2[0] + n2[1] * n2[1] + n2[2] * n2[2]);
    if (l1 == 0.0 || l2 == 0.0) return 0.0;
    double c = (n1[0] * n2[0] + n1[1] * n2[1] + n1[2] * n2[2]) / (l1 * l2);
    c = c < -1.0 ? -1.0 : (c > 1.0 ? 1.0 : c);
    const double e = sStiff * acos(c);
    return e * e;
}

// Three sums per block (stretch, shear, bend), then three sums over blocks,
// both in a fixed order.
__device__ __forceinline__ void block_sum3(double a, double b, double c, double (*red)[kWsBlock], double* out)
{
    red[0][threadIdx.x] = a;
    red[1][threadIdx.x] = b;
    red[2][threadIdx.x] = c;
    __syncthreads();
    for (int o = kWsBlock / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            for (int k = 0; k < 3; k++) red[k][threadIdx.x] += red[k][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; k++) out[k] = red[k][0];
}

__global__ void __launch_bounds__(kWsBlock) k_residual(const float4* __restrict__ pos, TriStreams tr, int nt,
                                                       const int4* __restrict__ bidx, int nb, double sStiff,
                                                       double* __restrict__ partial)
{
    __shared__ double red[3][kWsBlock];
    double st = 0.0, sh = 0.0, sb = 0.0;
    const int items = nt + nb;
    for (int i = blockIdx.x * kWsBlock + threadIdx.x; i < items; i += gridDim.x * kWsBlock) {
        if (i < nt) {
            const double2 r = tri_residual2(pos, tr, i);
            st += r.x;
            sh += r.y;
        } else {
            sb += bend_residual2(pos, bidx[i - nt], sStiff);
        }
    }
    block_sum3(st, sh, sb, red, partial + 3 * blockIdx.x);
}

__global__ void __launch_bounds__(kWsBlock) k_residual_final(const double* __restrict__ partial, int blocks,
                                                             double* __restrict__ out)
{
    __shared__ double red[3][kWsBlock];
    double a = 0.0, b = 0.0, c = 0.0;
    for (int i = threadIdx.x; i < blocks; i += kWsBlock) {
        a += partial[3 * i];
        b += partial[3 * i + 1];
        c += partial[3 * i + 2];
    }
    block_sum3(a, b, c, red, out);
}

// ----------------------------------------------------------- launchers ----

static inline int wblocks(long n) { return (int)((n + kWsBlock - 1) / kWsBlock); }

void launch_ws_tri_strain(hipStream_t s, const float4* pos, TriStreams tr, int nt, float4* out)
{
    if (nt <= 0) return;
    hipLaunchKernelGGL(k_ws_tri_strain, dim3(wblocks(nt)), dim3(kWsBlock), 0, s, pos, tr, nt, out);
}

void launch_ws_features(hipStream_t s, const float4* pos, const float4* prev, WarmBufs b, int nv, float invEll)
{
    if (nv <= 0) return;
    hipLaunchKernelGGL(k_ws_features, dim3(wblocks(nv)), dim3(kWsBlock), 0, s, pos, prev, b, nv, invEll);
}

static int forward_blocks(int n)
{
    const int tiles = (n + 15) / 16;
    int blocks = (tiles + 3) / 4;
    return blocks < 1024 ? (blocks < 1 ? 1 : blocks) : 1024;
}

void launch_mlp_apply(hipStream_t s, MlpWeights w, const uint16_t* X, int n, float4* pos, float outScale)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_mlp_forward<1>, dim3(forward_blocks(n)), dim3(kWsBlock), 0, s, w, X, n, (float*)nullptr, pos,
                       outScale);
}

void launch_mlp_forward(hipStream_t s, MlpWeights w, const uint16_t* X, int n, float* Y)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_mlp_forward<0>, dim3(forward_blocks(n)), dim3(kWsBlock), 0, s, w, X, n, Y,
                       (float4*)nullptr, 0.f);
}

void launch_ws_after(hipStream_t s, const float4* pos, WarmBufs b, int nv, int train, float targetScale)
{
    if (nv <= 0) return;
    hipLaunchKernelGGL(k_ws_after, dim3(wblocks(nv)), dim3(kWsBlock), 0, s, pos, b, nv, train, targetScale);
}

int mlp_backward_blocks(int n)
{
    const int groups = (n + 63) / 64;
    return groups < kMlpMaxBlocks ? (groups < 1 ? 1 : groups) : kMlpMaxBlocks;
}

void launch_mlp_backward(hipStream_t s, MlpWeights w, const uint16_t* X, const float4* T, int n, float invCount,
                         float* partial, int blocks)
{
    hipLaunchKernelGGL(k_mlp_backward, dim3(blocks), dim3(kWsBlock), 0, s, w, X, T, n, invCount, partial);
}

void launch_mlp_reduce(hipStream_t s, const float* partial, int blocks, float* grad)
{
    hipLaunchKernelGGL(k_mlp_reduce, dim3((kMlpStride + 63) / 64), dim3(1024), 0, s, partial, blocks, grad);
}

void launch_mlp_adam(hipStream_t s, float* master, float* m, float* v, const float* grad, const int* step,
                     const float* lr)
{
    hipLaunchKernelGGL(k_mlp_adam, dim3(wblocks(kMlpParams)), dim3(kWsBlock), 0, s, master, m, v, grad, step, lr);
}

void launch_mlp_pack(hipStream_t s, const float* master, uint16_t* packed, int* step)
{
    hipLaunchKernelGGL(k_mlp_pack, dim3(wblocks(kMlpOperands * 256)), dim3(kWsBlock), 0, s, master, packed, step);
}

void launch_pack_features(hipStream_t s, const float* Xf, int n, uint16_t* X)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_pack_features, dim3(wblocks((long)n * 16)), dim3(kWsBlock), 0, s, Xf, n, X);
}

// ------------------------------------------------- no-allocation forward ----
// The BF16 feature rows (16 values, 32 bytes each) of up to kNoAllocRows
// samples: 8 MiB per GPU, held until the process ends. Every row a call uses
// is written by k_pack_features before the forward pass reads it, so nothing
// needs clearing between calls. In host code `scratch` names a host-side
// stand-in; only the address from hipGetSymbolAddress is used.
static __device__ __align__(256) uint16_t scratch[(size_t)kNoAllocRows * 16];

enum { kNoAllocMaxDevices = 64 };
static std::mutex gScratchMutex;
static uint16_t* gScratchAddr[kNoAllocMaxDevices] = {};

// This GPU's copy of `scratch`, looked up once per device with it current.
static uint16_t* scratch_address()
{
    int dev = -1;
    HIP_OK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocMaxDevices) throw std::runtime_error("device index out of range for the scratch table");
    std::lock_guard<std::mutex> lock(gScratchMutex);
    if (!gScratchAddr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw std::runtime_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
        gScratchAddr[dev] = (uint16_t*)p;
    }
    return gScratchAddr[dev];
}

void launch_mlp_forward_noalloc(hipStream_t s, MlpWeights w, const float* Xf, int n, float* Y)
{
    if (n <= 0 || n > kNoAllocRows)
        throw std::runtime_error("the no-allocation forward pass takes 1 to " + std::to_string((int)kNoAllocRows) +
                                 " samples, got " + std::to_string(n));
    uint16_t* X = scratch_address();
    launch_pack_features(s, Xf, n, X);
    launch_mlp_forward(s, w, X, n, Y);
    HIP_OK(hipGetLastError());
}

void launch_unpack_features(hipStream_t s, const uint16_t* X, int n, float* Xf)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_unpack_features, dim3(wblocks((long)n * kMlpIn)), dim3(kWsBlock), 0, s, X, n, Xf);
}

int residual_blocks(int items)
{
    const int b = wblocks(items);
    return b < 1024 ? (b < 1 ? 1 : b) : 1024;
}

void launch_residual(hipStream_t s, const float4* pos, TriStreams tr, int nt, const int4* bidx, int nb,
                     double sStiff, double* partial, int blocks, double* out)
{
    hipLaunchKernelGGL(k_residual, dim3(blocks), dim3(kWsBlock), 0, s, pos, tr, nt, bidx, nb, sStiff, partial);
    hipLaunchKernelGGL(k_residual_final, dim3(1), dim3(kWsBlock), 0, s, partial, blocks, out);
}

}  // namespace sc
