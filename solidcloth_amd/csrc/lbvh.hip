// GPU LBVH builder: Morton codes, deterministic stable radix sort, Karras
// hierarchy and bottom-up refit.
#include "hiputil.h"
#include "lbvh.h"

namespace sc {

namespace {

constexpr int kBlock = 256;
constexpr int kItems = 8;                  // keys per thread per sort tile
constexpr int kTile = kBlock * kItems;     // keys per sort tile
constexpr int kScanBlock = 1024;

// float <-> unsigned with the same order, for atomicMin/atomicMax on floats
__device__ __forceinline__ unsigned ord_enc(float f)
{
    unsigned u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord_dec(unsigned u)
{
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

__device__ __forceinline__ float3 centroid(const float4* lo, const float4* hi, int i)
{
    const float4 a = lo[i], b = hi[i];
    return make_float3(0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z));
}

__global__ void k_bounds_init(unsigned* b)
{
    const int t = threadIdx.x;
    if (t < 3) b[t] = 0xffffffffu;
    else if (t < 6) b[t] = 0u;
}

// Bounds of the box centroids: wave reduction, one atomic per wave and axis.
__global__ void __launch_bounds__(kBlock) k_bounds(const float4* __restrict__ lo, const float4* __restrict__ hi, int n,
                                                   unsigned* b)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned mn[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, mx[3] = {0u, 0u, 0u};
    if (i < n) {
        const float3 c = centroid(lo, hi, i);
        mn[0] = mx[0] = ord_enc(c.x);
        mn[1] = mx[1] = ord_enc(c.y);
        mn[2] = mx[2] = ord_enc(c.z);
    }
    for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
        for (int a = 0; a < 3; a++) {
            mn[a] = min(mn[a], (unsigned)__shfl_xor((int)mn[a], o, 64));
            mx[a] = max(mx[a], (unsigned)__shfl_xor((int)mx[a], o, 64));
        }
    }
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int a = 0; a < 3; a++) {
            atomicMin(&b[a], mn[a]);
            atomicMax(&b[3 + a], mx[a]);
        }
    }
}

__device__ __forceinline__ unsigned spread10(unsigned v)
{
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__device__ __forceinline__ unsigned quant(float c, float lo, float ext)
{
    float t = ext > 0.f ? (c - lo) / ext : 0.f;
    t = fminf(fmaxf(t * 1024.f, 0.f), 1023.f);
    return (unsigned)t;
}

__global__ void __launch_bounds__(kBlock) k_morton(const float4* __restrict__ lo, const float4* __restrict__ hi, int n,
                                                   const unsigned* __restrict__ b, unsigned* keys, int* vals)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float bx = ord_dec(b[0]), by = ord_dec(b[1]), bz = ord_dec(b[2]);
    const float ex = ord_dec(b[3]) - bx, ey = ord_dec(b[4]) - by, ez = ord_dec(b[5]) - bz;
    const float3 c = centroid(lo, hi, i);
    keys[i] = (spread10(quant(c.x, bx, ex)) << 2) | (spread10(quant(c.y, by, ey)) << 1) | spread10(quant(c.z, bz, ez));
    vals[i] = i;
}

// Digit counts of one tile, stored digit-major: hist[digit * tiles + tile].
__global__ void __launch_bounds__(kBlock) k_hist(const unsigned* __restrict__ keys, int n, int shift, unsigned* hist,
                                                 int tiles)
{
    __shared__ unsigned cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * kTile;
    for (int r = 0; r < kItems; r++) {
        const int i = base + r * kBlock + threadIdx.x;
        if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * tiles + blockIdx.x] = cnt[threadIdx.x];
}

// Exclusive prefix sum of m counts in one workgroup.
__global__ void __launch_bounds__(kScanBlock) k_scan(unsigned* a, int m)
{
    __shared__ unsigned part[kScanBlock];
    const int per = (m + kScanBlock - 1) / kScanBlock;
    const int b = threadIdx.x * per, e = min(b + per, m);
    unsigned s = 0;
    for (int i = b; i < e; i++) s += a[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int o = 1; o < kScanBlock; o <<= 1) {
        unsigned v = threadIdx.x >= o ? part[threadIdx.x - o] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    unsigned run = part[threadIdx.x] - s;
    for (int i = b; i < e; i++) {
        const unsigned v = a[i];
        a[i] = run;
        run += v;
    }
}

// Stable scatter of one tile. Keys are taken in index order, 256 at a time;
// equal digits are ranked inside a wavefront by ballots over the digit's bits
// and across the workgroup's wavefronts by an LDS prefix per digit.
__global__ void __launch_bounds__(kBlock) k_scatter(const unsigned* __restrict__ kin, const int* __restrict__ vin,
                                                    unsigned* __restrict__ kout, int* __restrict__ vout, int n,
                                                    int shift, const unsigned* __restrict__ hist, int tiles)
{
    __shared__ unsigned base[256];
    __shared__ unsigned run[256];
    __shared__ unsigned wcnt[kBlock / 64][256];
    const int t = threadIdx.x, wave = t >> 6;
    base[t] = hist[t * tiles + blockIdx.x];
    run[t] = 0;
    const int tile0 = blockIdx.x * kTile;
    for (int r = 0; r < kItems; r++) {
#pragma unroll
        for (int w = 0; w < kBlock / 64; w++) wcnt[w][t] = 0;
        __syncthreads();
        const int i = tile0 + r * kBlock + t;
        const bool valid = i < n;
        const unsigned key = valid ? kin[i] : 0u;
        const int val = valid ? vin[i] : 0;
        const unsigned d = (key >> shift) & 255u;
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; bit++) {
            const bool on = (d >> bit) & 1u;
            const unsigned long long m = __ballot(on);
            peers &= on ? m : ~m;
        }
        const unsigned below = __builtin_amdgcn_mbcnt_hi((unsigned)(peers >> 32),
                                                         __builtin_amdgcn_mbcnt_lo((unsigned)peers, 0u));
        if (valid && below == 0) wcnt[wave][d] = (unsigned)__popcll(peers);
        __syncthreads();
        {
            unsigned acc = run[t];
#pragma unroll
            for (int w = 0; w < kBlock / 64; w++) {
                const unsigned c = wcnt[w][t];
                wcnt[w][t] = acc;
                acc += c;
            }
            run[t] = acc;
        }
        __syncthreads();
        if (valid) {
            const unsigned dst = base[d] + wcnt[wave][d] + below;
            kout[dst] = key;
            vout[dst] = val;
        }
        __syncthreads();
    }
}

// Common-prefix length of sorted keys i and j, with the index appended to
// break ties (Karras 2012, section 4); -1 outside [0, n).
__device__ __forceinline__ int delta(const unsigned* keys, int n, int i, int j)
{
    if (j < 0 || j >= n) return -1;
    const unsigned a = keys[i], b = keys[j];
    if (a == b) return 32 + __clz((unsigned)(i ^ j));
    return __clz(a ^ b);
}

__device__ __forceinline__ int enc_child(int c, bool leaf) { return leaf ? ~c : c; }

__global__ void __launch_bounds__(kBlock) k_karras(const unsigned* __restrict__ keys, int n, float4* lo, float4* hi,
                                                   int* parentNode, int* parentLeaf, int* flags)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const int d = delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1) > 0 ? 1 : -1;
    const int dmin = delta(keys, n, i, i - d);
    int lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = delta(keys, n, i, j);
    int s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int gamma = i + s * d + min(d, 0);
    const bool leftLeaf = min(i, j) == gamma, rightLeaf = max(i, j) == gamma + 1;
    lo[i] = make_float4(0.f, 0.f, 0.f, __int_as_float(enc_child(gamma, leftLeaf)));
    hi[i] = make_float4(0.f, 0.f, 0.f, __int_as_float(enc_child(gamma + 1, rightLeaf)));
    if (leftLeaf) parentLeaf[gamma] = i;
    else parentNode[gamma] = i;
    if (rightLeaf) parentLeaf[gamma + 1] = i;
    else parentNode[gamma + 1] = i;
    flags[i] = 0;
    if (i == 0) parentNode[0] = -1;
}

__device__ __forceinline__ void child_box(int c, const float4* lo, const float4* hi, const float4* leafLo,
                                          const float4* leafHi, float4& a, float4& b)
{
    if (c < 0) {
        a = leafLo[~c];
        b = leafHi[~c];
    } else {
        a = lo[c];
        b = hi[c];
    }
}

// Bottom-up boxes: each leaf thread walks towards the root; at every node the
// first arriving child stops and the second, which then sees both children's
// boxes, writes the node's box and goes on.
__global__ void __launch_bounds__(kBlock) k_refit(const float4* __restrict__ boxLo, const float4* __restrict__ boxHi,
                                                  const int* __restrict__ sorted, int n, float4* lo, float4* hi,
                                                  float4* leafLo, float4* leafHi, int* leafPrim,
                                                  const int* __restrict__ parentNode,
                                                  const int* __restrict__ parentLeaf, int* flags)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int prim = sorted[k];
    const float4 a = boxLo[prim], b = boxHi[prim];
    leafLo[k] = make_float4(a.x, a.y, a.z, __int_as_float(prim));
    leafHi[k] = make_float4(b.x, b.y, b.z, 0.f);
    leafPrim[k] = prim;
    if (n == 1) return;
    int p = parentLeaf[k];
    while (p >= 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const int old = __hip_atomic_fetch_add(&flags[p], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (old == 0) return;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        const float4 L = lo[p], H = hi[p];
        const int cl = __float_as_int(L.w), cr = __float_as_int(H.w);
        float4 a0, b0, a1, b1;
        child_box(cl, lo, hi, leafLo, leafHi, a0, b0);
        child_box(cr, lo, hi, leafLo, leafHi, a1, b1);
        lo[p] = make_float4(fminf(a0.x, a1.x), fminf(a0.y, a1.y), fminf(a0.z, a1.z), L.w);
        hi[p] = make_float4(fmaxf(b0.x, b1.x), fmaxf(b0.y, b1.y), fmaxf(b0.z, b1.z), H.w);
        p = parentNode[p];
    }
}

inline int nblocks(int n, int b) { return (n + b - 1) / b; }

}  // namespace

Lbvh::~Lbvh()
{
    dev_free(lo_); dev_free(hi_); dev_free(leafLo_); dev_free(leafHi_); dev_free(leafPrim_);
    dev_free(keys_); dev_free(keys2_); dev_free(vals_); dev_free(vals2_); dev_free(hist_);
    dev_free(parentNode_); dev_free(parentLeaf_); dev_free(flags_); dev_free(bounds_);
}

void Lbvh::reserve(int n)
{
    if (n <= cap_) return;
    const int tiles = nblocks(n, kTile);
    dev_alloc(lo_, n); dev_alloc(hi_, n); dev_alloc(leafLo_, n); dev_alloc(leafHi_, n); dev_alloc(leafPrim_, n);
    dev_alloc(keys_, n); dev_alloc(keys2_, n); dev_alloc(vals_, n); dev_alloc(vals2_, n);
    dev_alloc(hist_, (size_t)256 * tiles);
    dev_alloc(parentNode_, n); dev_alloc(parentLeaf_, n); dev_alloc(flags_, n);
    if (!bounds_) dev_alloc(bounds_, 6);
    cap_ = n;
}

void Lbvh::build(hipStream_t s, const float4* boxLo, const float4* boxHi, int n)
{
    if (n < 0) throw std::runtime_error("negative BVH size");
    n_ = n;
    if (n == 0) return;
    reserve(n);
    tiles_ = nblocks(n, kTile);
    const int nb = nblocks(n, kBlock);
    hipLaunchKernelGGL(k_bounds_init, dim3(1), dim3(64), 0, s, bounds_);
    hipLaunchKernelGGL(k_bounds, dim3(nb), dim3(kBlock), 0, s, boxLo, boxHi, n, bounds_);
    hipLaunchKernelGGL(k_morton, dim3(nb), dim3(kBlock), 0, s, boxLo, boxHi, n, bounds_, keys_, vals_);
    unsigned *ki = keys_, *ko = keys2_;
    int *vi = vals_, *vo = vals2_;
    for (int shift = 0; shift < 30; shift += 8) {
        hipLaunchKernelGGL(k_hist, dim3(tiles_), dim3(kBlock), 0, s, ki, n, shift, hist_, tiles_);
        hipLaunchKernelGGL(k_scan, dim3(1), dim3(kScanBlock), 0, s, hist_, 256 * tiles_);
        hipLaunchKernelGGL(k_scatter, dim3(tiles_), dim3(kBlock), 0, s, ki, vi, ko, vo, n, shift, hist_, tiles_);
        std::swap(ki, ko);
        std::swap(vi, vo);
    }
    // four passes: the sorted keys are back in keys_/vals_
    if (n > 1)
        hipLaunchKernelGGL(k_karras, dim3(nblocks(n - 1, kBlock)), dim3(kBlock), 0, s, ki, n, lo_, hi_, parentNode_,
                           parentLeaf_, flags_);
    hipLaunchKernelGGL(k_refit, dim3(nb), dim3(kBlock), 0, s, boxLo, boxHi, vi, n, lo_, hi_, leafLo_, leafHi_,
                       leafPrim_, parentNode_, parentLeaf_, flags_);
    HIP_OK(hipGetLastError());
}

BvhView Lbvh::view() const { return BvhView{lo_, hi_, leafLo_, leafHi_, leafPrim_, n_}; }

}  // namespace sc
