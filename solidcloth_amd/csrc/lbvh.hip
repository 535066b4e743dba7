// GPU LBVH builder: Morton codes, deterministic stable radix sort, Karras
// hierarchy and bottom-up refit.
#include <algorithm>
#include <vector>

#include "hiputil.h"
#include "lbvh.h"

namespace sc {

namespace {

constexpr int kBlock = 256;
constexpr int kItems = 8;                  // keys per thread per sort tile
constexpr int kTile = kBlock * kItems;     // keys per sort tile
constexpr int kScanBlock = 1024;
constexpr int kRefitBlock = 1024;          // leaves per refit chunk
constexpr float FLT_MAX_F = 3.402823466e+38f;

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

// Bounds of the box centroids: grid-stride loop, wave and workgroup
// reductions, one atomic per workgroup and axis.
__global__ void __launch_bounds__(kBlock) k_bounds(const float4* __restrict__ lo, const float4* __restrict__ hi, int n,
                                                   unsigned* b)
{
    __shared__ unsigned red[kBlock / 64][6];
    unsigned mn[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, mx[3] = {0u, 0u, 0u};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float3 c = centroid(lo, hi, i);
        const unsigned e[3] = {ord_enc(c.x), ord_enc(c.y), ord_enc(c.z)};
#pragma unroll
        for (int a = 0; a < 3; a++) {
            mn[a] = min(mn[a], e[a]);
            mx[a] = max(mx[a], e[a]);
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
        for (int a = 0; a < 3; a++) {
            mn[a] = min(mn[a], (unsigned)__shfl_xor((int)mn[a], o, 64));
            mx[a] = max(mx[a], (unsigned)__shfl_xor((int)mx[a], o, 64));
        }
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int a = 0; a < 3; a++) {
            red[w][a] = mn[a];
            red[w][3 + a] = mx[a];
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int a = threadIdx.x;
        unsigned v = red[0][a];
        for (int k = 1; k < kBlock / 64; k++) v = a < 3 ? min(v, red[k][a]) : max(v, red[k][a]);
        if (a < 3) atomicMin(&b[a], v);
        else atomicMax(&b[a], v);
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

// Exclusive prefix sum of m counts (m a multiple of 4) in one workgroup: in
// chunks of four consecutive counts per thread, so that a chunk is one
// coalesced pass; each chunk is scanned inside the waves with DPP (row
// shifts, then row broadcasts), across them by an LDS prefix of the wave
// totals, and offset by the running total of the chunks before it. The next
// chunk's counts are requested before the current one is scanned.
__global__ void __launch_bounds__(kScanBlock) k_scan(unsigned* a, int m)
{
    constexpr int kWaves = kScanBlock / 64, kChunk = 4 * kScanBlock;
    __shared__ unsigned wsum[kWaves];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    auto load = [&](int base) {
        const int i = base + 4 * t;
        return i < m ? *(const uint4*)(a + i) : make_uint4(0u, 0u, 0u, 0u);
    };
    unsigned carry = 0;
    uint4 cur = load(0);
    for (int base = 0; base < m; base += kChunk) {
        const uint4 nxt = load(base + kChunk);
        const unsigned s0 = cur.x, s1 = s0 + cur.y, s2 = s1 + cur.z, s3 = s2 + cur.w;
        unsigned incl = s3;  // inclusive scan of the thread totals over the wave
        incl += (unsigned)__builtin_amdgcn_update_dpp(0, (int)incl, 0x111, 0xf, 0xf, false);  // row_shr:1
        incl += (unsigned)__builtin_amdgcn_update_dpp(0, (int)incl, 0x112, 0xf, 0xf, false);  // row_shr:2
        incl += (unsigned)__builtin_amdgcn_update_dpp(0, (int)incl, 0x114, 0xf, 0xf, false);  // row_shr:4
        incl += (unsigned)__builtin_amdgcn_update_dpp(0, (int)incl, 0x118, 0xf, 0xf, false);  // row_shr:8
        incl += (unsigned)__builtin_amdgcn_update_dpp(0, (int)incl, 0x142, 0xa, 0xf, false);  // row_bcast:15, rows 1, 3
        incl += (unsigned)__builtin_amdgcn_update_dpp(0, (int)incl, 0x143, 0xc, 0xf, false);  // row_bcast:31, rows 2, 3
        if (lane == 63) wsum[w] = incl;
        __syncthreads();
        unsigned before = 0, total = 0;
#pragma unroll
        for (int k = 0; k < kWaves; k++) {
            const unsigned x = wsum[k];
            before += k < w ? x : 0u;
            total += x;
        }
        const unsigned off = carry + before + incl - s3;
        const int i = base + 4 * t;
        if (i < m) *(uint4*)(a + i) = make_uint4(off, off + s0, off + s1, off + s2);
        carry += total;
        __syncthreads();  // (wsum is rewritten by the next chunk)
        cur = nxt;
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
            // (always below n for a correct histogram scan; the check keeps a
            // wrong one from writing outside the arrays)
            if (dst < (unsigned)n) {
                kout[dst] = key;
                vout[dst] = val;
            }
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
                                                   int* parentNode, int* parentLeaf, int* flags, int2* range,
                                                   unsigned* enc)
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
    range[i] = make_int2(min(i, j), max(i, j));
    unsigned* e = enc + 6 * (size_t)i;
    e[0] = e[1] = e[2] = 0xffffffffu;
    e[3] = e[4] = e[5] = 0u;
    if (i == 0) parentNode[0] = -1;
}

// Escape of every node (threads [0, n-1) internal nodes, [n-1, 2n-1) leaves):
// climb while the node is its parent's right child; the escape is the right
// sibling of the first ancestor-or-self that is a left child.
__global__ void __launch_bounds__(kBlock) k_ropes(const float4* __restrict__ lo, const float4* __restrict__ hi,
                                                  const int* __restrict__ parentNode,
                                                  const int* __restrict__ parentLeaf, int n, int* ropeNode,
                                                  int* ropeLeaf)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= 2 * n - 1) return;
    const bool leaf = x >= n - 1;
    const int l = x - (n - 1);
    int cur = leaf ? ~l : x;
    int p = leaf ? parentLeaf[l] : parentNode[x];
    int esc = kRopeEnd;
    while (p >= 0) {
        if (__float_as_int(lo[p].w) == cur) {
            esc = __float_as_int(hi[p].w);
            break;
        }
        cur = p;
        p = parentNode[p];
    }
    if (leaf) ropeLeaf[l] = esc;
    else ropeNode[x] = esc;
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

__device__ __forceinline__ void node_union(int p, float4* lo, float4* hi, const float4* leafLo,
                                           const float4* leafHi)
{
    const float4 L = lo[p], H = hi[p];
    float4 a0, b0, a1, b1;
    child_box(__float_as_int(L.w), lo, hi, leafLo, leafHi, a0, b0);
    child_box(__float_as_int(H.w), lo, hi, leafLo, leafHi, a1, b1);
    lo[p] = make_float4(fminf(a0.x, a1.x), fminf(a0.y, a1.y), fminf(a0.z, a1.z), L.w);
    hi[p] = make_float4(fmaxf(b0.x, b1.x), fmaxf(b0.y, b1.y), fmaxf(b0.z, b1.z), H.w);
}

// Bottom-up boxes. Each leaf thread walks towards the root; at every node the
// first arriving child stops and the second, which then holds both children's
// boxes, completes the node and goes on (arrival parity: two arrivals per node
// and pass, so the counters need no reset between passes).
// A workgroup owns kRefitBlock consecutive sorted leaves. A node whose leaf
// range lies inside that chunk is reached only by the workgroup's own threads:
// the box goes through global memory under workgroup-scope ordering, which
// costs nothing. A node whose range crosses chunks is completed through
// device-scope atomics only: each arrival merges its box into the node's
// order-encoded min/max words, waits until the merges are performed, then
// counts itself; the second to count reads the union back with a neutral
// merge. No cache write-back or invalidation is needed.
__device__ __forceinline__ void merge_box(unsigned* enc, float4& a, float4& b)
{
    const unsigned m0 = __hip_atomic_fetch_min(&enc[0], ord_enc(a.x), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned m1 = __hip_atomic_fetch_min(&enc[1], ord_enc(a.y), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned m2 = __hip_atomic_fetch_min(&enc[2], ord_enc(a.z), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned x0 = __hip_atomic_fetch_max(&enc[3], ord_enc(b.x), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned x1 = __hip_atomic_fetch_max(&enc[4], ord_enc(b.y), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned x2 = __hip_atomic_fetch_max(&enc[5], ord_enc(b.z), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    a.x = fminf(a.x, ord_dec(m0));
    a.y = fminf(a.y, ord_dec(m1));
    a.z = fminf(a.z, ord_dec(m2));
    b.x = fmaxf(b.x, ord_dec(x0));
    b.y = fmaxf(b.y, ord_dec(x1));
    b.z = fmaxf(b.z, ord_dec(x2));
}

__global__ void __launch_bounds__(kRefitBlock) k_refit(const float4* __restrict__ boxLo,
                                                       const float4* __restrict__ boxHi,
                                                       const int* __restrict__ sorted, int n, float4* lo, float4* hi,
                                                       float4* leafLo, float4* leafHi, int* leafPrim,
                                                       const int* __restrict__ parentNode,
                                                       const int* __restrict__ parentLeaf, int* flags,
                                                       const int2* __restrict__ range, unsigned* enc)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int prim = sorted[k];
    float4 a = boxLo[prim], b = boxHi[prim];
    leafLo[k] = make_float4(a.x, a.y, a.z, __int_as_float(prim));
    leafHi[k] = make_float4(b.x, b.y, b.z, 0.f);
    leafPrim[k] = prim;
    if (n == 1) return;
    const int c0 = blockIdx.x * blockDim.x, c1 = c0 + blockDim.x;
    int p = parentLeaf[k];
    while (p >= 0) {
        const int2 r = range[p];
        if (r.x < c0 || r.y >= c1) break;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        const int old = __hip_atomic_fetch_add(&flags[p], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if ((old & 1) == 0) return;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        node_union(p, lo, hi, leafLo, leafHi);
        a = lo[p];
        b = hi[p];
        p = parentNode[p];
    }
    // (a, b) is the box of the child through which this thread reaches p
    while (p >= 0) {
        unsigned* e = enc + 6 * (size_t)p;
        // Both arrivals merge first and count afterwards; the merges are
        // waited for before counting, so the second to count sees, in the
        // node's words, both boxes: it re-reads them with a neutral merge.
        merge_box(e, a, b);
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        const int old = __hip_atomic_fetch_add(&flags[p], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((old & 1) == 0) return;
        float4 ua = make_float4(FLT_MAX_F, FLT_MAX_F, FLT_MAX_F, 0.f);
        float4 ub = make_float4(-FLT_MAX_F, -FLT_MAX_F, -FLT_MAX_F, 0.f);
        merge_box(e, ua, ub);  // reads the union of both children
        a = ua;
        b = ub;
        const float4 L = lo[p], H = hi[p];
        lo[p] = make_float4(a.x, a.y, a.z, L.w);
        hi[p] = make_float4(b.x, b.y, b.z, H.w);
        // words back to the empty box for the next pass (no other arrival
        // touches this node again in this pass)
        e[0] = e[1] = e[2] = 0xffffffffu;
        e[3] = e[4] = e[5] = 0u;
        p = parentNode[p];
    }
}

// Depth-first (pre-order, left first) position of every node and leaf
// (threads [0, n - 1) internal nodes, [n - 1, 2n - 1) leaves): the nodes
// before v are its ancestors and, for each ancestor whose right subtree holds
// v, that ancestor's whole left subtree; those subtrees hold the `first`
// leaves left of v's range and 2 first - R nodes, where R counts such
// ancestors, so the position is 2 first + (the ancestors whose left subtree
// holds v).
__global__ void __launch_bounds__(kBlock) k_dfs_pos(const float4* __restrict__ lo, const int* __restrict__ parentNode,
                                                    const int* __restrict__ parentLeaf,
                                                    const int2* __restrict__ range, int n, int* dfsNode,
                                                    int* dfsLeaf)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= 2 * n - 1) return;
    if (n == 1) {
        dfsLeaf[0] = 0;
        return;
    }
    const bool leaf = x >= n - 1;
    const int l = x - (n - 1);
    int cur = leaf ? ~l : x;
    int p = leaf ? parentLeaf[l] : parentNode[x];
    int left = 0;
    while (p >= 0) {
        left += __float_as_int(lo[p].w) == cur;
        cur = p;
        p = parentNode[p];
    }
    const int first = leaf ? l : range[x].x;
    if (leaf) dfsLeaf[l] = 2 * first + left;
    else dfsNode[x] = 2 * first + left;
}

// The packed depth-first records (BvhView::dfs), written after every build
// and refit: threads [0, n - 1) internal nodes, [n - 1, 2n - 1) leaves.
__global__ void __launch_bounds__(kBlock) k_pack(const float4* __restrict__ lo, const float4* __restrict__ hi,
                                                 const float4* __restrict__ leafLo, const float4* __restrict__ leafHi,
                                                 const int* __restrict__ leafPrim, const int2* __restrict__ range,
                                                 const int* __restrict__ dfsNode, const int* __restrict__ dfsLeaf,
                                                 const int4* __restrict__ payload, int n, float4* dfs)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x < n - 1) {
        const int2 r = range[x];
        const int p = dfsNode[x];
        const float4 a = lo[x], b = hi[x];
        float4* R = dfs + 4 * (size_t)p;
        R[0] = make_float4(a.x, a.y, a.z, __int_as_float(-1));
        R[1] = make_float4(b.x, b.y, b.z, __int_as_float(p + 2 * (r.y - r.x) + 1));
        R[2] = make_float4(0.f, 0.f, 0.f, 0.f);
        R[3] = make_float4(__int_as_float(r.y), 0.f, 0.f, 0.f);
    } else if (x < 2 * n - 1) {
        const int k = x - (n - 1);
        const int p = dfsLeaf[k];
        const int prim = leafPrim[k];
        const float4 a = leafLo[k], b = leafHi[k];
        const int4 q = payload ? payload[prim] : make_int4(0, 0, 0, 0);
        float4* R = dfs + 4 * (size_t)p;
        R[0] = make_float4(a.x, a.y, a.z, __int_as_float(prim));
        R[1] = make_float4(b.x, b.y, b.z, __int_as_float(p + 1));
        R[2] = make_float4(__int_as_float(q.x), __int_as_float(q.y), __int_as_float(q.z), __int_as_float(q.w));
        R[3] = make_float4(__int_as_float(k), 0.f, 0.f, 0.f);
    }
}

inline int nblocks(int n, int b) { return (n + b - 1) / b; }

}  // namespace

Lbvh::~Lbvh()
{
    dev_free(lo_); dev_free(hi_); dev_free(leafLo_); dev_free(leafHi_); dev_free(leafPrim_);
    dev_free(keys_); dev_free(keys2_); dev_free(vals_); dev_free(vals2_); dev_free(hist_);
    dev_free(parentNode_); dev_free(parentLeaf_); dev_free(flags_); dev_free(bounds_); dev_free(range_); dev_free(enc_);
    dev_free(ropeNode_); dev_free(ropeLeaf_); dev_free(dfsNode_); dev_free(dfsLeaf_); dev_free(dfs_);
}

void Lbvh::reserve(int n)
{
    if (n <= cap_) return;
    const int tiles = nblocks(n, kTile);
    dev_alloc(lo_, n); dev_alloc(hi_, n); dev_alloc(leafLo_, n); dev_alloc(leafHi_, n); dev_alloc(leafPrim_, n);
    dev_alloc(keys_, n); dev_alloc(keys2_, n); dev_alloc(vals_, n); dev_alloc(vals2_, n);
    dev_alloc(hist_, (size_t)256 * tiles);
    dev_alloc(parentNode_, n); dev_alloc(parentLeaf_, n); dev_alloc(flags_, n); dev_alloc(range_, n);
    dev_alloc(enc_, (size_t)6 * n);
    dev_alloc(ropeNode_, n); dev_alloc(ropeLeaf_, n);
    dev_alloc(dfsNode_, n); dev_alloc(dfsLeaf_, n); dev_alloc(dfs_, (size_t)8 * n);
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
    hipLaunchKernelGGL(k_bounds, dim3(nb < 256 ? nb : 256), dim3(kBlock), 0, s, boxLo, boxHi, n, bounds_);
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
                           parentLeaf_, flags_, range_, enc_);
    hipLaunchKernelGGL(k_refit, dim3(nblocks(n, kRefitBlock)), dim3(kRefitBlock), 0, s, boxLo, boxHi, vi, n, lo_, hi_,
                       leafLo_, leafHi_, leafPrim_, parentNode_, parentLeaf_, flags_, range_, enc_);
    if (n > 1)
        hipLaunchKernelGGL(k_ropes, dim3(nblocks(2 * n - 1, kBlock)), dim3(kBlock), 0, s, lo_, hi_, parentNode_,
                           parentLeaf_, n, ropeNode_, ropeLeaf_);
    hipLaunchKernelGGL(k_dfs_pos, dim3(nblocks(2 * n - 1, kBlock)), dim3(kBlock), 0, s, lo_, parentNode_, parentLeaf_,
                       range_, n, dfsNode_, dfsLeaf_);
    pack(s);
    HIP_OK(hipGetLastError());
}

void Lbvh::pack(hipStream_t s)
{
    if (n_ < 1) return;
    hipLaunchKernelGGL(k_pack, dim3(nblocks(2 * n_ - 1, kBlock)), dim3(kBlock), 0, s, lo_, hi_, leafLo_, leafHi_,
                       leafPrim_, range_, dfsNode_, dfsLeaf_, payload_, n_, dfs_);
}

void Lbvh::refit(hipStream_t s, const float4* boxLo, const float4* boxHi)
{
    if (n_ == 0) return;
    // the sorted primitive order of the last build lives in vals_
    hipLaunchKernelGGL(k_refit, dim3(nblocks(n_, kRefitBlock)), dim3(kRefitBlock), 0, s, boxLo, boxHi, vals_, n_,
                       lo_, hi_, leafLo_, leafHi_, leafPrim_, parentNode_, parentLeaf_, flags_, range_, enc_);
    pack(s);
}

int Lbvh::verify(hipStream_t s, const float4* boxLo, const float4* boxHi) const
{
    HIP_OK(hipStreamSynchronize(s));
    const int n = n_;
    if (n == 0) return 0;
    std::vector<float4> lo(n > 1 ? n - 1 : 0), hi(lo.size()), llo(n), lhi(n), blo(n), bhi(n);
    std::vector<int> prim(n), pn(n), pl(n);
    auto get = [](auto& dst, const void* src) {
        if (!dst.empty()) HIP_OK(hipMemcpy(dst.data(), src, dst.size() * sizeof(dst[0]), hipMemcpyDeviceToHost));
    };
    get(lo, lo_); get(hi, hi_); get(llo, leafLo_); get(lhi, leafHi_); get(prim, leafPrim_);
    get(blo, boxLo); get(bhi, boxHi);
    if (n > 1) { get(pn, parentNode_); get(pl, parentLeaf_); }
    int bad = 0;
    std::vector<int> seen(n, 0);
    for (int k = 0; k < n; k++) {
        const int p = prim[k];
        if (p < 0 || p >= n || seen[p]++) { bad++; continue; }
        const float4 a = blo[p], b = bhi[p];
        if (llo[k].x != a.x || llo[k].y != a.y || llo[k].z != a.z || lhi[k].x != b.x || lhi[k].y != b.y ||
            lhi[k].z != b.z)
            bad++;
    }
    if (n == 1) return bad;
    std::vector<int> visits(n, 0);
    for (int i = 0; i < n - 1; i++) {
        float4 cl[2], ch[2];
        for (int c = 0; c < 2; c++) {
            const int e = __builtin_bit_cast(int, c == 0 ? lo[i].w : hi[i].w);
            if (e < 0) {
                const int leaf = ~e;
                if (leaf >= n || pl[leaf] != i) { bad++; continue; }
                visits[leaf]++;
                cl[c] = llo[leaf]; ch[c] = lhi[leaf];
            } else {
                if (e >= n - 1 || pn[e] != i) { bad++; continue; }
                cl[c] = lo[e]; ch[c] = hi[e];
            }
        }
        if (bad) continue;
        if (lo[i].x != std::min(cl[0].x, cl[1].x) || lo[i].y != std::min(cl[0].y, cl[1].y) ||
            lo[i].z != std::min(cl[0].z, cl[1].z) || hi[i].x != std::max(ch[0].x, ch[1].x) ||
            hi[i].y != std::max(ch[0].y, ch[1].y) || hi[i].z != std::max(ch[0].z, ch[1].z))
            bad++;
    }
    for (int k = 0; k < n; k++) bad += visits[k] != 1;
    if (pn[0] != -1) bad++;
    if (bad) return bad;
    // escape links against a depth-first, left-first walk on the host
    std::vector<int> rn(n - 1), rl(n);
    get(rn, ropeNode_);
    get(rl, ropeLeaf_);
    std::vector<int> order;  // tagged ids in walk order
    order.reserve(2 * n - 1);
    std::vector<int> stack{0};
    while (!stack.empty()) {
        const int x = stack.back();
        stack.pop_back();
        order.push_back(x);
        if (x >= 0) {
            stack.push_back(__builtin_bit_cast(int, hi[x].w));
            stack.push_back(__builtin_bit_cast(int, lo[x].w));
        }
    }
    // escape(x) = the first node after x's subtree in that order
    std::vector<int> sizeInt(n - 1, 0);
    for (int k = (int)order.size() - 1; k >= 0; k--) {
        const int x = order[k];
        if (x < 0) continue;
        int sz = 1;
        for (const int c : {__builtin_bit_cast(int, lo[x].w), __builtin_bit_cast(int, hi[x].w)})
            sz += c < 0 ? 1 : sizeInt[c];
        sizeInt[x] = sz;
    }
    for (int k = 0; k < (int)order.size(); k++) {
        const int x = order[k];
        const int after = k + (x < 0 ? 1 : sizeInt[x]);
        const int expect = after < (int)order.size() ? order[after] : kRopeEnd;
        if ((x < 0 ? rl[~x] : rn[x]) != expect) bad++;
    }
    // the packed records: the walk order's k-th node at position k, its
    // escape k + subtree size, its box, primitive and leaf fields
    std::vector<int> dn(n - 1), dl(n);
    std::vector<float4> rec(4 * (size_t)(2 * n - 1));
    get(dn, dfsNode_);
    get(dl, dfsLeaf_);
    get(rec, dfs_);
    for (int k = 0; k < (int)order.size(); k++) {
        const int x = order[k];
        if ((x < 0 ? dl[~x] : dn[x]) != k) { bad++; continue; }
        const float4* R = rec.data() + 4 * (size_t)k;
        const int esc = k + (x < 0 ? 1 : sizeInt[x]);
        const float4 a = x < 0 ? llo[~x] : lo[x], b = x < 0 ? lhi[~x] : hi[x];
        if (R[0].x != a.x || R[0].y != a.y || R[0].z != a.z || R[1].x != b.x || R[1].y != b.y || R[1].z != b.z ||
            __builtin_bit_cast(int, R[1].w) != esc ||
            __builtin_bit_cast(int, R[0].w) != (x < 0 ? prim[~x] : -1))
            bad++;
    }
    return bad;
}

BvhView Lbvh::view() const
{
    return BvhView{lo_, hi_, leafLo_, leafHi_, leafPrim_, ropeNode_, ropeLeaf_, range_, dfs_, n_};
}

}  // namespace sc
