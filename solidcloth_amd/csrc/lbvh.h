// Linear BVH over axis-aligned boxes, built on the GPU:
//   1. scene bounds of the box centroids (block reduction + atomics),
//   2. 30-bit Morton codes of the centroids,
//   3. stable LSD radix sort of (code, primitive id), four 8-bit passes; the
//      scatter ranks equal digits inside a wavefront with ballots and across
//      wavefronts with LDS prefix sums, so the order is deterministic,
//   4. Karras (HPG 2012) hierarchy: internal node i covers a key range found by
//      binary search on common-prefix lengths (ties broken by index),
//   5. bottom-up refit: the second child to arrive at a node (atomic counter)
//      writes its box and continues upward; nodes inside one workgroup's
//      chunk of leaves synchronise at workgroup scope, the few whose leaf
//      range crosses chunks through device-scope atomic min/max words.
// Internal node k stores (min, left child) and (max, right child); a child
// c >= 0 is internal node c, c < 0 is leaf ~c, i.e. sorted position ~c.
#pragma once

#include <hip/hip_runtime.h>

namespace sc {

struct BvhView {
    const float4* lo;     // internal nodes: xyz = box min, w = left child (int bits)
    const float4* hi;     // internal nodes: xyz = box max, w = right child
    const float4* leafLo; // leaf boxes in sorted order
    const float4* leafHi;
    const int* leafPrim;  // sorted position -> primitive id
    // escape ("rope") of each internal node and leaf: the node a depth-first,
    // left-first walk visits after the subtree (internal c >= 0, leaf ~c,
    // kRopeEnd after the last); with them an overlap query needs no stack
    const int* ropeNode;
    const int* ropeLeaf;
    const int2* range;    // internal nodes: first and last sorted leaf below
    // The same hierarchy packed for the walks: all 2n - 1 nodes and leaves in
    // depth-first, left-first order, 64 bytes each (4 float4). Record p:
    // [0] = (min, primitive; -1 for an internal node), [1] = (max, escape
    // position: p + 1 for a leaf, p + its subtree's size for an internal
    // node, 2n - 1 past the last), [2] = the leaf's payload (int4, set with
    // Lbvh::set_payload; 0 for internal nodes), [3] = (sorted leaf index of a
    // leaf / last sorted leaf below an internal node, 0, 0, 0), int bits.
    // A walk that enters a node goes to p + 1, so it moves forward through
    // the array, mostly one record at a time.
    const float4* dfs;
    int n;                // number of primitives (leaves)
};

constexpr int kRopeEnd = -0x7fffffff - 1;

class Lbvh {
public:
    ~Lbvh();
    // Builds over n boxes (device arrays, xyz used). Asynchronous on `s`;
    // allocates on the first build and when n grows.
    void build(hipStream_t s, const float4* boxLo, const float4* boxHi, int n);
    // New boxes for the same primitives, same hierarchy (asynchronous).
    void refit(hipStream_t s, const float4* boxLo, const float4* boxHi);
    BvhView view() const;
    int size() const { return n_; }
    // int4 per primitive (by primitive id) copied into each leaf's record by
    // every build and refit from then on; the array must outlive the tree's
    // use (null: zeros).
    void set_payload(const int4* perPrimitive) { payload_ = perPrimitive; }
    // Downloads the hierarchy and checks it on the host: every leaf appears
    // once, parent links agree with child links, node boxes equal the union
    // of their children. Returns the number of violations (synchronous).
    int verify(hipStream_t s, const float4* boxLo, const float4* boxHi) const;

private:
    void reserve(int n);
    void pack(hipStream_t s);
    int n_ = 0, cap_ = 0, tiles_ = 0;
    float4 *lo_ = nullptr, *hi_ = nullptr, *leafLo_ = nullptr, *leafHi_ = nullptr;
    int* leafPrim_ = nullptr;
    unsigned *keys_ = nullptr, *keys2_ = nullptr;
    int *vals_ = nullptr, *vals2_ = nullptr;
    unsigned* hist_ = nullptr;
    int *parentNode_ = nullptr, *parentLeaf_ = nullptr, *flags_ = nullptr;
    unsigned* bounds_ = nullptr;  // 6 orderable-encoded floats
    int2* range_ = nullptr;       // sorted-leaf range of each internal node
    unsigned* enc_ = nullptr;     // per internal node: order-encoded min (3) / max (3) words
    int *ropeNode_ = nullptr, *ropeLeaf_ = nullptr;
    int *dfsNode_ = nullptr, *dfsLeaf_ = nullptr;  // depth-first positions of nodes and leaves
    float4* dfs_ = nullptr;                          // packed records (BvhView::dfs)
    const int4* payload_ = nullptr;
};

}  // namespace sc
