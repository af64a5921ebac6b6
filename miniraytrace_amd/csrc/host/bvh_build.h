// Binned-SAH BVH2 builder for the HIP traversal kernels.
//
// Produces the reference's 32-byte node layout and node order
// (core/src/accel/Bvh.hpp:15-26, core/src/accel/Bvh.cpp:41-156): 32 bins along
// the widest centroid axis, traversal cost 1.0, intersection cost 1.5, leaf
// when count <= maxLeaf or the centroid extent is below 1e-8, median split
// when SAH prefers a leaf that would exceed 2*maxLeaf, children allocated as a
// pair (right = left + 1) in depth-first order.
//
// Unlike the reference (one thread per BVH), subtrees above a size threshold
// are built on separate threads; the tree is then written out in the
// reference's depth-first allocation order, so the node array is identical to
// a serial build.
#pragma once

#include <cstdint>
#include <vector>

namespace mrt {

struct BvhNode {
    float    lo[3];
    uint32_t leftOrPrim;  // internal: left child index; leaf: first primitive
    float    hi[3];
    uint32_t packed;      // (count << 2) | split axis; count == 0 -> internal
};
static_assert(sizeof(BvhNode) == 32, "BvhNode is 32 bytes");

struct Box3 {
    float lo[3] = { 3.402823466e+38f, 3.402823466e+38f, 3.402823466e+38f };
    float hi[3] = { -3.402823466e+38f, -3.402823466e+38f, -3.402823466e+38f };
    void grow(const float p[3]) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = (p[a] < lo[a]) ? p[a] : lo[a];
            hi[a] = (hi[a] < p[a]) ? p[a] : hi[a];
        }
    }
    void grow(const Box3& b) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = (b.lo[a] < lo[a]) ? b.lo[a] : lo[a];
            hi[a] = (hi[a] < b.hi[a]) ? b.hi[a] : hi[a];
        }
    }
    bool valid() const { return lo[0] <= hi[0]; }
    float area() const;
};

struct BvhBuild {
    std::vector<BvhNode>  nodes;      // node 0 = root; empty for zero primitives
    std::vector<uint32_t> primOrder;  // leaf primitive slot -> input primitive index
    Box3                  rootBounds;
    uint32_t              maxDepth = 0;
    uint32_t              leafCount = 0;
};

// bounds: one Box3 per primitive.  threads <= 0 picks the hardware count.
BvhBuild buildBvh(const std::vector<Box3>& bounds, uint32_t maxLeaf, int threads = 0);

} // namespace mrt
