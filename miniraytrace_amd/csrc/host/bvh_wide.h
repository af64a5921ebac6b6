// Four-wide traversal BVH collapsed from the reference's binary BVH.
//
// The binary BVH (bvh_build.h, the reference's Bvh.cpp layout) stays the tree
// of record; traversal kernels may instead walk this collapsed copy.  Each
// wide node takes an inner binary node and replaces its inner children by
// their own children, the one with the largest surface area first, until it
// holds four children or only leaves.  Leaves keep their binary-tree
// primitive ranges, so the leaf-ordered triangle arrays serve both trees.
// Child boxes are the binary nodes' boxes, unchanged; box tests are
// conservative and ties in t resolve to the lowest triangle id, so a walk of
// this tree returns the binary walk's hits, bit for bit.
#pragma once

#include "bvh_build.h"

#include <cstdint>
#include <vector>

namespace mrt {

// Child codes (WideNode::code).
constexpr uint32_t kWideInner = 0;        // child = wide node
constexpr uint32_t kWideMaxLeaf = 29;     // 1..29: leaf of that many primitives, child = first primitive
constexpr uint32_t kWideBigLeaf = 0xFE;   // leaf of more: child = its binary node
constexpr uint32_t kWideEmpty = 0xFF;     // unused slot (its box is empty)

// 128 bytes, eight 16-byte rows: the four children's lo.x, lo.y, lo.z,
// hi.x, hi.y, hi.z, then their first node or primitive, then their codes.
// Row a (lower bound of axis a) and row 3 + a (upper bound) are the near
// and far planes of all four children for a ray whose direction on axis a
// is positive, the other way round when it is negative, so a kernel loads
// the planes it needs from offsets set once per ray.
struct WideNode {
    float bound[6][4];
    uint32_t child[4];
    uint32_t code[4];
};
static_assert(sizeof(WideNode) == 128, "wide node is 128 bytes");

// Root at index 0, nodes in depth-first order.  An empty tree gives one node
// with no children; a tree whose root is a leaf gives one node holding it.
std::vector<WideNode> collapseWide(const std::vector<BvhNode>& nodes);

} // namespace mrt
