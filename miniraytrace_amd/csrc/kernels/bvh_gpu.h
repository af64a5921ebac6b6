// Records and launchers of the GPU binned-SAH BVH builder (bvh_gpu.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mrt {
namespace dev {

constexpr uint32_t kSmallMax = 512;  // nodes with at most this many primitives: one wavefront builds the subtree
constexpr uint32_t kChunk = 1024;     // slots per chunk of a large node
constexpr uint32_t kSplitLeaf = 0xFFFFFFFFu;
constexpr uint32_t kSplitPending = 0xFFFFFFFEu;
constexpr uint32_t kBuildNeedsMedian = 1u;  // flag: a node needs the reference's median split
constexpr uint32_t kLargeAccWords = 18 + 32 * 7;  // accumulator words per large node (boxes, zero codes, 32 bins)

// The 32-byte node record (reference Bvh.hpp:15-26).
struct OutNode {
    float lo[3];
    uint32_t leftOrPrim;
    float hi[3];
    uint32_t packed;
};

// A large node of the current level: its slot range and, once decided, its
// box, split axis and bin mapping, and the split (bin index, kSplitLeaf) with
// the number of its primitives that go left.
struct LargeNode {
    uint32_t begin, count;
    float lo[3], hi[3];
    uint32_t axis;
    float k, origin;
    uint32_t split, mid;
    uint32_t chunk0, chunks;  // this node's chunks in the level's chunk list
};

struct LargeChunk {
    uint32_t node, begin, end;  // slot range [begin, end) of node `node`
};

// A node of at most kSmallMax primitives whose subtree one workgroup builds.
struct SmallTask {
    uint32_t begin, count;  // slot range
    uint32_t depth;         // depth of the subtree root in the whole tree
    uint32_t outBase;       // where its local node records go in the temporary array
};
struct SmallResult {
    uint32_t nodes, maxDepth, leaves;
};
struct SmallPlace {
    uint32_t root;  // final index of the subtree root
    uint32_t base;  // final index of its first child pair
};

struct LargeLevelArgs {
    const float* bounds;  // 6 floats per primitive (lo xyz, hi xyz)
    uint32_t* idx;        // slot -> primitive
    LargeNode* nodes;
    uint32_t nodeCount;
    const LargeChunk* chunks;
    uint32_t chunkCount;
    int* acc;             // per node accumulators
    uint32_t* counts;     // 2 per chunk
    uint32_t* totals;     // 2 per node: right-side, then left-side slots changing side
    uint32_t* pos;        // one per slot
    uint32_t maxLeaf;
    uint32_t* flags;
};

struct SmallArgs {
    const float* bounds;
    uint32_t* idx;
    const SmallTask* tasks;
    uint32_t taskCount;
    OutNode* tmp;
    SmallResult* results;
    uint32_t maxLeaf;
    uint32_t* flags;
};

struct PlaceArgs {
    const OutNode* large;       // final records of the large nodes
    const uint32_t* largeWhere; // their final indices
    uint32_t largeCount;
    const OutNode* tmp;
    const SmallTask* tasks;
    const SmallResult* results;
    const SmallPlace* place;
    uint32_t taskCount;
    OutNode* dst;
};

// idx = 0..n-1, flags = 0.
hipError_t launchBvhIota(uint32_t* idx, uint32_t n, uint32_t* flags, hipStream_t s);
// Decides and partitions every large node of one level.
hipError_t launchBvhLargeLevel(const LargeLevelArgs& a, hipStream_t s);
// Builds every small subtree.
hipError_t launchBvhSmall(const SmallArgs& a, hipStream_t s);
// Writes the final node array.
hipError_t launchBvhPlace(const PlaceArgs& a, hipStream_t s);

} // namespace dev
} // namespace mrt
