// Host-callable launchers for the HIP kernels (implemented in *.hip files).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mrt {
namespace dev {

struct Node;

struct QueryLaunch {
    const Node* nodes;     // 32-byte BVH nodes, root at 0
    const float4* wide;    // the four-wide tree collapsed from them (8 float4 per node), root at 0
    const float4* tris;    // 3 float4 per leaf-ordered triangle, v0.w = triangle id bits
    const float4* rays;    // 2 float4 per ray
    uint32_t n;
    void* out;             // uint2 hits (closest) or uint32 flags (any)
    uint32_t* status;      // bit 0 set on traversal stack overflow
};

// loop: 0 = one step per iteration, 1 = while-while; 2 and 3 the same over
// the four-wide tree.
hipError_t launchTraceClosest(const QueryLaunch& q, uint32_t block, uint32_t stack, uint32_t loop,
                              hipStream_t s);
hipError_t launchTraceAny(const QueryLaunch& q, uint32_t block, uint32_t stack, uint32_t loop,
                          hipStream_t s);
bool traversalShapeSupported(uint32_t block, uint32_t stack, uint32_t loop);

} // namespace dev
} // namespace mrt
