// GPU binned-SAH BVH builder: same nodes and primitive order as buildBvh
// (bvh_build.h), built by the kernels of kernels/bvh_gpu.hip.
#pragma once

#include "bvh_build.h"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace mrt {

struct GpuBvhInfo {
    uint32_t nodeCount = 0, maxDepth = 0, leafCount = 0;
    uint32_t largeNodes = 0, subtrees = 0;  // nodes built level by level / subtrees built by one wavefront each
    uint32_t levels = 0;                    // levels of large nodes
    double msLevels = 0, msSubtrees = 0, msPlace = 0;  // host wall time of the three phases
};

// Builds on the current device.  Reusable: keeps its workspace between calls.
class GpuBvhBuilder {
public:
    GpuBvhBuilder();
    ~GpuBvhBuilder();
    GpuBvhBuilder(const GpuBvhBuilder&) = delete;
    GpuBvhBuilder& operator=(const GpuBvhBuilder&) = delete;

    // dBounds: 6 floats per primitive in device memory.  Writes info.nodeCount
    // 32-byte nodes to dNodes (throws when that exceeds nodeCap) and n
    // primitive indices to dOrder.  Returns false, writing nothing, when a node
    // needs the reference's median split (std::nth_element), which this
    // builder does not restate: the caller then builds on the CPU.  Waits for
    // the stream between levels.  n must stay below 2^31.
    bool build(const float* dBounds, uint32_t n, uint32_t maxLeaf, void* dNodes, uint64_t nodeCap, uint32_t* dOrder,
               GpuBvhInfo& info, hipStream_t s);

private:
    struct Work;
    Work* w_;
};

// Host bounds in, host result out, built on the GPU by `builder` (on the CPU
// when a node needs the median split; *onGpu tells which).
BvhBuild buildBvhOnDevice(GpuBvhBuilder& builder, const std::vector<Box3>& bounds, uint32_t maxLeaf, hipStream_t s,
                          bool* onGpu = nullptr);

} // namespace mrt
