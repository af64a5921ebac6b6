// Host-callable launchers of the wavefront path-tracing kernels.
#pragma once

#include "pt_params.h"

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mrt {
namespace dev {

// Bounce levels 0..16 (maxBounces is clamped to 16) plus one for the queue
// a level-16 shade would fill.
constexpr uint32_t kMaxBounceLevels = 18;
constexpr uint32_t kStatusWord = kMaxBounceLevels * kCntPerBounce;  // bit 0: stack overflow
constexpr uint32_t kCounterWords = kStatusWord + 8;

struct TraceShape {
    uint32_t block;  // threads per workgroup
    uint32_t stack;  // LDS stack entries per lane
    uint32_t grid;   // persistent workgroups
};

hipError_t launchClearWords(uint32_t* words, uint32_t n, hipStream_t s);
hipError_t launchGenerate(const PathView& pv, const FrameConsts* g, uint32_t pixels, hipStream_t s);
hipError_t launchExtend(const SceneView& sc, const PathView& pv, const FrameConsts* g, uint32_t bounce,
                        const TraceShape& shape, hipStream_t s);
// guide: the guiding variant (needs the prep pass of this bounce).
hipError_t launchShade(const SceneView& sc, const PathView& pv, const FrameConsts* g, uint32_t bounce,
                       uint32_t grid, bool guide, hipStream_t s);
// Query records of guidable hits for the guiding network (grid-stride).
hipError_t launchGuidePrep(const SceneView& sc, const PathView& pv, const FrameConsts* g, uint32_t bounce,
                           uint32_t grid, hipStream_t s);
hipError_t launchShadow(const SceneView& sc, const PathView& pv, const FrameConsts* g, uint32_t bounce,
                        const TraceShape& shape, hipStream_t s);
hipError_t launchAccumulate(const PathView& pv, const FrameConsts* g, uint32_t pixels, hipStream_t s);
hipError_t launchResolve8(const PathView& pv, const FrameConsts* g, uint32_t pixels, hipStream_t s);
hipError_t launchResolve32(const PathView& pv, uint32_t pixels, uint32_t spp, hipStream_t s);

bool ptShapeSupported(uint32_t block, uint32_t stack);
hipError_t persistentGrid(uint32_t block, uint32_t stack, bool shadow, int* blocksPerCU);
hipError_t shadeBlocksPerCU(bool guide, int* blocksPerCU);

} // namespace dev
} // namespace mrt
