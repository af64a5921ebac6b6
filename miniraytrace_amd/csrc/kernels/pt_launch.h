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
constexpr uint32_t kNoLevel = 0xFFFFFFFFu;  // "no queue" for launchTraceBatch

struct TraceShape {
    uint32_t block;  // threads per workgroup
    uint32_t stack;  // LDS stack entries per lane
    uint32_t grid;   // persistent workgroups
};

hipError_t launchClearWords(uint32_t* words, uint32_t n, hipStream_t s);
hipError_t launchResolve32(const PathView& pv, uint32_t pixels, uint32_t spp, hipStream_t s);
// Guided shading of one engine's frame (needs the query records of this
// bounce's trace launch and the network's mixtures); skips a level that
// holds at most tailPaths rays (the tail kernel shades it, unguided).
hipError_t launchShadeGuided(const SceneView& sc, const PathView& pv, const FrameConsts* g, uint32_t bounce,
                             uint32_t tailPaths, uint32_t grid, hipStream_t s);
hipError_t guidedShadeBlocksPerCU(int* blocksPerCU);

// A frame over one or more requests (BatchView; an engine's own frame is a
// batch of one): clear, generate, then per bounce b one trace launch (closest
// hits of level b together with the shadow rays of level b - 1), shade and
// tail, a last trace launch for the deepest level's shadow rays, then
// accumulate + resolve.  Plain shading (guided engines shade with
// launchShadeGuided and use the same tail).
hipError_t launchClearBatch(const BatchView& bv, hipStream_t s);
hipError_t launchGenerateBatch(const BatchView& bv, hipStream_t s);
// extendLevel / shadowLevel: queue levels, either may be kNoLevel.
// guidePrep (a guided engine's batch of one): also write the guiding
// network's query records of the closest hits when the level holds more than
// prepAbove rays.
hipError_t launchTraceBatch(const BatchView& bv, uint32_t extendLevel, uint32_t shadowLevel, const TraceShape& shape,
                            hipStream_t s, bool guidePrep = false, uint32_t prepAbove = 0);
// Plain shading of level `bounce`; skipped when all requests together hold at
// most tailPaths rays there, in which case launchTailBatch finishes the
// frame for them (both are launched for every level; 0 = never the tail).
hipError_t launchShadeBatch(const BatchView& bv, uint32_t bounce, uint32_t tailPaths, uint32_t grid, hipStream_t s);
hipError_t launchTailBatch(const BatchView& bv, uint32_t bounce, uint32_t tailPaths, const TraceShape& shape,
                           hipStream_t s);
hipError_t batchTailBlocksPerCU(uint32_t block, uint32_t stack, int* blocksPerCU);
hipError_t launchFinishBatch(const BatchView& bv, hipStream_t s);
hipError_t batchShadeBlocksPerCU(int* blocksPerCU);
hipError_t batchTraceBlocksPerCU(uint32_t block, uint32_t stack, bool guidePrep, int* blocksPerCU);

bool ptShapeSupported(uint32_t block, uint32_t stack);

} // namespace dev
} // namespace mrt
