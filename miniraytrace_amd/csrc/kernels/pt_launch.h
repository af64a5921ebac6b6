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
constexpr uint32_t kStatusWord = cntWord(kMaxBounceLevels, 0, 0);  // bit 0: stack overflow
constexpr uint32_t kCounterWords = kStatusWord + kLineWords;
constexpr uint32_t kCounterLines = kCounterWords / kLineWords;  // words in use: the first of each line
// Work cursors of a frame, one line each: stage 0 trace (closest hits, then
// the shadow rays sharing the launch), 1 shade, 2 shadow-only trace, 3 tail.
__host__ __device__ constexpr uint32_t cursorWord(uint32_t level, uint32_t stage, uint32_t seg) {
    return ((level * 4 + stage) * kQueueSegs + seg) * kLineWords;
}
constexpr uint32_t kCursorLines = kMaxBounceLevels * 4 * kQueueSegs;
constexpr uint32_t kCursorWords = kCursorLines * kLineWords;
constexpr uint32_t kNoLevel = 0xFFFFFFFFu;  // "no queue" for launchTraceBatch

struct TraceShape {
    uint32_t block;  // threads per workgroup
    uint32_t stack;  // LDS stack entries per lane
    uint32_t grid;   // persistent workgroups
};

hipError_t launchClearWords(uint32_t* words, uint32_t n, hipStream_t s);
hipError_t launchResolve32(const PathView& pv, uint32_t pixels, uint32_t spp, hipStream_t s);
// A frame over one or more requests (BatchView; an engine's own frame is a
// batch of one): clear, generate, then per bounce b one trace launch (closest
// hits of level b together with the shadow rays of level b - 1), for guided
// requests the network inference of their query records, shade and tail, a
// last trace launch for the deepest level's shadow rays, then accumulate +
// resolve.  tailPaths: the tail kernel serves the requests of a level once
// all requests together hold at most tailPaths rays there (0 = never), each
// further limited by BatchEntry::tailAt.
hipError_t launchClearBatch(const BatchView& bv, hipStream_t s);
hipError_t launchGenerateBatch(const BatchView& bv, hipStream_t s);
// extendLevel / shadowLevel: queue levels, either may be kNoLevel.
// guidePrep (some request guided): guided requests also get the network's
// query records of their closest hits at levels the tail does not take.
hipError_t launchTraceBatch(const BatchView& bv, uint32_t extendLevel, uint32_t shadowLevel, const TraceShape& shape,
                            uint32_t tailPaths, bool guidePrep, hipStream_t s);
// guided: some request shades with its learned mixtures.
hipError_t launchShadeBatch(const BatchView& bv, uint32_t bounce, uint32_t tailPaths, bool guided, uint32_t grid,
                            hipStream_t s);
hipError_t launchTailBatch(const BatchView& bv, uint32_t bounce, uint32_t tailPaths, const TraceShape& shape,
                           hipStream_t s);
hipError_t launchFinishBatch(const BatchView& bv, hipStream_t s);
hipError_t batchShadeBlocksPerCU(bool guided, int* blocksPerCU);
hipError_t batchTraceBlocksPerCU(uint32_t block, uint32_t stack, bool guidePrep, int* blocksPerCU);
hipError_t batchTailBlocksPerCU(uint32_t block, uint32_t stack, int* blocksPerCU);

bool ptShapeSupported(uint32_t block, uint32_t stack);

} // namespace dev
} // namespace mrt
