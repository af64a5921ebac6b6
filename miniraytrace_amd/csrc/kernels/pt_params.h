// Parameter blocks of the wavefront path-tracing kernels (plain device
// pointers; passed by value to every launch).
#pragma once

#include "common/gpu_types.h"

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mrt {
namespace dev {

struct Node;

struct SceneView {
    const Node* nodes;               // all BLAS nodes, then the TLAS at FrameConsts::tlasRoot
    const float4* wide;              // their four-wide copies (8 float4 per node), TLAS at tlasWideRoot
    const float4* triVerts;          // 3 float4 per leaf-ordered triangle (object space)
    const GpuTriangle* tris;         // leaf-ordered vertex indices (global)
    const float4* positions;
    const float4* normals;
    const float2* uvs;
    const float4* tangents;
    const GpuInstance* instances;    // TLAS leaf order
    const TraceInstance* traceInst;  // same order, compact traversal copy
    const GpuMaterial* materials;
    const GpuLight* lights;
    const float4* envMap;            // envWidth x envHeight RGBA32F, row 0 = top (theta = 0)
    const AliasEntry* envAlias;      // envWidth * envHeight entries
    const EnvStratumEntry* envA;     // FrameConsts::envCountA entries
    const EnvStratumEntry* envB;     // FrameConsts::envCountB entries
    const GpuTexture* texTable;      // 64 slots
    const uint32_t* tex8;            // RGBA8 texel pool
    const float4* tex32;             // RGBA32F texel pool
};

// Queue segments.  Every ray and shadow-ray queue of a request is split into
// kQueueSegs segments of PathView::segCap slots (slot = segment x segCap +
// item), each with its own counters and work cursors.  Device-scope atomics
// on one address serialise (about 11 ns each on the MI355X, whatever the
// number of waves), so a queue with one counter and one work cursor is
// capped near 90M appends or chunk fetches per second; eight segments, each
// word on a 128-byte line of its own, spread them over eight addresses.  A
// wave starts on the segment of its XCD and moves on to the others when that
// one is spent.  Level-0 segments are contiguous ranges of camera-ray slots
// (image bands), and a chunk's new rays and shadow rays go to its own
// segment, so a segment never holds more than its level-0 rays.
constexpr uint32_t kQueueSegs = 8;
constexpr uint32_t kLineWords = 32;  // one 128-byte line

// Per-bounce counters: word cntWord(level, slot, segment).
enum CounterSlot : uint32_t {
    kCntRays = 0,          // rays queued for this bounce, per segment
    kCntShadows = 1,       // shadow rays queued at this bounce, per segment
    kCntGuideRecords = 2,  // guiding query records of this bounce (segment 0)
    kCntTailRays = 3,      // rays traced at this bounce by the tail kernel (statistics, per segment)
    kCntTailShadows = 4,   // shadow rays traced at this bounce by the tail kernel (statistics, per segment)
    kCntSlots = 5,
};
__host__ __device__ constexpr uint32_t cntWord(uint32_t level, uint32_t slot, uint32_t seg) {
    return ((level * kCntSlots + slot) * kQueueSegs + seg) * kLineWords;
}
// Slots per queue segment for `paths` paths: a multiple of 64, so a wave's
// 64 slots never straddle two segments.
__host__ __device__ constexpr uint32_t queueSegCap(uint32_t paths) {
    return ((paths + kQueueSegs - 1) / kQueueSegs + 63u) & ~63u;
}

struct PathView {
    uint32_t capacity;               // paths (= pixels) per frame
    uint32_t segCap;                 // queueSegCap(capacity): queues hold kQueueSegs x segCap slots
    float4* beta;                    // rgb throughput + previous BSDF pdf
    float4* radiance;                // rgb radiance + BSDF pdf of the last continuation (guided shading)
    uint32_t* rngCounter;            // draws consumed so far
    float4* rayO[2];                 // queued ray origin xyz + path index bits
    float4* rayD[2];                 // queued ray direction xyz
    float4* hit;                     // t, b1, b2, triangle index bits (per queue slot)
    uint32_t* hitInst;               // instance index (per queue slot)
    float4* shadowO;                 // origin xyz + tmax
    float4* shadowD;                 // direction xyz + path index bits
    float4* shadowC;                 // unoccluded contribution rgb
    uint32_t* counters;              // kCounterWords: cntWord(level, slot, segment), then the status word
    float4* accum;                   // rgb sum + sample count, per pixel
    uint32_t* outRgba8;              // resolved image
    float4* outRgba32;               // accum / spp, alpha 1
    // Path guiding (null when off).
    uint32_t* guideIdx;              // per queue slot: row of guideRec / guideMix, or ~0u
    float4* guideRec;                // 3 float4 per guided row: (pos01), (n), (wo)
    float4* guideMix;                // 12 float4 per guided row: decoded mixture
    float4* pathRec;                 // [level][path] x 4 float4 training records
    uint32_t* recMask;               // per training path: bit b set when level b holds a record of this frame
    // Learned environment-stratum selection (null when guided sampling is off).
    uint2* shadowX;                  // per shadow slot: (cell << 1 | stratum, contribution bits) or ~0u
    float4* neeStats;                // kNeeCells cells
    // Guided sampling (null when off): per path, the product over its
    // vertices of (pdf the continuation was drawn with) / (its BSDF pdf), so
    // beta * rrScale is the throughput BSDF sampling alone would carry.
    // Russian roulette reads that throughput.
    float* rrScale;
    // Host-mapped pinned words (null when not recorded): kFinishBatch writes
    // the rays queued at each level of the frame (kMaxBounceLevels words).
    uint32_t* levelRays;
};

// One request of a batched frame (requests keep their own queues, counters
// and frame buffers; the batch kernels only spread their waves over them).
struct BatchEntry {
    SceneView sc;
    PathView pv;
    const FrameConsts* g;            // this request's constants for this frame
    uint32_t pixels;                 // width * height
    uint32_t pixelBase;              // first slot in the batch pixel space (multiple of 64)
    uint32_t active;                 // 1 when the request renders this frame
    uint32_t tailAt;                 // the tail kernel serves this request at a level holding at
                                     // most this many of its rays (~0u: no limit of its own)
};

constexpr uint32_t kMaxBatchRequests = 64;  // one wave lane per request in the work split

struct BatchView {
    const BatchEntry* entries;       // count entries, pixelBase ascending
    uint32_t count;                  // <= kMaxBatchRequests
    uint32_t pixelSlots;             // padded pixel space of all requests
    uint32_t* cursors;               // kCursorWords: cursorWord(level, stage, segment), cleared per frame
};

} // namespace dev
} // namespace mrt
