// Host-visible launchers of the path-guiding network kernels (guide.hip).
//
// Record formats (float4 rows):
//   query record     (pos01.xyz, -), (n.xyz, -), (wo.xyz, -)            stride >= 3
//   training record  (pos01.xyz, weight), (n.xyz, -), (wo.xyz, -), (wi.xyz, -)
// pos01 is the position normalised to the scene box, n the shading normal
// facing wo, wo the direction towards the previous vertex and wi the sampled
// incident direction, all in world space.
#pragma once

#include "common/gpu_types.h"

#include <hip/hip_runtime.h>
#include <cstdint>

namespace mrt {
namespace dev {

constexpr int kGuideParamCount = 20136;  // flat FP32 parameter vector
constexpr int kGuideOutputs = 40;        // raw outputs per query
constexpr int kGuideInputsPadded = 80;   // bf16 features per row (78 used)
constexpr int kGuideMixFloat4 = 12;      // decoded mixture per query (see guide_device.h)
// Packed forward image: bf16 weights (24,064 with row padding) then FP32 biases (5 x 64), in bytes.
constexpr size_t kGuideImageBytes = 24064 * 2 + 320 * 4;
constexpr uint32_t kGuideGradChunk = 1024;  // training rows per weight-gradient partial (4 waves x 256 rows)

__host__ __device__ constexpr uint32_t guideRowStride(uint32_t cap) { return (cap + 31u) & ~31u; }
__host__ __device__ constexpr uint32_t guideGradChunks(uint32_t cap) {
    return (guideRowStride(cap) + kGuideGradChunk - 1) / kGuideGradChunk;
}

// Device state words of one network (uint32 / float bit patterns).
enum GuideWord : uint32_t {
    kGwWeightSum = 0,  // f32: sum of the current batch's weights
    kGwLoss = 1,       // f32: normalised loss of the current batch
    kGwBatch = 2,      // rows in the current batch
    kGwStep = 3,       // Adam steps taken (t)
    kGwLastLoss = 4,   // f32: loss of the last Adam step
    kGwLastBatch = 5,  // rows of the last Adam step
    kGwCand0 = 6,      // positive-weight records found, per stamp parity (2 words)
    kGwPool = 8,       // records appended to the training pool this frame
    kGwWords = 16,
};

// Rows i = 0..count-1 of a record array: base + (phase + i * period) * stride.
struct RecView {
    const float4* base = nullptr;
    uint32_t stride = 4;
    uint32_t phase = 0;
    uint32_t period = 1;
};

// One training batch drawn from a pool of training records whose size lives
// in device memory: rows phase, phase + period, ... of the first
// min(*total, poolCap) records, at most cap of them.
struct GuideBatch {
    const float4* pool = nullptr;
    const uint32_t* total = nullptr;
    uint32_t poolCap = 0;
    uint32_t phase = 0;
    uint32_t period = 1;
    uint32_t cap = 0;
};

// Network parameters, Adam moments and the training workspace for batches of
// up to `cap` rows.  Activations and deltas are stored feature-major
// ([feature][row], row stride `stride` = cap rounded up to 32) so the matrix
// kernels read 16-byte fragments.
struct GuideNet {
    float* params = nullptr;     // kGuideParamCount
    float* adamM = nullptr;
    float* adamV = nullptr;
    float* grad = nullptr;       // written by the gradient reduction
    void* image = nullptr;       // packed forward image (kGuideImageBytes), kept in step with params
    uint16_t* actIn = nullptr;   // bf16 [96][stride] encoded inputs (rows 78..95 zero)
    uint16_t* act = nullptr;     // bf16 [4][64][stride] hidden-layer outputs
    float* raw = nullptr;        // FP32 [40][stride] raw outputs
    uint16_t* delta[5] = {};     // bf16 [64][stride] per layer: error at layer l's outputs
    float* partial = nullptr;    // FP32 [guideGradChunks(cap)][kGuideParamCount]
    float* weightPart = nullptr; // FP32 [guideGradChunks(cap)]: batch weight sum per row chunk
    uint32_t* words = nullptr;   // kGwWords
    uint32_t cap = 0;
    uint32_t stride = 0;
};

struct GuideAdam {
    float lr = 1e-3f, beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
};

// Packs params into the forward image (bf16 weights, padded, and biases).
hipError_t launchGuidePack(const float* params, void* image, hipStream_t s);

// Network outputs for the first *countPtr rows of rec (countPtr in device
// memory, at most maxCount).  decoded: kGuideMixFloat4 float4 per row
// (the renderer's format); otherwise kGuideOutputs raw floats per row.
hipError_t launchGuideInfer(const RecView& rec, const uint32_t* countPtr, uint32_t maxCount, const void* image,
                            void* out, bool decoded, hipStream_t s);

// grad = d(loss)/d(params) of one batch, loss = sum w (-log q) / sum w.
// countStep: advance the Adam step counter when the batch is not empty.
hipError_t launchGuideGradient(const GuideBatch& b, const GuideNet& net, bool countStep, hipStream_t s);

// Adam update from grad (skipped when the last batch was empty), then repacks
// the forward image.
hipError_t launchGuideAdam(const GuideNet& net, const GuideAdam& h, hipStream_t s);

// One training step (gradient + Adam) on `count` device-resident training
// records (4 float4 each) with the workspace in the library's fixed device
// array (guide_noalloc.hip) instead of state's own: only the parameters,
// Adam moments, gradient, image and state words of `state` are used.
// count must be 1..guideNoAllocMaxRows(); `device` is the current device.
// Allocates nothing.  Calls that overlap on one device race on the array.
hipError_t launchGuideTrainNoAlloc(const GuideNet& state, const float4* records, uint32_t count, const GuideAdam& h,
                                   int device, hipStream_t s);
uint32_t guideNoAllocMaxRows();

// Decodes raw outputs and samples the mixture with the given uniforms:
// out row i = (direction.xyz, mixture pdf of that direction).
hipError_t launchGuideSample(const float* raw, const float* u, uint32_t count, float4* out, hipStream_t s);

// Mixture pdf of given directions: raw outputs + dirs (float4 xyz) -> pdf.
hipError_t launchGuidePdf(const float* raw, const float4* dirs, uint32_t count, float* out, hipStream_t s);

// Per-frame training-pool bookkeeping (clears this frame's counters).
hipError_t launchGuideFrameBegin(uint32_t* words, const FrameConsts* g, hipStream_t s);

// Path records written by the shade kernel for the frame's training paths
// (paths p with p % guideTrainStride == guideTrainPhase), per bounce level
// and training path q = p / guideTrainStride:
//   (pos01.xyz, luminance of the throughput before the vertex),
//   (n.xyz, luminance of the path radiance when the vertex was shaded),
//   (wo.xyz, FrameConsts::guideStamp as bits), (wi.xyz, -)
struct GuideCollectArgs {
    const float4* pathRec = nullptr;  // [level][trainPaths] x 4 float4
    const float4* radiance = nullptr; // per path: rgb radiance of this frame
    uint32_t paths = 0;               // all paths of the frame
    uint32_t trainPaths = 0;          // ceil(paths / guideTrainStride)
    uint32_t levels = 0;
    float4* pool = nullptr;           // training records
    uint32_t poolCap = 0;
    uint32_t* words = nullptr;
    const FrameConsts* g = nullptr;
};
// Turns path records into weighted training records (weight = luminance of
// the radiance gathered after the vertex / throughput luminance before it),
// keeping a random subset sized to the pool.
hipError_t launchGuideCollect(const GuideCollectArgs& a, hipStream_t s);

} // namespace dev
} // namespace mrt
