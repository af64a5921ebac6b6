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
// Transposed image that follows it: bf16 W_l^T of layers 1..4 (64 inputs x
// 72-element rows of outputs), the backward pass's operand, in bytes.
constexpr size_t kGuideImageTBytes = 4 * 64 * 72 * 2;
constexpr uint32_t kGuideGradChunk = 2048;  // training rows per weight-gradient partial (8 waves x 256 rows)

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
// up to `cap` rows (`stride` = cap rounded up to 32).  Inputs, activations
// and deltas are stored row-block-tiled: element (feature f, row r) of an
// array of F features at ((r / 16) F + f) 16 + r % 16, so the weight-gradient
// kernel reads each operand fragment of a wave as one contiguous kilobyte.
struct GuideNet {
    float* params = nullptr;     // kGuideParamCount
    float* adamM = nullptr;
    float* adamV = nullptr;
    float* grad = nullptr;       // written by the gradient reduction
    void* image = nullptr;       // packed forward image (kGuideImageBytes), kept in step with params
    void* imageT = nullptr;      // transposed image (kGuideImageTBytes), kept in step with params
    uint16_t* actIn = nullptr;   // bf16, 96 features x stride rows: encoded inputs (features 78..95 zero)
    uint16_t* act = nullptr;     // bf16, 4 x (64 features x stride rows): hidden-layer outputs
    uint16_t* delta[5] = {};     // bf16, 64 features x stride rows per layer: error at layer l's outputs (unnormalised)
    float* tilePart = nullptr;   // FP32 [stride / 32][2]: weight and loss sums per 32-row tile
    float* partial = nullptr;    // FP32 [guideGradChunks(cap)][kGuideParamCount]
    float* weightPart = nullptr; // FP32 [guideGradChunks(cap)][2]: weight and loss sums per row chunk
    uint32_t* words = nullptr;   // kGwWords
    uint32_t cap = 0;
    uint32_t stride = 0;
};

struct GuideAdam {
    float lr = 1e-3f, beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
};

// Packs params into the forward image (bf16 weights, padded, and biases)
// and the transposed image.
hipError_t launchGuidePack(const float* params, void* image, void* imageT, hipStream_t s);

// Network outputs for the first *countPtr rows of rec (countPtr in device
// memory, at most maxCount).  decoded: kGuideMixFloat4 float4 per row
// (the renderer's format); otherwise kGuideOutputs raw floats per row.
hipError_t launchGuideInfer(const RecView& rec, const uint32_t* countPtr, uint32_t maxCount, const void* image,
                            void* out, bool decoded, hipStream_t s);

// grad = d(loss)/d(params) of one batch, loss = sum w (-log q) / sum w.
// countStep: advance the Adam step counter when the batch is not empty.
hipError_t launchGuideGradient(const GuideBatch& b, const GuideNet& net, bool countStep, hipStream_t s);

// One training step: the gradient as launchGuideGradient (advancing the
// step counter), then the Adam update (skipped for an empty batch) and the
// images' bf16 weights.  Three launches.
hipError_t launchGuideStep(const GuideBatch& b, const GuideNet& net, const GuideAdam& h, hipStream_t s);

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
    const uint32_t* recMask = nullptr; // per training path: levels that wrote a record this frame
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
