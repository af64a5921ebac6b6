// Host-visible launchers of the path-guiding network kernels (guide.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace mrt {
namespace dev {

constexpr int kGuideParamCount = 20136;  // flat FP32 parameter vector
constexpr int kGuideOutputs = 40;        // raw outputs per query
constexpr int kGuideInputsPadded = 80;   // bf16 features per row (78 used)

// Buffers of one training step over `count` records.
//   rec       recStride float4 per record: (pos01, weight), (n, 0), (wo, 0), (wi, 0)
//   acts      bf16: count x 80 (encoded inputs) then 4 x count x 64 (hidden activations)
//   raw       FP32 count x 40 raw outputs
//   delta[2]  bf16 count x 64 each (ping-pong of the back-propagated error)
//   grad      FP32 gradient accumulator (zero on entry; Adam clears it)
//   scratch   FP32 [0] = sum of weights, [1] = loss (both cleared by the caller)
struct GuideTrainArgs {
    const float4* rec = nullptr;
    uint32_t recStride = 4;
    uint32_t count = 0;
    uint32_t grid = 1024;
    float* params = nullptr;
    float* adamM = nullptr;
    float* adamV = nullptr;
    float* grad = nullptr;
    uint16_t* acts = nullptr;
    float* raw = nullptr;
    uint16_t* delta[2] = { nullptr, nullptr };
    float* scratch = nullptr;
    float lr = 1e-3f, beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    float biasCorr1 = 0.1f, biasCorr2 = 0.001f;  // 1 - beta^t
};

// Network outputs for the first *countPtr records (countPtr is device memory,
// at most maxCount).  rec rows: (pos01, .), (n, .), (wo, .).
hipError_t launchGuideInfer(const float4* rec, uint32_t recStride, const uint32_t* countPtr, uint32_t maxCount,
                            const float* params, float* out, uint32_t grid, hipStream_t s);

// Forward + backward + Adam over one batch.
hipError_t launchGuideTrainStep(const GuideTrainArgs& a, hipStream_t s);

// Training-style forward only (activations and raw outputs), for tests.
hipError_t launchGuideForwardBatch(const float4* rec, uint32_t recStride, uint32_t count, const float* params,
                                   uint16_t* acts, float* raw, uint32_t grid, hipStream_t s);

} // namespace dev
} // namespace mrt
