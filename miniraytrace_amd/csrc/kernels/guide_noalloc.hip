// One Adam step of the path-guiding network with its training workspace in a
// fixed device array instead of allocated memory, for callers that cannot
// allocate device memory (work under HIP graph capture, tight serving loops).
//
// The array `scratch` holds, for batches of up to kNoAllocRows records:
// the encoded inputs, the hidden activations, the five per-layer delta
// blocks, the per-tile and per-chunk weight and loss sums, the per-chunk
// weight-gradient partials and the batch-size word that the training
// kernels read.  It is a file-scope device variable, so every GPU the
// library runs on holds its own copy (about 12 MB) until the process ends.
// Host code never passes the array's name to a kernel: on the first call on
// a device, with that device current, hipGetSymbolAddress yields that
// device's address, which is kept per device and used for every kernel
// argument.
//
// Every call on one device shares that device's array and it keeps the last
// call's contents; calls that overlap on one device (on different streams,
// or from different threads) race on it.  Nothing in it must start at zero:
// the training kernels write every element they read (the zero input
// features 80..95 included), and the batch size is published by a small
// kernel, so the call can be captured into a graph (no memset node).
#include "guide_launch.h"
#include "mini_raytrace_ext.h"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

namespace {

using namespace mrt::dev;

constexpr uint32_t kNoAllocRows = MRTX_GUIDE_NOALLOC_MAX_RECORDS;  // largest batch
constexpr uint32_t kStride = guideRowStride(kNoAllocRows);
constexpr uint32_t kChunks = guideGradChunks(kNoAllocRows);

constexpr size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Byte offsets of the workspace blocks inside the array (row-block-tiled,
// kStride rows, as GuideNet describes them).
constexpr size_t kOffActIn = 0;                                                   // bf16 [96][stride]
constexpr size_t kOffAct = kOffActIn + align256(size_t(96) * kStride * 2);        // bf16 [4][64][stride]
constexpr size_t kOffTile = kOffAct + align256(size_t(4) * 64 * kStride * 2);     // f32 [stride / 32][2]
constexpr size_t kDeltaBytes = align256(size_t(64) * kStride * 2);                // bf16 [64][stride]
constexpr size_t kOffDelta = kOffTile + align256(size_t(kStride / 32) * 2 * 4);   // x5
constexpr size_t kOffPartial = kOffDelta + 5 * kDeltaBytes;                       // f32 [chunks][params]
constexpr size_t kOffWeightPart = kOffPartial + align256(size_t(kChunks) * kGuideParamCount * 4);  // f32 [chunks][2]
constexpr size_t kOffCount = kOffWeightPart + align256(size_t(kChunks) * 2 * 4);  // u32 batch size
constexpr size_t kScratchBytes = kOffCount + 256;

constexpr int kMaxDevices = 64;
std::mutex gScratchMutex;
void* gScratchPtr[kMaxDevices] = {};

} // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

// Publishes the batch size (the training kernels read it from device memory).
__global__ void kNoAllocBegin(uint32_t* __restrict__ countWord, uint32_t count) { *countWord = count; }

// This device's address of `scratch` (the device must be current).
hipError_t scratchAddress(int device, unsigned char** out) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(gScratchMutex);
    if (!gScratchPtr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        gScratchPtr[device] = p;
    }
    *out = static_cast<unsigned char*>(gScratchPtr[device]);
    return hipSuccess;
}

} // namespace

namespace mrt {
namespace dev {

uint32_t guideNoAllocMaxRows() { return kNoAllocRows; }

hipError_t launchGuideTrainNoAlloc(const GuideNet& state, const float4* records, uint32_t count, const GuideAdam& h,
                                   int device, hipStream_t s) {
    if (count == 0 || count > kNoAllocRows || !records) return hipErrorInvalidValue;
    unsigned char* base = nullptr;
    hipError_t e = scratchAddress(device, &base);
    if (e != hipSuccess) return e;

    GuideNet net = state;  // parameters, Adam moments, gradient, images and state words of the network
    net.actIn = reinterpret_cast<uint16_t*>(base + kOffActIn);
    net.act = reinterpret_cast<uint16_t*>(base + kOffAct);
    net.tilePart = reinterpret_cast<float*>(base + kOffTile);
    for (int l = 0; l < 5; ++l) net.delta[l] = reinterpret_cast<uint16_t*>(base + kOffDelta + l * kDeltaBytes);
    net.partial = reinterpret_cast<float*>(base + kOffPartial);
    net.weightPart = reinterpret_cast<float*>(base + kOffWeightPart);
    net.cap = kNoAllocRows;
    net.stride = kStride;
    uint32_t* countWord = reinterpret_cast<uint32_t*>(base + kOffCount);

    hipLaunchKernelGGL(kNoAllocBegin, dim3(1), dim3(1), 0, s, countWord, count);
    if ((e = hipGetLastError()) != hipSuccess) return e;

    GuideBatch b;
    b.pool = records;
    b.total = countWord;
    b.poolCap = count;
    b.phase = 0;
    b.period = 1;
    b.cap = count;
    return launchGuideStep(b, net, h, s);
}

} // namespace dev
} // namespace mrt
