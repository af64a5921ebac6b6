// One Adam step of the path-guiding network with its training workspace in a
// fixed device array instead of allocated memory, for callers that cannot
// allocate device memory (work under HIP graph capture, tight serving loops).
//
// The array `scratch` holds, for batches of up to kNoAllocRows records:
// the encoded inputs, the hidden activations, the raw outputs, the five
// per-layer delta blocks, the per-chunk weight-gradient partials, the
// per-chunk batch-weight sums and the batch-size word that the training
// kernels read.  It is a file-scope device variable, so every GPU the
// library runs on holds its own copy (about 13 MB) until the process ends.
// Host code never passes the array's name to a kernel: on the first call on
// a device, with that device current, hipGetSymbolAddress yields that
// device's address, which is kept per device and used for every kernel
// argument.
//
// Every call on one device shares that device's array and it keeps the last
// call's contents; calls that overlap on one device (on different streams,
// or from different threads) race on it.  Each call clears what must start
// at zero (the unused input-feature rows 80..95) with a fill kernel, so the
// call can be captured into a graph (no memset node).
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

// Byte offsets of the workspace blocks inside the array (feature-major, row
// stride kStride, as GuideNet describes them).
constexpr size_t kOffActIn = 0;                                                   // bf16 [96][stride]
constexpr size_t kOffAct = kOffActIn + align256(size_t(96) * kStride * 2);        // bf16 [4][64][stride]
constexpr size_t kOffRaw = kOffAct + align256(size_t(4) * 64 * kStride * 2);      // f32 [40][stride]
constexpr size_t kDeltaBytes = align256(size_t(64) * kStride * 2);                // bf16 [64][stride]
constexpr size_t kOffDelta = kOffRaw + align256(size_t(kGuideOutputs) * kStride * 4);  // x5
constexpr size_t kOffPartial = kOffDelta + 5 * kDeltaBytes;                       // f32 [chunks][params]
constexpr size_t kOffWeightPart = kOffPartial + align256(size_t(kChunks) * kGuideParamCount * 4);
constexpr size_t kOffCount = kOffWeightPart + align256(size_t(kChunks) * 4);      // u32 batch size
constexpr size_t kScratchBytes = kOffCount + 256;

// Input-feature rows that no kernel writes but the layer-0 weight-gradient
// tiles read (as zeros).
constexpr uint32_t kZeroRowFirst = 80, kZeroRows = 16;
constexpr uint32_t kZeroWords = kZeroRows * kStride / 2;  // u32 words (two bf16 each)

constexpr int kMaxDevices = 64;
std::mutex gScratchMutex;
void* gScratchPtr[kMaxDevices] = {};

} // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

// Clears the unused input-feature rows and publishes the batch size.
__global__ __launch_bounds__(256) void kNoAllocBegin(uint32_t* __restrict__ zeroRows, uint32_t* __restrict__ countWord,
                                                     uint32_t count) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < kZeroWords) zeroRows[i] = 0u;
    if (i == 0) *countWord = count;
}

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

    GuideNet net = state;  // parameters, Adam moments, gradient, image and state words of the network
    net.actIn = reinterpret_cast<uint16_t*>(base + kOffActIn);
    net.act = reinterpret_cast<uint16_t*>(base + kOffAct);
    net.raw = reinterpret_cast<float*>(base + kOffRaw);
    for (int l = 0; l < 5; ++l) net.delta[l] = reinterpret_cast<uint16_t*>(base + kOffDelta + l * kDeltaBytes);
    net.partial = reinterpret_cast<float*>(base + kOffPartial);
    net.weightPart = reinterpret_cast<float*>(base + kOffWeightPart);
    net.cap = kNoAllocRows;
    net.stride = kStride;
    uint32_t* countWord = reinterpret_cast<uint32_t*>(base + kOffCount);

    uint32_t* zeroRows = reinterpret_cast<uint32_t*>(net.actIn + size_t(kZeroRowFirst) * kStride);
    hipLaunchKernelGGL(kNoAllocBegin, dim3((kZeroWords + 255) / 256), dim3(256), 0, s, zeroRows, countWord, count);
    if ((e = hipGetLastError()) != hipSuccess) return e;

    GuideBatch b;
    b.pool = records;
    b.total = countWord;
    b.poolCap = count;
    b.phase = 0;
    b.period = 1;
    b.cap = count;
    if ((e = launchGuideGradient(b, net, true, s)) != hipSuccess) return e;
    return launchGuideAdam(net, h, s);
}

} // namespace dev
} // namespace mrt
