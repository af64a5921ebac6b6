// C entry points of the path-guiding network (mini_raytrace_ext.h, section 5):
// the engine's guiding controls and the standalone network object.
#include "capi_handles.h"
#include "common.h"
#include "engine.h"
#include "guide.h"
#include "mini_raytrace.h"
#include "mini_raytrace_ext.h"

#include <algorithm>
#include <cstddef>
#include <cstring>
#include <memory>
#include <string>

struct mrtxGuideNetwork {
    int device = 0;
    hipStream_t stream = nullptr;
    std::unique_ptr<mrt::GuideNetwork> net;
    mrt::DeviceBuffer a, b, out, count;  // staging for host calls
};

namespace {

constexpr size_t kRecFloats = 16;  // floats per record in host buffers

using mrt::unwrap;
using mrt::validStruct;

// Uploads count records and the count word; returns the record view.
mrt::dev::RecView stageRecords(mrtxGuideNetwork* n, const float* records, uint32_t count) {
    n->a.ensure(size_t(count) * kRecFloats * sizeof(float));
    n->a.upload(records, size_t(count) * kRecFloats * sizeof(float), 0, n->stream);
    n->count.ensure(sizeof(uint32_t));
    n->count.upload(&count, sizeof(uint32_t), 0, n->stream);
    mrt::dev::RecView v;
    v.base = n->a.as<const float4>();
    v.stride = 4;
    return v;
}

mrt::dev::GuideBatch wholeBatch(mrtxGuideNetwork* n, uint32_t count) {
    mrt::dev::GuideBatch b;
    b.pool = n->a.as<const float4>();
    b.total = n->count.as<const uint32_t>();
    b.poolCap = count;
    b.phase = 0;
    b.period = 1;
    b.cap = count;
    return b;
}

float wordAsFloat(const uint32_t* w, uint32_t i) {
    float f;
    std::memcpy(&f, w + i, sizeof(float));
    return f;
}

} // namespace

extern "C" {

// ---- engine guiding ----------------------------------------------------------

// Callers built before the warm-up fields pass the shorter struct: the
// fields it lacks take their defaults (set) or are not written (get).
constexpr size_t kGuidingOptionsV1 = offsetof(mrtxGuidingOptions, warmupFrames);

MRT_API mrtResult mrtxSetGuiding(mrtEngine* e, const mrtxGuidingOptions* o) {
    mrt::Engine* eng = unwrap(e);
    if (!eng) return MRT_ERROR_INVALID_HANDLE;
    if (!o || o->structSize < kGuidingOptionsV1) return MRT_ERROR_INVALID_ARGUMENT;
    return mrt::guarded([&] {
        mrtxGuidingOptions full{};
        std::memcpy(&full, o, std::min(o->structSize, sizeof(full)));
        full.structSize = sizeof(full);
        eng->setGuiding(full);
        return MRT_SUCCESS;
    });
}

MRT_API mrtResult mrtxGetGuiding(mrtEngine* e, mrtxGuidingOptions* o) {
    mrt::Engine* eng = unwrap(e);
    if (!eng) return MRT_ERROR_INVALID_HANDLE;
    if (!o || o->structSize < kGuidingOptionsV1) return MRT_ERROR_INVALID_ARGUMENT;
    mrtxGuidingOptions full;
    eng->getGuiding(full);
    const size_t n = std::min(o->structSize, sizeof(full));
    full.structSize = o->structSize;
    std::memcpy(o, &full, n);
    return MRT_SUCCESS;
}

MRT_API mrtResult mrtxGetGuidingStats(mrtEngine* e, mrtxGuidingStats* o) {
    mrt::Engine* eng = unwrap(e);
    if (!eng) return MRT_ERROR_INVALID_HANDLE;
    if (!o || o->structSize < offsetof(mrtxGuidingStats, envBrightShare)) return MRT_ERROR_INVALID_ARGUMENT;
    return mrt::guarded([&] {
        mrtxGuidingStats full;
        eng->guidingStats(full);
        const size_t n = std::min(o->structSize, sizeof(full));
        full.structSize = o->structSize;
        std::memcpy(o, &full, n);
        return MRT_SUCCESS;
    });
}

MRT_API mrtResult mrtxGetGuideParams(mrtEngine* e, float* dst, size_t count) {
    mrt::Engine* eng = unwrap(e);
    if (!eng) return MRT_ERROR_INVALID_HANDLE;
    if (!dst || count != MRTX_GUIDE_PARAM_COUNT) return MRT_ERROR_INVALID_ARGUMENT;
    return mrt::guarded([&] {
        eng->getGuideParams(dst);
        return MRT_SUCCESS;
    });
}

MRT_API mrtResult mrtxSetGuideParams(mrtEngine* e, const float* src, size_t count) {
    mrt::Engine* eng = unwrap(e);
    if (!eng) return MRT_ERROR_INVALID_HANDLE;
    if (!src || count != MRTX_GUIDE_PARAM_COUNT) return MRT_ERROR_INVALID_ARGUMENT;
    return mrt::guarded([&] {
        eng->setGuideParams(src);
        return MRT_SUCCESS;
    });
}

// ---- standalone network --------------------------------------------------------

MRT_API mrtResult mrtxCreateGuideNetwork(int32_t gpuIndex1, uint32_t maxBatch, uint32_t seed,
                                         mrtxGuideNetwork** out) {
    if (!out || gpuIndex1 < 0 || maxBatch == 0 || maxBatch > (1u << 22)) return MRT_ERROR_INVALID_ARGUMENT;
    return mrt::guarded([&] {
        auto n = std::make_unique<mrtxGuideNetwork>();
        n->device = mrt::selectDevice(uint32_t(gpuIndex1));
        mrt::DeviceScope scope(n->device);
        MRT_HIP(hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking));
        n->net = std::make_unique<mrt::GuideNetwork>(maxBatch);
        n->net->initParams(seed);
        *out = n.release();
        return MRT_SUCCESS;
    });
}

MRT_API void mrtxDestroyGuideNetwork(mrtxGuideNetwork* n) {
    if (!n) return;
    (void)mrt::guarded([&] {
        mrt::DeviceScope scope(n->device);
        if (n->stream) (void)hipStreamSynchronize(n->stream);
        n->net.reset();
        n->a.release();
        n->b.release();
        n->out.release();
        n->count.release();
        if (n->stream) (void)hipStreamDestroy(n->stream);
        delete n;
        return MRT_SUCCESS;
    });
}

MRT_API mrtResult mrtxGuideNetworkSetParams(mrtxGuideNetwork* n, const float* src, size_t count) {
    if (!n) return MRT_ERROR_INVALID_HANDLE;
    if (!src || count != MRTX_GUIDE_PARAM_COUNT) return MRT_ERROR_INVALID_ARGUMENT;
    return mrt::guarded([&] {
        mrt::DeviceScope scope(n->device);
        MRT_HIP(hipStreamSynchronize(n->stream));
        n->net->setParams(src);
        return MRT_SUCCESS;
    });
}

MRT_API mrtResult mrtxGuideNetworkGetParams(mrtxGuideNetwork* n, float* dst, size_t count) {
    if (!n) return MRT_ERROR_INVALID_HANDLE;
    if (!dst || count != MRTX_GUIDE_PARAM_COUNT) return MRT_ERROR_INVALID_ARGUMENT;
    return mrt::guarded([&] {
        mrt::DeviceScope scope(n->device);
        MRT_HIP(hipStreamSynchronize(n->stream));
        n->net->getParams(dst);
        return MRT_SUCCESS;
    });
}

MRT_API mrtResult mrtxGuideNetworkGetAdam(mrtxGuideNetwork* n, float* m, float* v, size_t count, uint32_t* step) {
    if (!n) return MRT_ERROR_INVALID_HANDLE;
    if ((m || v) && count != MRTX_GUIDE_PARAM_COUNT) return MRT_ERROR_INVALID_ARGUMENT;
    return mrt::guarded([&] {
        mrt::DeviceScope scope(n->device);
        MRT_HIP(hipStreamSynchronize(n->stream));
        n->net->getAdam(m, v, step);
        return MRT_SUCCESS;
    });
}

MRT_API mrtResult mrtxGuideNetworkInfer(mrtxGuideNetwork* n, const float* records, uint32_t count, float* raw) {
    if (!n) return MRT_ERROR_INVALID_HANDLE;
    if (!records || !raw) return MRT_ERROR_INVALID_ARGUMENT;
    if (count == 0) return MRT_SUCCESS;
    return mrt::guarded([&] {
        mrt::DeviceScope scope(n->device);
        const mrt::dev::RecView v = stageRecords(n, records, count);
        n->out.ensure(size_t(count) * MRTX_GUIDE_OUTPUTS * sizeof(float));
        MRT_HIP(mrt::dev::launchGuideInfer(v, n->count.as<const uint32_t>(), count, n->net->deviceImage(),
                                           n->out.get(), false, n->stream));
        n->out.download(raw, size_t(count) * MRTX_GUIDE_OUTPUTS * sizeof(float), 0, n->stream);
        return MRT_SUCCESS;
    });
}

MRT_API mrtResult mrtxGuideNetworkGradient(mrtxGuideNetwork* n, const float* records, uint32_t count, float* grad,
                                           size_t gradCount, float* loss) {
    if (!n) return MRT_ERROR_INVALID_HANDLE;
    if (!records || count == 0 || (grad && gradCount != MRTX_GUIDE_PARAM_COUNT)) return MRT_ERROR_INVALID_ARGUMENT;
    return mrt::guarded([&] {
        mrt::DeviceScope scope(n->device);
        n->net->ensureCap(count);
        (void)stageRecords(n, records, count);
        const mrt::dev::GuideNet net = n->net->view();
        MRT_HIP(mrt::dev::launchGuideGradient(wholeBatch(n, count), net, false, n->stream));
        if (grad) {
            MRT_HIP(hipMemcpyAsync(grad, net.grad, MRTX_GUIDE_PARAM_COUNT * sizeof(float), hipMemcpyDeviceToHost,
                                   n->stream));
        }
        uint32_t w[mrt::dev::kGwWords];
        n->net->readWords(w, n->stream);
        if (loss) *loss = wordAsFloat(w, mrt::dev::kGwLoss);
        return MRT_SUCCESS;
    });
}

MRT_API mrtResult mrtxGuideNetworkTrain(mrtxGuideNetwork* n, const float* records, uint32_t count,
                                        float learningRate, float* loss) {
    if (!n) return MRT_ERROR_INVALID_HANDLE;
    if (!records || count == 0 || !(learningRate > 0.0f)) return MRT_ERROR_INVALID_ARGUMENT;
    return mrt::guarded([&] {
        mrt::DeviceScope scope(n->device);
        n->net->ensureCap(count);
        (void)stageRecords(n, records, count);
        const mrt::dev::GuideNet net = n->net->view();
        mrt::dev::GuideAdam h;
        h.lr = learningRate;
        MRT_HIP(mrt::dev::launchGuideStep(wholeBatch(n, count), net, h, n->stream));
        uint32_t w[mrt::dev::kGwWords];
        n->net->readWords(w, n->stream);
        if (loss) *loss = wordAsFloat(w, mrt::dev::kGwLastLoss);
        return MRT_SUCCESS;
    });
}

MRT_API mrtResult mrtxGuideNetworkTrainNoAlloc(mrtxGuideNetwork* n, const float* dRecords, uint32_t count,
                                               float learningRate, void* hipStream, float* loss) {
    if (!n) return MRT_ERROR_INVALID_HANDLE;
    if (!dRecords || count == 0 || !(learningRate > 0.0f)) return MRT_ERROR_INVALID_ARGUMENT;
    if (count > MRTX_GUIDE_NOALLOC_MAX_RECORDS) {
        mrt::setLastErrorMessage("mrtxGuideNetworkTrainNoAlloc: " + std::to_string(count) +
                                 " records exceed the fixed workspace of " +
                                 std::to_string(MRTX_GUIDE_NOALLOC_MAX_RECORDS));
        return MRT_ERROR_INVALID_ARGUMENT;
    }
    return mrt::guarded([&] {
        mrt::DeviceScope scope(n->device);
        hipStream_t s = hipStream ? static_cast<hipStream_t>(hipStream) : n->stream;
        mrt::dev::GuideAdam h;
        h.lr = learningRate;
        MRT_HIP(mrt::dev::launchGuideTrainNoAlloc(n->net->view(), reinterpret_cast<const float4*>(dRecords), count,
                                                  h, n->device, s));
        if (loss) {
            uint32_t w = 0;
            MRT_HIP(hipMemcpyAsync(&w, n->net->deviceWords() + mrt::dev::kGwLastLoss, sizeof(w),
                                   hipMemcpyDeviceToHost, s));
            MRT_HIP(hipStreamSynchronize(s));
            *loss = wordAsFloat(&w, 0);
        }
        return MRT_SUCCESS;
    });
}

MRT_API mrtResult mrtxGuideNetworkSample(mrtxGuideNetwork* n, const float* raw, const float* u, uint32_t count,
                                         float* out) {
    if (!n) return MRT_ERROR_INVALID_HANDLE;
    if (!raw || !u || !out) return MRT_ERROR_INVALID_ARGUMENT;
    if (count == 0) return MRT_SUCCESS;
    return mrt::guarded([&] {
        mrt::DeviceScope scope(n->device);
        n->a.ensure(size_t(count) * MRTX_GUIDE_OUTPUTS * sizeof(float));
        n->a.upload(raw, size_t(count) * MRTX_GUIDE_OUTPUTS * sizeof(float), 0, n->stream);
        n->b.ensure(size_t(count) * 3 * sizeof(float));
        n->b.upload(u, size_t(count) * 3 * sizeof(float), 0, n->stream);
        n->out.ensure(size_t(count) * 4 * sizeof(float));
        MRT_HIP(mrt::dev::launchGuideSample(n->a.as<const float>(), n->b.as<const float>(), count,
                                            n->out.as<float4>(), n->stream));
        n->out.download(out, size_t(count) * 4 * sizeof(float), 0, n->stream);
        return MRT_SUCCESS;
    });
}

MRT_API mrtResult mrtxGuideNetworkPdf(mrtxGuideNetwork* n, const float* raw, const float* dirs, uint32_t count,
                                      float* pdf) {
    if (!n) return MRT_ERROR_INVALID_HANDLE;
    if (!raw || !dirs || !pdf) return MRT_ERROR_INVALID_ARGUMENT;
    if (count == 0) return MRT_SUCCESS;
    return mrt::guarded([&] {
        mrt::DeviceScope scope(n->device);
        n->a.ensure(size_t(count) * MRTX_GUIDE_OUTPUTS * sizeof(float));
        n->a.upload(raw, size_t(count) * MRTX_GUIDE_OUTPUTS * sizeof(float), 0, n->stream);
        n->b.ensure(size_t(count) * 4 * sizeof(float));
        n->b.upload(dirs, size_t(count) * 4 * sizeof(float), 0, n->stream);
        n->out.ensure(size_t(count) * sizeof(float));
        MRT_HIP(mrt::dev::launchGuidePdf(n->a.as<const float>(), n->b.as<const float4>(), count,
                                         n->out.as<float>(), n->stream));
        n->out.download(pdf, size_t(count) * sizeof(float), 0, n->stream);
        return MRT_SUCCESS;
    });
}

} // extern "C"
