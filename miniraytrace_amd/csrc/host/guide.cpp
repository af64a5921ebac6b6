// GuideNetwork: device buffers of the path-guiding network.
#include "guide.h"

#include <cmath>
#include <random>
#include <vector>

namespace mrt {

namespace {

constexpr size_t kParamBytes = size_t(dev::kGuideParamCount) * sizeof(float);

// Layer shapes of the flat parameter vector: W [out x in] then b [out].
constexpr int kLayerIn[5] = { 78, 64, 64, 64, 64 };
constexpr int kLayerOut[5] = { 64, 64, 64, 64, 40 };

} // namespace

GuideNetwork::GuideNetwork(uint32_t cap) {
    params_.allocate(kParamBytes);
    m_.allocate(kParamBytes);
    v_.allocate(kParamBytes);
    grad_.allocate(kParamBytes);
    image_.allocate(dev::kGuideImageBytes + dev::kGuideImageTBytes);
    words_.allocate(dev::kGwWords * sizeof(uint32_t));
    grad_.zero();
    words_.zero();
    params_.zero();
    allocWorkspace(cap);
    resetAdam();
    MRT_HIP(dev::launchGuidePack(params_.as<float>(), image_.get(), imageT(), nullptr));
    MRT_HIP(hipDeviceSynchronize());
}

// Row-block-tiled workspace (dev::GuideNet), rows padded to a multiple of
// 32.  The input block has 96 features so the weight-gradient tiles of
// layer 0 cover it with three 32-feature tiles; the training kernel writes
// features 80..95 as zeros.
void GuideNetwork::allocWorkspace(uint32_t cap) {
    if (cap == 0) throw Error(MRT_ERROR_INVALID_ARGUMENT, "guide batch capacity must be > 0");
    const uint32_t stride = dev::guideRowStride(cap);
    actIn_.allocate(size_t(96) * stride * sizeof(uint16_t));
    actIn_.zero();
    act_.allocate(size_t(4) * 64 * stride * sizeof(uint16_t));
    tile_.allocate(size_t(stride / 32) * 2 * sizeof(float));
    // one delta block per layer, so that one launch forms every layer's
    // weight gradient after the backward pass
    for (DeviceBuffer& d : delta_) d.allocate(size_t(64) * stride * sizeof(uint16_t));
    // weight-gradient partials, then the per-chunk weight and loss sums
    partial_.allocate(size_t(dev::guideGradChunks(cap)) * (dev::kGuideParamCount + 2) * sizeof(float));
    MRT_HIP(hipStreamSynchronize(nullptr));
    cap_ = cap;
    stride_ = stride;
}

void GuideNetwork::ensureCap(uint32_t cap) {
    if (cap > cap_) allocWorkspace(cap);
}

void GuideNetwork::initParams(uint32_t seed) {
    std::vector<float> p(dev::kGuideParamCount, 0.0f);
    std::mt19937_64 rng(0x5EEDull ^ (uint64_t(seed) * 0x9E3779B97F4A7C15ull));
    size_t off = 0;
    for (int l = 0; l < 5; ++l) {
        const float lim = std::sqrt(6.0f / float(kLayerIn[l])) * (l == 4 ? 0.1f : 1.0f);
        std::uniform_real_distribution<float> u(-lim, lim);
        for (int i = 0; i < kLayerIn[l] * kLayerOut[l]; ++i) p[off++] = u(rng);
        off += kLayerOut[l];  // zero biases
    }
    setParams(p.data());
}

void GuideNetwork::setParams(const float* src) {
    params_.upload(src, kParamBytes);
    MRT_HIP(dev::launchGuidePack(params_.as<float>(), image_.get(), imageT(), nullptr));
    resetAdam();
}

void GuideNetwork::getParams(float* dst) const { params_.download(dst, kParamBytes); }

void GuideNetwork::getAdam(float* m, float* v, uint32_t* step) const {
    if (m) m_.download(m, kParamBytes);
    if (v) v_.download(v, kParamBytes);
    if (step) words_.download(step, sizeof(uint32_t), dev::kGwStep * sizeof(uint32_t));
}

void GuideNetwork::resetAdam() {
    m_.zero();
    v_.zero();
    grad_.zero();
    words_.zero();
    MRT_HIP(hipStreamSynchronize(nullptr));
}

void GuideNetwork::readWords(uint32_t out[dev::kGwWords], hipStream_t s) const {
    words_.download(out, dev::kGwWords * sizeof(uint32_t), 0, s);
}

dev::GuideNet GuideNetwork::view() const {
    dev::GuideNet n;
    n.params = params_.as<float>();
    n.adamM = m_.as<float>();
    n.adamV = v_.as<float>();
    n.grad = grad_.as<float>();
    n.image = image_.get();
    n.imageT = imageT();
    n.actIn = actIn_.as<uint16_t>();
    n.act = act_.as<uint16_t>();
    n.tilePart = tile_.as<float>();
    for (int l = 0; l < 5; ++l) n.delta[l] = delta_[l].as<uint16_t>();
    n.partial = partial_.as<float>();
    n.weightPart = n.partial + size_t(dev::guideGradChunks(cap_)) * dev::kGuideParamCount;
    n.words = words_.as<uint32_t>();
    n.cap = cap_;
    n.stride = stride_;
    return n;
}

} // namespace mrt
