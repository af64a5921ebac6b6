// Path-guiding network state on one HIP device: FP32 master parameters,
// Adam moments, the training workspace and the device state words.  Used by
// the engine (guided rendering and online training) and by the standalone
// network API of mini_raytrace_ext.h.
#pragma once

#include "common.h"
#include "kernels/guide_launch.h"

#include <cstdint>

namespace mrt {

class GuideNetwork {
public:
    // Workspace for batches of up to `cap` rows, on the current device.
    explicit GuideNetwork(uint32_t cap);

    uint32_t cap() const { return cap_; }
    void ensureCap(uint32_t cap);

    // He-uniform weights from a seeded generator (the output layer scaled by
    // 0.1 so an untrained network is a set of broad lobes), zero biases.
    void initParams(uint32_t seed);
    void setParams(const float* src);  // kGuideParamCount floats; resets Adam
    void getParams(float* dst) const;
    void getAdam(float* m, float* v, uint32_t* step) const;
    void resetAdam();
    void readWords(uint32_t out[dev::kGwWords], hipStream_t s) const;

    dev::GuideNet view() const;
    const void* deviceImage() const { return image_.get(); }
    void* imageT() const { return static_cast<char*>(image_.get()) + dev::kGuideImageBytes; }
    uint32_t* deviceWords() const { return words_.as<uint32_t>(); }
    float* deviceGrad() const { return grad_.as<float>(); }

private:
    void allocWorkspace(uint32_t cap);

    uint32_t cap_ = 0, stride_ = 0;
    DeviceBuffer params_, m_, v_, grad_, image_, words_;
    DeviceBuffer actIn_, act_, tile_, delta_[5], partial_;
};

} // namespace mrt
