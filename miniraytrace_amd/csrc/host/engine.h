// Engine: scene + device resources + the wavefront frame loop behind the C
// ABI.  Responsibilities match the reference's Engine/Renderer pair
// (core/src/Engine.cpp:44-60, core/src/render/Renderer.cpp:290-909); the GPU
// side is the HIP pipeline in kernels/pathtrace.hip.
#pragma once

#include "bvh_build.h"
#include "bvh_gpu_build.h"
#include "bvh_wide.h"
#include "common.h"
#include "common/gpu_types.h"
#include "guide.h"
#include "kernels/pt_launch.h"
#include "mini_raytrace_ext.h"
#include "scene.h"

#include <algorithm>
#include <map>
#include <memory>
#include <vector>

namespace mrt {

// Default launch shape of the path tracer's trace and tail kernels (threads
// per workgroup, LDS traversal-stack entries per lane).
constexpr uint32_t kDefaultTraceBlock = 256;
constexpr uint32_t kDefaultTraceStack = 8;
// Tail kernel shape: its registers (shading and traversal in one kernel)
// allow 3 waves per SIMD, which 128-thread blocks with 16 LDS stack entries
// per lane (16 KB per block) keep.
constexpr uint32_t kTailBlock = 128;
constexpr uint32_t kTailStack = 16;
// Default tail threshold: the tail kernel takes over a bounce level once all
// requests together hold at most this many rays per compute unit there
// (measured on the 1.256M-triangle terrain at 1080p: the tail then covers the
// levels whose launches are dominated by their slowest rays).
constexpr uint32_t kTailPathsPerCU = 1280;
// Upper bound of the guiding path-record buffer (levels x recorded paths x
// 64 B); 1280x720 at 9 levels records every path within it.
constexpr uint64_t kPathRecordBytes = 1ull << 30;

struct Settings {
    uint32_t width = 1280, height = 720;
    uint32_t maxBounces = 8;
    uint32_t sppLimit = 1024;
    float exposureEV = 0.0f;
    float fireflyClamp = 50.0f;
    uint32_t tonemap = MRT_TONEMAP_ACES;
    bool denoise = true;
};

struct FrameStats {
    float frameMs = 0.0f;
    float commitMs = 0.0f;        // last commit that changed something
    uint64_t commitUploadBytes = 0;
    uint32_t geometryRepacks = 0;  // commits that repacked every mesh (since creation)
    uint32_t meshesWritten = 0;    // meshes rewritten alone, in place or appended (since creation)
    float stageMs[8] = {};  // generate, extend, shade, shadow, accumulate + resolve, tail, guide, train
    uint32_t rays[dev::kMaxBounceLevels] = {};
    uint32_t shadowRays[dev::kMaxBounceLevels] = {};
    uint32_t status = 0;
    uint32_t guideRows = 0;  // guiding-network query records of the frame (all levels)
    uint32_t graphTailLevel = ~0u;  // tailLevel of the graph the last frame replayed (~0u: none)
};

class Engine {
public:
    explicit Engine(const mrtEngineDesc& desc);
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    Scene& scene() { return scene_; }
    void setSettings(const Settings& s);
    const Settings& settings() const { return settings_; }
    void commit(mrtCommitStats* out);
    void renderFrame(mrtFrameInfo* out);
    void readFramebuffer(mrtPixelFormat fmt, void* dst, size_t dstSize);
    void resetAccumulation() { frameIndex_ = 0; }
    uint32_t frameIndex() const { return frameIndex_; }

    // extensions
    void setPathOptions(const mrtxPathOptions& o);
    void getPathOptions(mrtxPathOptions& o) const;
    void setProfiling(bool on) { profiling_ = on; }
    const FrameStats& stats() const { return stats_; }
    void exportScene(mrtxSceneExport& out);
    void setGuiding(const mrtxGuidingOptions& o);
    void getGuiding(mrtxGuidingOptions& o) const { o = guide_; }
    void guidingStats(mrtxGuidingStats& o);
    void getGuideParams(float* dst);
    void setGuideParams(const float* src);
    // batching
    uint32_t captureFrames(uint32_t frames, hipGraph_t* out);
    uint32_t prepareFrames(uint32_t frames, const FrameConsts** consts);
    dev::SceneView sceneView() const;
    dev::PathView pathView() const;
    bool guideActive() const { return guide_.enableSampling || guide_.enableTraining; }
    const dev::TraceShape& extendShape() const { return extendShape_; }
    const dev::TraceShape& shadowShape() const { return shadowShape_; }
    uint32_t tailPaths() const { return tailPaths_; }
    // Guiding passes of a frame (no-ops for an unguided frame), also used by
    // the merged batch for each of its requests.
    bool guideFrame() const;
    void enqueueGuideFrameBegin(hipStream_t s, const FrameConsts* g);
    void enqueueGuideInfer(hipStream_t s, uint32_t bounce);
    // offset: the frame's position after the next frame to render (captured
    // batches enqueue several frames before any of them runs).
    void enqueueGuideTraining(hipStream_t s, const FrameConsts* g, uint32_t offset = 0);
    // BatchEntry::tailAt of this engine's requests.
    uint32_t tailAt() const;
    // Trace-launch shape (guidePrep: the variant that writes query records).
    dev::TraceShape traceShape(bool guidePrep) const;
    uint32_t shadeGrid(bool guided) const { return guided ? shadeGuideGrid_ : shadeGrid_; }
    dev::TraceShape tailShape() const { return dev::TraceShape{ tailBlock_, tailStack_, tailGrid_ }; }
    uint32_t pixels() const { return frameW_ * frameH_; }
    uint32_t bounceLevels() const { return std::min(settings_.maxBounces, 16u) + 1; }
    int cuCount() const { return cuCount_; }
    void framesDone(uint32_t frames);
    void autotuneShapes(mrtxPathOptions& chosen);
    // Launch shapes the autotuner tries (block, LDS stack depth), and this
    // engine's measured trace + tail time per frame for each (false until
    // the engine has been autotuned).
    static constexpr int kShapeCandidates = 6;
    static const uint32_t kShapeCand[kShapeCandidates][2];
    bool shapeTimes(float out[kShapeCandidates]) const;
    int device() const { return device_; }
    hipStream_t stream() const { return stream_; }

private:
    struct Blas {
        std::vector<BvhNode> nodes;
        std::vector<WideNode> wide;  // four-wide copy (local indices)
        std::vector<uint32_t> triOrder;
        Box3 bounds;
        uint32_t depth = 0;
        uint32_t rootGlobal = 0;
        uint32_t wideRootGlobal = 0;
    };
    // A mesh's ranges in the vertex, triangle, node and wide-node arrays:
    // first index and capacity (its roots are node nb and wide node wb).
    struct MeshSlot {
        uint32_t vb = 0, vcap = 0, tb = 0, tcap = 0, nb = 0, ncap = 0, wb = 0, wcap = 0;
    };
    struct Holes {  // entries of the four arrays no mesh uses
        uint64_t verts = 0, tris = 0, nodes = 0, wides = 0;
        void add(const MeshSlot& s) {
            verts += s.vcap;
            tris += s.tcap;
            nodes += s.ncap;
            wides += s.wcap;
        }
    };
    struct GeomArrays {
        std::vector<float> pos, nrm, uv, tan, triVerts;
        std::vector<GpuTriangle> tris;
        std::vector<BvhNode> nodes;
        std::vector<WideNode> wide;
    };
    void appendMesh(const MeshData& m, Blas& b, const MeshSlot& sl, GeomArrays& g);
    void repackGeometry();
    bool placeIncremental(const std::vector<uint32_t>& dirty);

    void sync(mrtCommitStats* out);
    uint32_t packGeometry();
    void packInstances();
    void packMaterials();
    void packLights();
    void packEnvironment();
    void uploadTextures();
    void allocFrame();
    void fillConsts(FrameConsts& g) const;
    // tailLevel: the tail kernel takes every path left at that level and the
    // frame ends there (~0u: per-level launches to the last bounce).
    void enqueueFrame(hipStream_t s, bool timed, const FrameConsts* g, const dev::BatchEntry* entry,
                      uint32_t offset = 0, uint32_t tailLevel = ~0u);
    uint32_t fittedTailLevel() const;
    bool tailLevelFits(uint32_t level) const;
    // Adam steps of the frame `offset` frames after the next one (0 when
    // the engine does not train): the warm-up's count for the network's
    // first warmupFrames trained frames, then stepsPerFrame.
    uint32_t trainSteps(uint32_t offset) const;
    void fillEntries(dev::BatchEntry* out, const FrameConsts* g0, uint32_t frames) const;
    void invalidateGraph();
    void chooseShapes();
    uint32_t guideLevels() const { return std::min(settings_.maxBounces, 16u) + 1; }
    // Slots of each ray and shadow-ray queue (kQueueSegs segments of the frame's paths).
    size_t queueSlots() const { return size_t(dev::kQueueSegs) * dev::queueSegCap(frameW_ * frameH_); }
    void ensureGuideNet();
    void allocGuide();


    int device_ = 0;
    int cuCount_ = 0;
    hipStream_t stream_ = nullptr;
    Scene scene_;
    Settings settings_;
    uint32_t frameIndex_ = 0;
    uint32_t seed_ = 0;
    bool useGraph_ = true;
    bool packetCamera_ = true;  // camera rays by wave packets (mrtxPathOptions::cameraPerLane == 0)
    bool profiling_ = false;
    FrameStats stats_;

    // packing state
    std::map<uint32_t, Blas> blas_;
    std::map<uint32_t, uint32_t> materialIndex_;
    std::map<uint32_t, MeshSlot> slots_;
    Holes holes_;
    bool layoutValid_ = false;             // geometry arrays packed (slots_ valid)
    uint32_t vertEnd_ = 0, triEnd_ = 0;    // first entries behind every slot
    uint32_t nodeEnd_ = 0;                 // first node behind every slot and the TLAS region
    uint32_t tlasCap_ = 0;                 // nodes reserved for the TLAS at tlasRoot_
    uint32_t wideEnd_ = 0;                 // first wide node behind every slot and the TLAS's wide region
    uint32_t tlasWideRoot_ = 0, tlasWideCap_ = 0;  // the TLAS's four-wide copy: first node, reserved nodes
    uint32_t geometryRepacks_ = 0, meshesWritten_ = 0;  // since creation
    uint64_t uploadBytes_ = 0;             // by the last commit
    uint32_t tlasRoot_ = 0;
    uint32_t instanceCount_ = 0;
    uint32_t tlasNodeCount_ = 0;
    uint32_t maxBlasDepth_ = 0;
    uint32_t tlasDepth_ = 0;
    uint32_t lightCount_ = 0;
    bool hasEnv_ = false;
    uint32_t envW_ = 0, envH_ = 0;
    float envIntegral_ = 1.0f;
    size_t totalTris_ = 0, totalVerts_ = 0;

    // scene buffers
    GpuBvhBuilder gpuBvh_;  // large meshes' BLAS builds (workspace kept between commits)
    DeviceBuffer dNodes_, dWide_, dTriVerts_, dTris_, dPos_, dNrm_, dUv_, dTan_;
    DeviceBuffer dInst_, dTraceInst_, dMats_, dLights_, dEnv_, dAlias_, dEnvA_, dEnvB_;
    float envSplitA_ = 0.0f;
    uint32_t envCountA_ = 0, envCountB_ = 0;
    DeviceBuffer dTexTable_, dTex8_, dTex32_;
    size_t tex8Texels_ = 0, tex32Texels_ = 0;  // texels held by the two pools
    // frame buffers
    DeviceBuffer dConsts_, dBeta_, dRad_, dRng_, dRayO_[2], dRayD_[2], dHit_, dHitInst_;
    DeviceBuffer dShO_, dShD_, dShC_, dCounters_, dAccum_, dOut8_, dOut32_;
    PinnedBuffer hConsts_, hReadback_, hCounters_, hBatchConsts_, hEntry_;
    DeviceBuffer dBatchConsts_, dBatchEntries_, dEntry_, dCursors_;
    uint32_t frameW_ = 0, frameH_ = 0;

    dev::TraceShape extendShape_{ kDefaultTraceBlock, kDefaultTraceStack, 0 };
    dev::TraceShape shadowShape_{ kDefaultTraceBlock, kDefaultTraceStack, 0 };
    uint32_t tailOption_ = 0;  // mrtxPathOptions::tailPaths as set
    uint32_t tailPaths_ = 0;   // tail kernel threshold passed to the kernels (0 = never)
    uint32_t tailGrid_ = 0;
    uint32_t prepTraceGrid_ = 0;  // grid of the trace launches that also write guide query records
    uint32_t tailBlock_ = kTailBlock, tailStack_ = kTailStack;
    uint32_t shadeGrid_ = 0;
    uint32_t shadeGuideGrid_ = 0;

    // path guiding
    mrtxGuidingOptions guide_{};
    std::unique_ptr<GuideNetwork> guideNet_;
    DeviceBuffer dGuideIdx_, dGuideRec_, dGuideMix_, dPathRec_, dRecMask_, dGuidePool_;
    DeviceBuffer dShX_, dNeeStats_;  // learned environment-stratum selection
    DeviceBuffer dRrScale_;                      // guided paths: roulette throughput scale
    uint32_t guidePaths_ = 0, guideLevels_ = 0;
    uint32_t guideTrainStride_ = 1;  // every guideTrainStride_-th path writes training records
    uint32_t guideStamp_ = 0;
    uint32_t trainedFrames_ = 0;   // frames trained since the network was created or its parameters set
    float shapeMs_[kShapeCandidates] = {};
    bool shapeTimed_ = false;
    uint32_t graphTrainSteps_ = 0; // Adam steps per frame in the captured frame graph
    uint32_t graphTailLevel_ = ~0u;  // tailLevel of the captured frame graph
    PinnedBuffer hLevelRays_;        // rays queued per level in the last frame (written by the GPU)
    bool levelRaysFresh_ = false;    // hLevelRays_ holds a frame of the current frame setup
    uint32_t lastPoolCap_ = 0;     // training pool size of the last frame enqueued
    float sceneBoxMin_[3] = { 0.0f, 0.0f, 0.0f };
    float sceneBoxScale_ = 1.0f;

    hipGraph_t graph_ = nullptr;
    hipGraphExec_t graphExec_ = nullptr;
    hipEvent_t ev_[16] = {};
};

} // namespace mrt
