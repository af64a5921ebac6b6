// Batched rendering and the path-tracer autotuner (mini_raytrace_ext.h,
// section 5).
#include "capi_handles.h"
#include "common.h"
#include "engine.h"
#include "mini_raytrace.h"
#include "mini_raytrace_ext.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <map>
#include <set>
#include <vector>

namespace {

using Clock = std::chrono::steady_clock;

float msSince(Clock::time_point t0) {
    return std::chrono::duration<float, std::milli>(Clock::now() - t0).count();
}

// Owns the graphs of one batch call.
struct BatchGraphs {
    std::vector<hipGraph_t> chains;
    hipGraph_t parent = nullptr;
    hipGraphExec_t exec = nullptr;
    ~BatchGraphs() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (parent) (void)hipGraphDestroy(parent);
        for (hipGraph_t g : chains)
            if (g) (void)hipGraphDestroy(g);
    }
};

// The launch shape that most of the batch's work was tuned for (weighted by
// pixels x frames of each request); guidePrep: the grid of the trace variant
// that also writes guided requests' query records.
mrt::dev::TraceShape voteShape(const std::vector<mrt::Engine*>& engines, const std::vector<uint32_t>& frames,
                               bool shadow, bool guidePrep) {
    constexpr int K = mrt::Engine::kShapeCandidates;
    // When every request was autotuned: the candidate with the least total
    // measured time, each request's per-frame time times its frames (a sum
    // over many requests evens out their single-frame measurements).
    std::array<double, K> total{};
    bool timed = true;
    for (size_t i = 0; i < engines.size() && timed; ++i) {
        float ms[K];
        timed = engines[i]->shapeTimes(ms);
        for (int k = 0; timed && k < K; ++k) total[k] += double(ms[k]) * frames[i];
    }
    mrt::dev::TraceShape t{ 0, 0, 0 };
    if (timed) {
        const int k = int(std::min_element(total.begin(), total.end()) - total.begin());
        t.block = mrt::Engine::kShapeCand[k][0];
        t.stack = mrt::Engine::kShapeCand[k][1];
    } else {
        // Otherwise the shape most of the work (pixels x frames) runs with.
        std::map<std::pair<uint32_t, uint32_t>, double> weight;
        for (size_t i = 0; i < engines.size(); ++i) {
            const mrt::dev::TraceShape& s = shadow ? engines[i]->shadowShape() : engines[i]->extendShape();
            weight[{ s.block, s.stack }] += double(engines[i]->pixels()) * frames[i];
        }
        auto best = std::max_element(weight.begin(), weight.end(),
                                     [](const auto& a, const auto& b) { return a.second < b.second; });
        t.block = best->first.first;
        t.stack = best->first.second;
    }
    int perCU = 0;
    MRT_HIP(mrt::dev::batchTraceBlocksPerCU(t.block, t.stack, guidePrep, &perCU));
    t.grid = uint32_t(std::max(1, perCU) * engines[0]->cuCount());
    return t;
}

// Mode 0: all requests' queues served by one set of persistent kernels per
// batched frame.  Guided requests add their own passes to the frame: the
// network inference of their query records after each trace launch and the
// training after the frame (each request has its own network).
void renderMerged(const std::vector<mrt::Engine*>& engines, std::vector<uint32_t>& frames, mrtxBatchStats& st) {
    const size_t n = engines.size();
    std::vector<const mrt::FrameConsts*> consts(n, nullptr);
    for (size_t i = 0; i < n; ++i) frames[i] = engines[i]->prepareFrames(frames[i], &consts[i]);
    const uint32_t F = *std::max_element(frames.begin(), frames.end());
    if (F == 0) return;
    std::vector<mrt::dev::BatchEntry> table(size_t(F) * n);
    uint32_t slots = 0, levels = 0;
    std::vector<uint32_t> base(n);
    for (size_t i = 0; i < n; ++i) {
        base[i] = slots;
        slots += (engines[i]->pixels() + 63u) & ~63u;
        levels = std::max(levels, engines[i]->bounceLevels());
    }
    for (uint32_t f = 0; f < F; ++f)
        for (size_t i = 0; i < n; ++i) {
            mrt::dev::BatchEntry& e = table[size_t(f) * n + i];
            e.sc = engines[i]->sceneView();
            e.pv = engines[i]->pathView();
            e.g = frames[i] > 0 ? consts[i] + std::min(f, frames[i] - 1) : consts[0];
            e.pixels = engines[i]->pixels();
            e.pixelBase = base[i];
            e.active = f < frames[i] ? 1u : 0u;
            e.tailAt = engines[i]->tailAt();
        }
    bool guided = false;
    for (mrt::Engine* e : engines) guided = guided || e->guideFrame();
    mrt::DeviceBuffer dTable, dCursors;
    dTable.allocate(table.size() * sizeof(mrt::dev::BatchEntry));
    dTable.upload(table.data(), table.size() * sizeof(mrt::dev::BatchEntry));
    dCursors.allocate(mrt::dev::kCursorWords * sizeof(uint32_t));
    const mrt::dev::TraceShape ext = voteShape(engines, frames, false, guided);
    const mrt::dev::TraceShape sha = voteShape(engines, frames, true, false);
    const uint32_t shadeGrid = engines[0]->shadeGrid(guided);
    const mrt::dev::TraceShape tail = engines[0]->tailShape();
    // The tail threshold of the first request (the batch shares one tail).
    const uint32_t tailPaths = engines[0]->tailPaths();
    hipStream_t s = engines[0]->stream();
    BatchGraphs g;
    g.chains.assign(1, nullptr);
    MRT_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
        for (uint32_t f = 0; f < F; ++f) {
            mrt::dev::BatchView bv;
            bv.entries = dTable.as<const mrt::dev::BatchEntry>() + size_t(f) * n;
            bv.count = uint32_t(n);
            bv.pixelSlots = slots;
            bv.cursors = dCursors.as<uint32_t>();
            MRT_HIP(mrt::dev::launchClearBatch(bv, s));
            for (size_t i = 0; i < n; ++i)
                if (f < frames[i]) engines[i]->enqueueGuideFrameBegin(s, consts[i] + f);
            MRT_HIP(mrt::dev::launchGenerateBatch(bv, s));
            for (uint32_t b = 0; b < levels; ++b) {
                MRT_HIP(mrt::dev::launchTraceBatch(bv, b, b == 0 ? mrt::dev::kNoLevel : b - 1, ext, tailPaths, guided,
                                                   s));
                for (size_t i = 0; i < n; ++i)
                    if (f < frames[i]) engines[i]->enqueueGuideInfer(s, b);
                MRT_HIP(mrt::dev::launchShadeBatch(bv, b, tailPaths, guided, shadeGrid, s));
                MRT_HIP(mrt::dev::launchTailBatch(bv, b, tailPaths, tail, s));
            }
            MRT_HIP(mrt::dev::launchTraceBatch(bv, mrt::dev::kNoLevel, levels - 1, sha, tailPaths, false, s));
            for (size_t i = 0; i < n; ++i)
                if (f < frames[i]) engines[i]->enqueueGuideTraining(s, consts[i] + f, f);
            MRT_HIP(mrt::dev::launchFinishBatch(bv, s));
        }
    } catch (...) {
        hipGraph_t g2 = nullptr;
        (void)hipStreamEndCapture(s, &g2);
        if (g2) (void)hipGraphDestroy(g2);
        throw;
    }
    MRT_HIP(hipStreamEndCapture(s, &g.chains[0]));
    MRT_HIP(hipGraphInstantiate(&g.exec, g.chains[0], nullptr, nullptr, 0));
    const auto t1 = Clock::now();
    MRT_HIP(hipGraphLaunch(g.exec, s));
    MRT_HIP(hipStreamSynchronize(s));
    st.launchMs = msSince(t1);
    st.mode = 0;
    st.extendBlock = ext.block;
    st.extendStack = ext.stack;
    st.shadowBlock = sha.block;
    st.shadowStack = sha.stack;
}

// Mode 1: one captured chain per request, chains as independent graph nodes.
void renderChains(const std::vector<mrt::Engine*>& engines, std::vector<uint32_t>& frames, mrtxBatchStats& st) {
    const size_t n = engines.size();
    BatchGraphs g;
    g.chains.assign(n, nullptr);
    for (size_t i = 0; i < n; ++i) frames[i] = engines[i]->captureFrames(frames[i], &g.chains[i]);
    MRT_HIP(hipGraphCreate(&g.parent, 0));
    bool any = false;
    for (size_t i = 0; i < n; ++i) {
        if (!g.chains[i]) continue;
        hipGraphNode_t node = nullptr;
        MRT_HIP(hipGraphAddChildGraphNode(&node, g.parent, nullptr, 0, g.chains[i]));
        any = true;
    }
    st.mode = 1;
    if (!any) return;
    MRT_HIP(hipGraphInstantiate(&g.exec, g.parent, nullptr, nullptr, 0));
    hipStream_t s = engines[0]->stream();
    const auto t1 = Clock::now();
    MRT_HIP(hipGraphLaunch(g.exec, s));
    MRT_HIP(hipStreamSynchronize(s));
    st.launchMs = msSince(t1);
}

} // namespace

extern "C" {

MRT_API mrtResult mrtxRenderBatch(mrtxBatchRequest* reqs, uint32_t count, mrtxBatchStats* stats) {
    if (count > 0 && !reqs) return MRT_ERROR_INVALID_ARGUMENT;
    if (stats && stats->structSize < sizeof(mrtxBatchStats)) return MRT_ERROR_INVALID_ARGUMENT;
    std::set<const mrtEngine*> seen;
    int device = -1;
    for (uint32_t i = 0; i < count; ++i) {
        mrt::Engine* e = mrt::unwrap(reqs[i].engine);
        if (!e) return MRT_ERROR_INVALID_HANDLE;
        if (!seen.insert(reqs[i].engine).second) {
            mrt::setLastErrorMessage("an engine appears twice in the batch");
            return MRT_ERROR_INVALID_ARGUMENT;
        }
        if (device >= 0 && e->device() != device) {
            mrt::setLastErrorMessage("batched engines must share one device");
            return MRT_ERROR_INVALID_ARGUMENT;
        }
        device = e->device();
        reqs[i].framesDone = 0;
    }
    return mrt::guarded([&] {
        mrtxBatchStats st{};
        st.structSize = sizeof(st);
        if (count > 0) {
            mrt::DeviceScope scope(device);
            const auto t0 = Clock::now();
            std::vector<mrt::Engine*> engines(count);
            std::vector<uint32_t> frames(count);
            for (uint32_t i = 0; i < count; ++i) {
                engines[i] = mrt::unwrap(reqs[i].engine);
                frames[i] = reqs[i].frames;
            }
            if (count <= mrt::dev::kMaxBatchRequests) renderMerged(engines, frames, st);
            else renderChains(engines, frames, st);
            st.captureMs = msSince(t0) - st.launchMs;
            for (uint32_t i = 0; i < count; ++i) {
                engines[i]->framesDone(frames[i]);
                reqs[i].framesDone = frames[i];
                if (frames[i] > 0) ++st.requests;
                st.frames += frames[i];
            }
        }
        if (stats) *stats = st;
        return MRT_SUCCESS;
    });
}

MRT_API mrtResult mrtxAutotunePath(mrtEngine* e, mrtxPathOptions* chosen) {
    mrt::Engine* eng = mrt::unwrap(e);
    if (!eng) return MRT_ERROR_INVALID_HANDLE;
    if (chosen && chosen->structSize < sizeof(mrtxPathOptions)) return MRT_ERROR_INVALID_ARGUMENT;
    return mrt::guarded([&] {
        mrtxPathOptions o{};
        o.structSize = sizeof(o);
        eng->autotuneShapes(o);
        if (chosen) *chosen = o;
        return MRT_SUCCESS;
    });
}

} // extern "C"
