// Engine implementation: scene packing, device buffers, frame loop, readback.
#include "engine.h"

#include "bvh_gpu_build.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <mutex>
#include <thread>
#include <tuple>

namespace mrt {

namespace {

using Clock = std::chrono::steady_clock;

float msSince(Clock::time_point t0) {
    return std::chrono::duration<float, std::milli>(Clock::now() - t0).count();
}

// Deepest stack a traversal can need: TLAS depth + instances pushed from one
// TLAS leaf + BLAS depth.  The kernels hold 64 scratch entries plus the LDS
// part (at least 8).
constexpr uint32_t kMaxTraversalStack = 64 + 8;
// Traversal stack entries hold node, triangle and instance indices in 27 bits.
constexpr size_t kMaxTraversalIndex = (size_t(1) << 27) - 1;
// Meshes with at least this many triangles build their BVH on the GPU.
constexpr uint32_t kGpuBuildMinTriangles = 1u << 17;

// A node whose inverted box fails every slab test (reference Renderer.cpp:51-58).
BvhNode sentinelNode() {
    BvhNode n{};
    for (int a = 0; a < 3; ++a) {
        n.lo[a] = std::numeric_limits<float>::max();
        n.hi[a] = -std::numeric_limits<float>::max();
    }
    n.leftOrPrim = 0;
    n.packed = 1u << 2;
    return n;
}

// Column-major m (4x4) times (x, y, z, 1), in glm's evaluation order:
// (c0*x + c1*y) + (c2*z + c3).
void xformPoint(const float m[16], const float p[3], float out[3]) {
    for (int i = 0; i < 3; ++i)
        out[i] = (m[0 + i] * p[0] + m[4 + i] * p[1]) + (m[8 + i] * p[2] + m[12 + i] * 1.0f);
}

// Conservative world box of an object box: all 8 corners transformed
// (reference MathTypes.hpp:38-48).
Box3 transformBox(const Box3& b, const float m[16]) {
    Box3 out;
    for (int i = 0; i < 8; ++i) {
        const float c[3] = { (i & 1) ? b.hi[0] : b.lo[0], (i & 2) ? b.hi[1] : b.lo[1], (i & 4) ? b.hi[2] : b.lo[2] };
        float w[3];
        xformPoint(m, c, w);
        out.grow(w);
    }
    return out;
}

// General 4x4 inverse in double (Gauss-Jordan with partial pivoting),
// rounded to float.  Column-major in and out.
void invert4(const float m[16], float out[16]) {
    double a[4][8];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            a[r][c] = m[c * 4 + r];
            a[r][c + 4] = (r == c) ? 1.0 : 0.0;
        }
    for (int col = 0; col < 4; ++col) {
        int piv = col;
        for (int r = col + 1; r < 4; ++r)
            if (std::fabs(a[r][col]) > std::fabs(a[piv][col])) piv = r;
        if (piv != col)
            for (int c = 0; c < 8; ++c) std::swap(a[col][c], a[piv][c]);
        const double p = a[col][col];
        const double inv = p != 0.0 ? 1.0 / p : 0.0;  // singular: zero inverse, rays miss
        for (int c = 0; c < 8; ++c) a[col][c] *= inv;
        for (int r = 0; r < 4; ++r) {
            if (r == col) continue;
            const double f = a[r][col];
            if (f != 0.0)
                for (int c = 0; c < 8; ++c) a[r][c] -= f * a[col][c];
        }
    }
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) out[c * 4 + r] = float(a[r][c + 4]);
}

// Vose's alias method over non-negative weights (double).
void buildAlias(const std::vector<double>& w, std::vector<AliasEntry>& out) {
    const size_t n = w.size();
    out.assign(n, AliasEntry{ 1.0f, 0u });
    double total = 0.0;
    for (double x : w) total += x;
    if (n == 0) return;
    if (!(total > 0.0)) {
        for (size_t i = 0; i < n; ++i) out[i] = AliasEntry{ 1.0f, uint32_t(i) };
        return;
    }
    std::vector<double> q(n);
    std::vector<uint32_t> small, large;
    small.reserve(n);
    large.reserve(n);
    for (size_t i = 0; i < n; ++i) {
        q[i] = w[i] * double(n) / total;
        (q[i] < 1.0 ? small : large).push_back(uint32_t(i));
    }
    while (!small.empty() && !large.empty()) {
        const uint32_t s = small.back();
        small.pop_back();
        const uint32_t l = large.back();
        out[s] = AliasEntry{ float(q[s]), l };
        q[l] = (q[l] + q[s]) - 1.0;
        if (q[l] < 1.0) {
            large.pop_back();
            small.push_back(l);
        }
    }
    for (uint32_t i : large) out[i] = AliasEntry{ 1.0f, i };
    for (uint32_t i : small) out[i] = AliasEntry{ 1.0f, i };
}

void uploadVec(DeviceBuffer& b, const void* data, size_t bytes) {
    b.ensure(std::max<size_t>(bytes, 16));
    if (bytes) b.upload(data, bytes);
}

} // namespace

// ------------------------------------------------------------------ lifecycle --
Engine::Engine(const mrtEngineDesc& desc) {
    device_ = selectDevice(desc.gpuIndex1);
    DeviceScope scope(device_);
    MRT_HIP(hipDeviceGetAttribute(&cuCount_, hipDeviceAttributeMultiprocessorCount, device_));
    MRT_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    for (auto& e : ev_) MRT_HIP(hipEventCreate(&e));
    if (desc.width > 0 && desc.height > 0) {
        settings_.width = desc.width;
        settings_.height = desc.height;
    }
    guide_.structSize = sizeof(guide_);
    guide_.selectProb = 0.5f;
    guide_.learningRate = 1e-3f;
    guide_.batchSize = 65536;
    guide_.stepsPerFrame = 4;
    guide_.warmupFrames = 32;
    guide_.warmupStepsPerFrame = 16;
    guide_.minRoughness = 0.2f;
    guide_.seed = 1;
    guide_.positionScale = 1.0f;
    // Minimal scene: default material, no instances, empty texture table.
    scene_.markDirty(kDirtyMaterials | kDirtyTextures | kDirtyLights | kDirtyGeometry | kDirtyInstances);
    allocFrame();
    chooseShapes();
    sync(nullptr);
    frameIndex_ = 0;
    logMessage(LogLevel::Info, "engine on HIP device " + std::to_string(device_) + " (" +
                                   std::to_string(cuCount_) + " CUs)");
}

Engine::~Engine() {
    DeviceScope scope(device_);
    if (stream_) (void)hipStreamSynchronize(stream_);
    invalidateGraph();
    for (auto& e : ev_)
        if (e) (void)hipEventDestroy(e);
    if (stream_) (void)hipStreamDestroy(stream_);
}

void Engine::invalidateGraph() {
    if (graphExec_) (void)hipGraphExecDestroy(graphExec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    graphExec_ = nullptr;
    graph_ = nullptr;
}

// Persistent grids: as many workgroups as fit on the device at once.
void Engine::chooseShapes() {
    int perCU = 0;
    MRT_HIP(dev::batchTraceBlocksPerCU(extendShape_.block, extendShape_.stack, false, &perCU));
    extendShape_.grid = uint32_t(std::max(1, perCU) * cuCount_);
    MRT_HIP(dev::batchTraceBlocksPerCU(extendShape_.block, extendShape_.stack, true, &perCU));
    prepTraceGrid_ = uint32_t(std::max(1, perCU) * cuCount_);
    MRT_HIP(dev::batchTraceBlocksPerCU(shadowShape_.block, shadowShape_.stack, false, &perCU));
    shadowShape_.grid = uint32_t(std::max(1, perCU) * cuCount_);
    MRT_HIP(dev::batchShadeBlocksPerCU(false, &perCU));
    shadeGrid_ = uint32_t(std::max(1, perCU) * cuCount_);
    MRT_HIP(dev::batchTailBlocksPerCU(tailBlock_, tailStack_, &perCU));
    tailGrid_ = uint32_t(std::max(1, perCU) * cuCount_);
    tailPaths_ = tailOption_ == MRTX_TAIL_DISABLED ? 0u
                 : tailOption_ != 0                ? tailOption_
                                                   : kTailPathsPerCU * uint32_t(cuCount_);
    MRT_HIP(dev::batchShadeBlocksPerCU(true, &perCU));
    shadeGuideGrid_ = uint32_t(std::max(1, perCU) * cuCount_);
    invalidateGraph();
}

// Batch-table entries of this engine for frames using consts g0, g0 + 1, ...
// (an engine's own frame runs the batch kernels as a batch of one).
void Engine::fillEntries(dev::BatchEntry* out, const FrameConsts* g0, uint32_t frames) const {
    for (uint32_t i = 0; i < frames; ++i) {
        dev::BatchEntry& e = out[i];
        std::memset(&e, 0, sizeof(e));
        e.sc = sceneView();
        e.pv = pathView();
        e.g = g0 + i;
        e.pixels = frameW_ * frameH_;
        e.pixelBase = 0;
        e.active = 1;
        e.tailAt = tailAt();
    }
}

// ------------------------------------------------------------------- settings --
void Engine::setSettings(const Settings& s) {
    if (s.width == 0 || s.height == 0) throw Error(MRT_ERROR_INVALID_ARGUMENT, "width and height must be > 0");
    const bool resize = s.width != settings_.width || s.height != settings_.height;
    const bool reset = resize || s.maxBounces != settings_.maxBounces || s.fireflyClamp != settings_.fireflyClamp;
    const bool levels = std::min(s.maxBounces, 16u) != std::min(settings_.maxBounces, 16u);
    if (levels) invalidateGraph();
    settings_ = s;
    if (resize) {
        DeviceScope scope(device_);
        MRT_HIP(hipStreamSynchronize(stream_));
        allocFrame();
    } else if (levels && guideActive()) {
        DeviceScope scope(device_);
        MRT_HIP(hipStreamSynchronize(stream_));
        allocGuide();
    }
    if (reset) resetAccumulation();
}

void Engine::setPathOptions(const mrtxPathOptions& o) {
    seed_ = o.seed;
    useGraph_ = o.useGraph != 0;
    packetCamera_ = o.cameraPerLane == 0;
    const uint32_t eb = o.extendBlock ? o.extendBlock : kDefaultTraceBlock;
    const uint32_t es = o.extendStack ? o.extendStack : kDefaultTraceStack;
    const uint32_t sb = o.shadowBlock ? o.shadowBlock : kDefaultTraceBlock;
    const uint32_t ss = o.shadowStack ? o.shadowStack : kDefaultTraceStack;
    if (!dev::ptShapeSupported(eb, es) || !dev::ptShapeSupported(sb, ss))
        throw Error(MRT_ERROR_INVALID_ARGUMENT, "unsupported traversal launch shape");
    extendShape_.block = eb;
    extendShape_.stack = es;
    shadowShape_.block = sb;
    shadowShape_.stack = ss;
    tailOption_ = o.tailPaths;
    DeviceScope scope(device_);
    chooseShapes();
}

void Engine::getPathOptions(mrtxPathOptions& o) const {
    o.seed = seed_;
    o.useGraph = useGraph_ ? 1 : 0;
    o.extendBlock = extendShape_.block;
    o.extendStack = extendShape_.stack;
    o.shadowBlock = shadowShape_.block;
    o.shadowStack = shadowShape_.stack;
    o.tailPaths = tailOption_;
    o.cameraPerLane = packetCamera_ ? 0u : 1u;
}

// --------------------------------------------------------------------- commit --
void Engine::commit(mrtCommitStats* out) {
    const uint32_t light = kDirtyCamera | kDirtyEnvParams;
    const uint32_t d = scene_.dirty();
    if (d != kDirtyNone && (d & ~light) == 0) {
        scene_.clearDirty();
        resetAccumulation();
        if (out) {
            out->blasRebuilt = 0;
            out->tlasRebuilt = 0;
            out->commitMs = 0.0f;
        }
        return;
    }
    sync(out);
}

void Engine::sync(mrtCommitStats* out) {
    const auto t0 = Clock::now();
    const uint32_t d = scene_.dirty();
    if (d == kDirtyNone) {
        if (out) {
            out->blasRebuilt = 0;
            out->tlasRebuilt = 0;
            out->commitMs = 0.0f;
        }
        return;
    }
    DeviceScope scope(device_);
    MRT_HIP(hipStreamSynchronize(stream_));
    uploadBytes_ = 0;
    uint32_t rebuilt = 0;
    if (d & kDirtyTextures) uploadTextures();
    if (d & kDirtyMaterials) packMaterials();
    if (d & kDirtyGeometry) rebuilt = packGeometry();
    const bool tlas = (d & (kDirtyGeometry | kDirtyInstances | kDirtyMaterials)) != 0;
    if (tlas) packInstances();
    if (d & (kDirtyLights | kDirtyEnv)) packLights();
    if (d & kDirtyEnv) packEnvironment();
    // the learned environment-stratum statistics describe the old scene
    if (dNeeStats_.get()) dNeeStats_.zero(stream_);
    scene_.clearDirty();
    resetAccumulation();
    stats_.commitMs = msSince(t0);
    stats_.commitUploadBytes = uploadBytes_;
    stats_.geometryRepacks = geometryRepacks_;
    stats_.meshesWritten = meshesWritten_;
    if (out) {
        out->blasRebuilt = rebuilt;
        out->tlasRebuilt = tlas ? 1 : 0;
        out->commitMs = stats_.commitMs;
    }
}

// -------------------------------------------------------------------- packing --
// Reference Renderer.cpp:333-446.
uint32_t Engine::packGeometry() {
    auto& meshes = scene_.meshes();
    for (auto it = blas_.begin(); it != blas_.end();)
        it = meshes.count(it->first) ? std::next(it) : blas_.erase(it);

    std::vector<std::pair<uint32_t, MeshData*>> jobs;
    for (auto& [id, m] : meshes)
        if (m.blasDirty || !blas_.count(id)) jobs.emplace_back(id, &m);
    std::vector<Blas> built(jobs.size());
    // Large meshes build on the GPU (same nodes, kernels/bvh_gpu.hip), one
    // after the other on the engine's stream; the rest on host threads.
    std::vector<char> onGpu(jobs.size(), 0);
    for (size_t j = 0; j < jobs.size(); ++j) {
        const MeshData& m = *jobs[j].second;
        const uint32_t nt = m.triCount();
        if (nt < kGpuBuildMinTriangles) continue;
        std::vector<Box3> bounds(nt);
        for (uint32_t t = 0; t < nt; ++t)
            for (int k = 0; k < 3; ++k) bounds[t].grow(&m.positions[size_t(m.indices[t * 3 + k]) * 4]);
        BvhBuild b = buildBvhOnDevice(gpuBvh_, bounds, 4, stream_);
        built[j].wide = collapseWide(b.nodes);
        built[j].nodes = std::move(b.nodes);
        built[j].triOrder = std::move(b.primOrder);
        built[j].bounds = b.rootBounds;
        built[j].depth = b.maxDepth;
        onGpu[j] = 1;
    }
    std::atomic<size_t> next{ 0 };
    auto worker = [&] {
        for (size_t j; (j = next.fetch_add(1)) < jobs.size();) {
            if (onGpu[j]) continue;
            const MeshData& m = *jobs[j].second;
            const uint32_t nt = m.triCount();
            std::vector<Box3> bounds(nt);
            for (uint32_t t = 0; t < nt; ++t)
                for (int k = 0; k < 3; ++k) bounds[t].grow(&m.positions[size_t(m.indices[t * 3 + k]) * 4]);
            // Large meshes parallelise inside the builder; small ones across meshes.
            BvhBuild b = buildBvh(bounds, 4, nt >= 65536 ? 0 : 1);
            built[j].wide = collapseWide(b.nodes);
            built[j].nodes = std::move(b.nodes);
            built[j].triOrder = std::move(b.primOrder);
            built[j].bounds = b.rootBounds;
            built[j].depth = b.maxDepth;
        }
    };
    const unsigned nthreads = std::max(1u, std::min<unsigned>(std::thread::hardware_concurrency(), unsigned(jobs.size())));
    std::vector<std::thread> pool;
    for (unsigned i = 1; i < nthreads; ++i) pool.emplace_back(worker);
    worker();
    for (auto& t : pool) t.join();
    for (size_t j = 0; j < jobs.size(); ++j) {
        blas_[jobs[j].first] = std::move(built[j]);
        jobs[j].second->blasDirty = false;
    }

    std::vector<uint32_t> dirty;
    dirty.reserve(jobs.size());
    for (auto& j : jobs) dirty.push_back(j.first);
    maxBlasDepth_ = 0;
    totalTris_ = totalVerts_ = 0;
    for (auto& [id, m] : meshes) {
        maxBlasDepth_ = std::max(maxBlasDepth_, blas_[id].depth);
        totalTris_ += m.triCount();
        totalVerts_ += m.vertexCount();
    }
    if (totalTris_ > kMaxTraversalIndex)
        throw Error(MRT_ERROR_OUT_OF_MEMORY, "scene holds more than 2^27 - 1 triangles");
    if (!layoutValid_ || !placeIncremental(dirty)) repackGeometry();
    return uint32_t(jobs.size());
}

// Wide node w with local indices -> global: inner children by wideBase,
// leaves' first primitives by primBase, big leaves' binary nodes by nodeBase.
static void rebaseWide(WideNode& w, uint32_t wideBase, uint32_t primBase, uint32_t nodeBase) {
    for (int k = 0; k < 4; ++k) {
        const uint32_t code = w.code[k];
        if (code == kWideEmpty) continue;
        w.child[k] += code == kWideInner ? wideBase : (code == kWideBigLeaf ? nodeBase : primBase);
    }
}

// One mesh's vertices, its triangles in BLAS leaf order and its BLAS nodes
// (binary and four-wide) rebased to the slot's ranges, appended to g.
void Engine::appendMesh(const MeshData& m, Blas& b, const MeshSlot& sl, GeomArrays& g) {
    b.rootGlobal = sl.nb;
    g.pos.insert(g.pos.end(), m.positions.begin(), m.positions.end());
    g.nrm.insert(g.nrm.end(), m.normals.begin(), m.normals.end());
    if (m.uvs.empty()) g.uv.insert(g.uv.end(), size_t(m.vertexCount()) * 2, 0.0f);
    else g.uv.insert(g.uv.end(), m.uvs.begin(), m.uvs.end());
    if (m.tangents.empty()) {
        for (uint32_t i = 0; i < m.vertexCount(); ++i) g.tan.insert(g.tan.end(), { 1.0f, 0.0f, 0.0f, 1.0f });
    } else {
        g.tan.insert(g.tan.end(), m.tangents.begin(), m.tangents.end());
    }
    uint32_t slot = sl.tb;
    for (uint32_t local : b.triOrder) {
        g.tris.push_back(GpuTriangle{ m.indices[local * 3] + sl.vb, m.indices[local * 3 + 1] + sl.vb,
                                      m.indices[local * 3 + 2] + sl.vb, 0u });
        const size_t at = g.triVerts.size();
        for (int k = 0; k < 3; ++k) {
            const float* p = &m.positions[size_t(m.indices[local * 3 + k]) * 4];
            g.triVerts.insert(g.triVerts.end(), { p[0], p[1], p[2], 0.0f });
        }
        std::memcpy(&g.triVerts[at + 3], &slot, 4);
        ++slot;
    }
    for (BvhNode n : b.nodes) {
        n.leftOrPrim += (n.packed >> 2) == 0 ? sl.nb : sl.tb;
        g.nodes.push_back(n);
    }
    b.wideRootGlobal = sl.wb;
    for (WideNode w : b.wide) {
        rebaseWide(w, sl.wb, sl.tb, sl.nb);
        g.wide.push_back(w);
    }
}

// Every mesh contiguous in id order (the reference's layout): vertices,
// triangles, then the BLAS nodes with every root at an odd index (an unused
// node in front where needed: child pairs root + 1 + 2k then start on 64-byte
// boundaries, so the two children a traversal step reads share one cache
// line), then the TLAS region, reserved for half as many instances again as
// the scene holds.  Buffers keep spare room (a quarter of the used size, and
// at least the largest mesh plus an eighth) where later commits move edited
// meshes that outgrow their ranges.
void Engine::repackGeometry() {
    auto& meshes = scene_.meshes();
    GeomArrays g;
    g.pos.reserve(totalVerts_ * 4);
    g.nrm.reserve(totalVerts_ * 4);
    g.uv.reserve(totalVerts_ * 2);
    g.tan.reserve(totalVerts_ * 4);
    g.tris.reserve(totalTris_);
    g.triVerts.reserve(totalTris_ * 12);
    slots_.clear();
    holes_ = Holes{};
    for (auto& [id, m] : meshes) {
        Blas& b = blas_[id];
        if (g.nodes.size() % 2 == 0) g.nodes.push_back(sentinelNode());
        MeshSlot sl;
        sl.vb = uint32_t(g.pos.size() / 4);
        sl.vcap = m.vertexCount();
        sl.tb = uint32_t(g.tris.size());
        sl.tcap = m.triCount();
        sl.nb = uint32_t(g.nodes.size());
        sl.ncap = uint32_t(b.nodes.size());
        sl.wb = uint32_t(g.wide.size());
        sl.wcap = uint32_t(b.wide.size());
        appendMesh(m, b, sl, g);
        slots_[id] = sl;
    }
    if (g.nodes.size() % 2 == 0) g.nodes.push_back(sentinelNode());
    tlasRoot_ = uint32_t(g.nodes.size());
    const size_t inst = scene_.instances().size();
    tlasCap_ = uint32_t(std::max<size_t>(2 * (inst + inst / 2) + 1, 4));
    vertEnd_ = uint32_t(g.pos.size() / 4);
    triEnd_ = uint32_t(g.tris.size());
    nodeEnd_ = tlasRoot_ + tlasCap_;
    // a wide tree has at most as many nodes as its binary tree
    tlasWideRoot_ = uint32_t(g.wide.size());
    tlasWideCap_ = tlasCap_;
    wideEnd_ = tlasWideRoot_ + tlasWideCap_;
    if (size_t(nodeEnd_) > kMaxTraversalIndex || size_t(wideEnd_) > kMaxTraversalIndex)
        throw Error(MRT_ERROR_OUT_OF_MEMORY, "scene BVH holds more than 2^27 - 1 nodes");
    size_t bigV = 0, bigT = 0, bigN = 0, bigW = 0;
    for (auto& [id, m] : meshes) {
        bigV = std::max<size_t>(bigV, m.vertexCount());
        bigT = std::max<size_t>(bigT, m.triCount());
        bigN = std::max<size_t>(bigN, blas_[id].nodes.size());
        bigW = std::max<size_t>(bigW, blas_[id].wide.size());
    }
    auto roomOf = [](size_t n, size_t big) { return n + std::max(n / 4, big + big / 8 + 8) + 64; };
    auto put = [&](DeviceBuffer& b, const void* data, size_t bytes, size_t capBytes) {
        b.ensure(capBytes);
        b.upload(data, bytes);
        uploadBytes_ += bytes;
    };
    const size_t vRoom = roomOf(vertEnd_, bigV), tRoom = roomOf(triEnd_, bigT), nRoom = roomOf(nodeEnd_, bigN);
    put(dPos_, g.pos.data(), g.pos.size() * 4, vRoom * 16);
    put(dNrm_, g.nrm.data(), g.nrm.size() * 4, vRoom * 16);
    put(dUv_, g.uv.data(), g.uv.size() * 4, vRoom * 8);
    put(dTan_, g.tan.data(), g.tan.size() * 4, vRoom * 16);
    put(dTris_, g.tris.data(), g.tris.size() * sizeof(GpuTriangle), tRoom * sizeof(GpuTriangle));
    put(dTriVerts_, g.triVerts.data(), g.triVerts.size() * 4, tRoom * 48);
    put(dNodes_, g.nodes.data(), g.nodes.size() * sizeof(BvhNode), nRoom * sizeof(BvhNode));
    put(dWide_, g.wide.data(), g.wide.size() * sizeof(WideNode), roomOf(wideEnd_, bigW) * sizeof(WideNode));
    layoutValid_ = true;
    ++geometryRepacks_;
    invalidateGraph();
}

// The rebuilt meshes placed without moving the others: each of a mesh's
// three ranges (vertices, triangles, nodes) is reused when its new data fits
// it, else moved behind the used ranges (the old range becomes a hole, as do
// removed meshes' ranges).  Only those meshes are uploaded.  Returns false,
// changing nothing, when that needs more than the allocated capacity or
// would leave holes over half the used space.
bool Engine::placeIncremental(const std::vector<uint32_t>& dirty) {
    auto& meshes = scene_.meshes();
    std::map<uint32_t, MeshSlot> slots = slots_;
    Holes holes = holes_;
    for (auto it = slots.begin(); it != slots.end();) {
        if (meshes.count(it->first)) {
            ++it;
            continue;
        }
        holes.add(it->second);
        it = slots.erase(it);
    }
    uint64_t ve = vertEnd_, te = triEnd_, ne = nodeEnd_, we = wideEnd_;
    struct Place {
        uint32_t id;
        MeshSlot sl;
        bool pad;  // an unused node in front of a moved root (at an even index)
    };
    std::vector<Place> place;
    for (uint32_t id : dirty) {
        const MeshData& m = meshes.at(id);
        const uint32_t nv = m.vertexCount(), nt = m.triCount(), nn = uint32_t(blas_[id].nodes.size());
        const uint32_t nw = uint32_t(blas_[id].wide.size());
        auto it = slots.find(id);
        MeshSlot sl;
        const bool had = it != slots.end();
        if (had) sl = it->second;
        // a moved range keeps 1/16 spare: a re-edited mesh whose BLAS has a
        // few more nodes (or which gains a few triangles) then stays put
        auto spare = [](uint32_t n) { return n + n / 16 + 4; };
        if (!had || nv > sl.vcap) {
            if (had) holes.verts += sl.vcap;
            sl.vb = uint32_t(ve);
            sl.vcap = spare(nv);
            ve += sl.vcap;
        }
        if (!had || nt > sl.tcap) {
            if (had) holes.tris += sl.tcap;
            sl.tb = uint32_t(te);
            sl.tcap = spare(nt);
            te += sl.tcap;
        }
        bool pad = false;
        if (!had || nn > sl.ncap) {
            if (had) holes.nodes += sl.ncap;
            sl.nb = uint32_t(ne | 1u);  // roots at odd indices
            sl.ncap = spare(nn);
            pad = sl.nb != ne;
            ne = uint64_t(sl.nb) + sl.ncap;
        }
        if (!had || nw > sl.wcap) {
            if (had) holes.wides += sl.wcap;
            sl.wb = uint32_t(we);
            sl.wcap = spare(nw);
            we += sl.wcap;
        }
        slots[id] = sl;
        place.push_back(Place{ id, sl, pad });
    }
    if (ne > kMaxTraversalIndex || te > kMaxTraversalIndex || we > kMaxTraversalIndex) return false;
    if (ve * 16 > dPos_.size() || ve * 16 > dNrm_.size() || ve * 16 > dTan_.size() || ve * 8 > dUv_.size() ||
        te * sizeof(GpuTriangle) > dTris_.size() || te * 48 > dTriVerts_.size() ||
        ne * sizeof(BvhNode) > dNodes_.size() || we * sizeof(WideNode) > dWide_.size())
        return false;
    if (2 * holes.verts > ve || 2 * holes.tris > te || 2 * holes.nodes > ne || 2 * holes.wides > we) return false;
    for (const Place& pl : place) {
        const MeshSlot& sl = pl.sl;
        GeomArrays g;
        appendMesh(meshes.at(pl.id), blas_[pl.id], sl, g);
        dPos_.upload(g.pos.data(), g.pos.size() * 4, size_t(sl.vb) * 16);
        dNrm_.upload(g.nrm.data(), g.nrm.size() * 4, size_t(sl.vb) * 16);
        dUv_.upload(g.uv.data(), g.uv.size() * 4, size_t(sl.vb) * 8);
        dTan_.upload(g.tan.data(), g.tan.size() * 4, size_t(sl.vb) * 16);
        dTris_.upload(g.tris.data(), g.tris.size() * sizeof(GpuTriangle), size_t(sl.tb) * sizeof(GpuTriangle));
        dTriVerts_.upload(g.triVerts.data(), g.triVerts.size() * 4, size_t(sl.tb) * 48);
        if (pl.pad) {
            const BvhNode pad = sentinelNode();
            dNodes_.upload(&pad, sizeof(BvhNode), size_t(sl.nb - 1) * sizeof(BvhNode));
        }
        dNodes_.upload(g.nodes.data(), g.nodes.size() * sizeof(BvhNode), size_t(sl.nb) * sizeof(BvhNode));
        dWide_.upload(g.wide.data(), g.wide.size() * sizeof(WideNode), size_t(sl.wb) * sizeof(WideNode));
        uploadBytes_ += (g.pos.size() + g.nrm.size() + g.uv.size() + g.tan.size() + g.triVerts.size()) * 4 +
                        g.tris.size() * sizeof(GpuTriangle) + g.nodes.size() * sizeof(BvhNode) +
                        g.wide.size() * sizeof(WideNode);
        ++meshesWritten_;
    }
    slots_ = std::move(slots);
    holes_ = holes;
    vertEnd_ = uint32_t(ve);
    triEnd_ = uint32_t(te);
    nodeEnd_ = uint32_t(ne);
    wideEnd_ = uint32_t(we);
    return true;
}

// Reference Renderer.cpp:448-508: the TLAS goes into the region reserved
// behind the BLAS nodes; instances outgrowing it force a full repack, which
// reserves room for the new count.
void Engine::packInstances() {
    std::vector<const InstanceData*> flat;
    std::vector<Box3> bounds;
    for (auto& [id, inst] : scene_.instances()) {
        auto it = blas_.find(inst.mesh);
        if (it == blas_.end() || !it->second.bounds.valid()) continue;
        flat.push_back(&inst);
        bounds.push_back(transformBox(it->second.bounds, inst.objectToWorld));
    }
    BvhBuild tlas = buildBvh(bounds, 1);
    if (tlas.nodes.empty()) tlas.nodes.push_back(sentinelNode());
    std::vector<WideNode> tlasWide = collapseWide(tlas.nodes);
    if (!layoutValid_ || tlas.nodes.size() > tlasCap_ || tlasWide.size() > tlasWideCap_) repackGeometry();
    std::vector<GpuInstance> gi;
    std::vector<TraceInstance> ti;
    gi.reserve(flat.size());
    ti.reserve(flat.size());
    uint32_t maxLeaf = 1;
    for (uint32_t idx : tlas.primOrder) {
        const InstanceData* inst = flat[idx];
        GpuInstance g{};
        float inv[16];
        invert4(inst->objectToWorld, inv);
        for (int c = 0; c < 4; ++c) {
            g.objectToWorld[c] = make_float4(inst->objectToWorld[c * 4], inst->objectToWorld[c * 4 + 1],
                                             inst->objectToWorld[c * 4 + 2], inst->objectToWorld[c * 4 + 3]);
            g.worldToObject[c] = make_float4(inv[c * 4], inv[c * 4 + 1], inv[c * 4 + 2], inv[c * 4 + 3]);
        }
        g.blasRoot = blas_[inst->mesh].rootGlobal;
        auto mi = materialIndex_.find(inst->material);
        g.materialIndex = mi != materialIndex_.end() ? mi->second : 0u;
        gi.push_back(g);
        TraceInstance t{};
        t.row0 = make_float4(inv[0], inv[4], inv[8], inv[12]);
        t.row1 = make_float4(inv[1], inv[5], inv[9], inv[13]);
        t.row2 = make_float4(inv[2], inv[6], inv[10], inv[14]);
        t.blasRoot = g.blasRoot;
        t.materialIndex = g.materialIndex;
        t.wideRoot = blas_[inst->mesh].wideRootGlobal;
        ti.push_back(t);
    }
    for (BvhNode& n : tlas.nodes) {
        const uint32_t cnt = n.packed >> 2;
        if (cnt == 0) n.leftOrPrim += tlasRoot_;
        else maxLeaf = std::max(maxLeaf, cnt);
    }
    instanceCount_ = uint32_t(gi.size());
    tlasNodeCount_ = uint32_t(tlas.nodes.size());
    if (gi.size() > kMaxTraversalIndex)
        throw Error(MRT_ERROR_OUT_OF_MEMORY, "scene holds more than 2^27 - 1 instances");
    // Guiding positions: the scene box scaled uniformly into [0, 1].
    if (tlas.rootBounds.valid()) {
        float extent = 0.0f;
        for (int a = 0; a < 3; ++a) {
            sceneBoxMin_[a] = tlas.rootBounds.lo[a];
            extent = std::max(extent, tlas.rootBounds.hi[a] - tlas.rootBounds.lo[a]);
        }
        sceneBoxScale_ = 1.0f / std::max(extent, 1e-6f);
    }
    tlasDepth_ = tlas.maxDepth;
    if (tlasDepth_ + maxLeaf + maxBlasDepth_ + 2 > kMaxTraversalStack)
        throw Error(MRT_ERROR_UNKNOWN, "scene BVH too deep for the traversal stack (TLAS depth " +
                                           std::to_string(tlasDepth_) + ", BLAS depth " +
                                           std::to_string(maxBlasDepth_) + ")");
    dNodes_.upload(tlas.nodes.data(), tlas.nodes.size() * sizeof(BvhNode), size_t(tlasRoot_) * sizeof(BvhNode));
    // leaves hold instance ranges (TLAS leaf order), big leaves the binary TLAS node
    for (WideNode& w : tlasWide) rebaseWide(w, tlasWideRoot_, 0u, tlasRoot_);
    dWide_.upload(tlasWide.data(), tlasWide.size() * sizeof(WideNode), size_t(tlasWideRoot_) * sizeof(WideNode));
    const void* inst0 = dInst_.get();
    const void* trace0 = dTraceInst_.get();
    uploadVec(dInst_, gi.data(), gi.size() * sizeof(GpuInstance));
    uploadVec(dTraceInst_, ti.data(), ti.size() * sizeof(TraceInstance));
    uploadBytes_ += tlas.nodes.size() * sizeof(BvhNode) + tlasWide.size() * sizeof(WideNode) +
                    gi.size() * (sizeof(GpuInstance) + sizeof(TraceInstance));
    if (dInst_.get() != inst0 || dTraceInst_.get() != trace0) invalidateGraph();
}

// Reference Renderer.cpp:510-550.
void Engine::packMaterials() {
    auto slotOf = [&](uint32_t id) -> int32_t {
        const uint32_t s = scene_.textureSlot(id);
        return (s == ~0u || s >= kMaxTextureSlots) ? -1 : int32_t(s);
    };
    std::vector<GpuMaterial> mats;
    mats.reserve(scene_.materials().size() + 1);
    materialIndex_.clear();
    GpuMaterial def{};
    def.baseColorOpacity = make_float4(0.8f, 0.8f, 0.8f, 1.0f);
    def.emissionRoughness = make_float4(0.0f, 0.0f, 0.0f, 0.5f);
    def.params = make_float4(0.0f, 0.0f, 1.45f, 0.0f);
    def.baseColorTex = def.roughnessTex = def.metallicTex = def.normalTex = def.emissionTex = -1;
    mats.push_back(def);
    for (auto& [id, m] : scene_.materials()) {
        GpuMaterial g{};
        g.baseColorOpacity = make_float4(m.baseColor[0], m.baseColor[1], m.baseColor[2], m.opacity);
        g.emissionRoughness = make_float4(m.emission[0], m.emission[1], m.emission[2], std::clamp(m.roughness, 0.0f, 1.0f));
        g.params = make_float4(std::clamp(m.metallic, 0.0f, 1.0f), std::clamp(m.transmission, 0.0f, 1.0f),
                               std::max(m.ior, 1.0f), 0.0f);
        g.baseColorTex = slotOf(m.baseColorTex);
        g.roughnessTex = slotOf(m.roughnessTex);
        g.metallicTex = slotOf(m.metallicTex);
        g.normalTex = slotOf(m.normalTex);
        g.emissionTex = slotOf(m.emissionTex);
        materialIndex_[id] = uint32_t(mats.size());
        mats.push_back(g);
    }
    uploadVec(dMats_, mats.data(), mats.size() * sizeof(GpuMaterial));
    invalidateGraph();
}

// Reference Renderer.cpp:552-603 (environment at slot 0 when present).
void Engine::packLights() {
    std::vector<GpuLight> lights;
    if (scene_.hasEnvironment()) {
        GpuLight g{};
        g.type = kLightEnv;
        lights.push_back(g);
    }
    for (auto& [id, l] : scene_.lights()) {
        GpuLight g{};
        if (l.type == MRT_LIGHT_SUN) {
            g.type = kLightSun;
            const float d[3] = { l.direction[0], l.direction[1], l.direction[2] };
            const float len = std::sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
            const float inv = 1.0f / len;
            g.p0 = make_float4(d[0] * inv, d[1] * inv, d[2] * inv, std::cos(std::max(l.angularRadius, 1e-4f)));
            g.p1 = make_float4(l.radiance[0], l.radiance[1], l.radiance[2], 0.0f);
        } else if (l.type == MRT_LIGHT_POINT) {
            g.type = kLightPoint;
            g.p0 = make_float4(l.position[0], l.position[1], l.position[2], 0.0f);
            g.p1 = make_float4(l.radiance[0], l.radiance[1], l.radiance[2], 0.0f);
        } else if (l.type == MRT_LIGHT_RECT) {
            g.type = kLightRect;
            const float* e0 = l.edge0;
            const float* e1 = l.edge1;
            const float cx = e0[1] * e1[2] - e1[1] * e0[2], cy = e0[2] * e1[0] - e1[2] * e0[0],
                        cz = e0[0] * e1[1] - e1[0] * e0[1];
            const float area = std::sqrt((cx * cx + cy * cy) + cz * cz);
            g.p0 = make_float4(l.corner[0], l.corner[1], l.corner[2], area);
            g.p1 = make_float4(e0[0], e0[1], e0[2], 0.0f);
            g.p2 = make_float4(e1[0], e1[1], e1[2], 0.0f);
            g.p3 = make_float4(l.radiance[0], l.radiance[1], l.radiance[2], l.twoSided ? 1.0f : 0.0f);
        } else {
            continue;  // unknown light type: ignored
        }
        lights.push_back(g);
    }
    lightCount_ = uint32_t(lights.size());
    uploadVec(dLights_, lights.data(), lights.size() * sizeof(GpuLight));
    invalidateGraph();
}

// Reference Renderer.cpp:605-662: luminance x sin(theta) texel weights and
// the solid-angle integral (FP64 sums); an alias table replaces the CDFs.
void Engine::packEnvironment() {
    hasEnv_ = scene_.hasEnvironment();
    if (!hasEnv_) {
        envW_ = envH_ = 0;
        envSplitA_ = 0.0f;
        envCountA_ = envCountB_ = 0;
        invalidateGraph();
        return;
    }
    const TextureData& env = scene_.environment();
    const uint32_t w = env.width, h = env.height;
    const float* px = reinterpret_cast<const float*>(env.data.data());
    const float kPi = 3.14159265358979323846f;
    const float dTheta = kPi / float(h);
    const float dPhi = 2.0f * kPi / float(w);
    std::vector<double> weights(size_t(w) * h);
    double integral = 0.0;
    for (uint32_t y = 0; y < h; ++y) {
        const float sinT = std::sin((float(y) + 0.5f) * dTheta);
        double row = 0.0;
        for (uint32_t x = 0; x < w; ++x) {
            const float* p = px + (size_t(y) * w + x) * 4;
            const float lum = (0.2126f * p[0] + 0.7152f * p[1]) + 0.0722f * p[2];
            const double wgt = double(lum) * double(sinT);
            weights[size_t(y) * w + x] = wgt;
            row += wgt;
        }
        integral += row * double(dTheta) * double(dPhi);
    }
    std::vector<AliasEntry> alias;
    buildAlias(weights, alias);
    envW_ = w;
    envH_ = h;
    envIntegral_ = std::max(float(integral), 1e-8f);
    uploadVec(dEnv_, px, size_t(w) * h * 16);
    uploadVec(dAlias_, alias.data(), alias.size() * sizeof(AliasEntry));

    // Strata for guided shading's environment NEE: texels of at least 64 x
    // the mean weight (a sun) and the rest, each with its own alias table over
    // the same weights, so that sampling a stratum with probability s and
    // then its table gives texel t the probability P(t) s / share (bright) or
    // P(t) (1 - s) / (1 - share).
    const size_t n = weights.size();
    double total = 0.0;
    for (double x : weights) total += x;
    const double cut = 64.0 * total / double(std::max<size_t>(n, 1));
    std::vector<double> wa, wb;
    std::vector<uint32_t> ta, tb;
    double sumA = 0.0;
    for (size_t i = 0; i < n; ++i) {
        if (weights[i] >= cut && weights[i] > 0.0) {
            wa.push_back(weights[i]);
            ta.push_back(uint32_t(i));
            sumA += weights[i];
        } else {
            wb.push_back(weights[i]);
            tb.push_back(uint32_t(i));
        }
    }
    const bool split = !wa.empty() && !wb.empty() && total > 0.0 && sumA < total;
    envSplitA_ = split ? float(sumA / total) : 0.0f;
    envCountA_ = split ? uint32_t(wa.size()) : 0u;
    envCountB_ = split ? uint32_t(wb.size()) : 0u;
    auto table = [](const std::vector<double>& wts, const std::vector<uint32_t>& tex) {
        std::vector<AliasEntry> al;
        buildAlias(wts, al);
        std::vector<EnvStratumEntry> out(al.size());
        for (size_t i = 0; i < al.size(); ++i) out[i] = EnvStratumEntry{ al[i].prob, al[i].alias, tex[i], 0u };
        return out;
    };
    if (split) {
        const auto a = table(wa, ta), b = table(wb, tb);
        uploadVec(dEnvA_, a.data(), a.size() * sizeof(EnvStratumEntry));
        uploadVec(dEnvB_, b.data(), b.size() * sizeof(EnvStratumEntry));
    }
    invalidateGraph();
}

// Texture pools: RGBA8 texels as uint32 (bytes R,G,B,A), RGBA32F as float4.
void Engine::uploadTextures() {
    const auto& tex = scene_.textures();
    std::vector<GpuTexture> table(kMaxTextureSlots, GpuTexture{ 0, 0, 0, 0 });
    std::vector<uint32_t> pool8;
    std::vector<float> pool32;
    for (size_t slot = 0; slot < tex.size() && slot < kMaxTextureSlots; ++slot) {
        const TextureData& t = tex[slot];
        if (t.data.empty()) continue;
        GpuTexture& g = table[slot];
        g.width = t.width;
        g.height = t.height;
        g.format = t.format;
        const size_t n = size_t(t.width) * t.height;
        if (t.format == MRT_TEX_RGBA32F) {
            g.offset = uint32_t(pool32.size() / 4);
            const float* f = reinterpret_cast<const float*>(t.data.data());
            pool32.insert(pool32.end(), f, f + n * 4);
        } else {
            g.offset = uint32_t(pool8.size());
            const size_t base = pool8.size();
            pool8.resize(base + n);
            std::memcpy(&pool8[base], t.data.data(), n * 4);
        }
    }
    if (tex.size() > kMaxTextureSlots)
        logMessage(LogLevel::Warn, "texture count exceeds 64 slots; extra textures ignored");
    tex8Texels_ = pool8.size();
    tex32Texels_ = pool32.size() / 4;
    uploadVec(dTexTable_, table.data(), table.size() * sizeof(GpuTexture));
    uploadVec(dTex8_, pool8.data(), pool8.size() * 4);
    uploadVec(dTex32_, pool32.data(), pool32.size() * 4);
    invalidateGraph();
}

// ------------------------------------------------------------- frame buffers --
void Engine::allocFrame() {
    const size_t n = size_t(settings_.width) * settings_.height;
    frameW_ = settings_.width;
    frameH_ = settings_.height;
    const size_t q = queueSlots();  // ray and shadow queues: kQueueSegs segments
    dConsts_.ensure(sizeof(FrameConsts));
    dBeta_.allocate(n * 16);
    dRad_.allocate(n * 16);
    dRng_.allocate(n * 4);
    for (int i = 0; i < 2; ++i) {
        dRayO_[i].allocate(q * 16);
        dRayD_[i].allocate(q * 16);
    }
    dHit_.allocate(q * 16);
    dHitInst_.allocate(q * 4);
    dShO_.allocate(q * 16);
    dShD_.allocate(q * 16);
    dShC_.allocate(q * 16);
    dCounters_.ensure(dev::kCounterWords * 4);
    dCursors_.ensure(dev::kCursorWords * 4);
    dEntry_.ensure(sizeof(dev::BatchEntry));
    hEntry_.ensure(sizeof(dev::BatchEntry));
    dAccum_.allocate(n * 16);
    dOut8_.allocate(n * 4);
    dOut32_.allocate(n * 16);
    dAccum_.zero(stream_);
    dOut8_.zero(stream_);
    dCounters_.zero(stream_);
    MRT_HIP(hipStreamSynchronize(stream_));
    hConsts_.ensure(sizeof(FrameConsts));
    hCounters_.ensure(dev::kCounterWords * 4);
    hLevelRays_.ensure(dev::kMaxBounceLevels * 4);
    levelRaysFresh_ = false;
    if (guideActive()) allocGuide();
    invalidateGraph();
}

// Per-frame guiding buffers, sized for the frame and the bounce levels.
void Engine::allocGuide() {
    const uint32_t n = frameW_ * frameH_;
    const uint32_t levels = guideLevels();
    dGuideIdx_.allocate(queueSlots() * 4);  // per ray-queue slot
    dGuideRec_.allocate(size_t(n) * 3 * 16);
    dGuideMix_.allocate(size_t(n) * dev::kGuideMixFloat4 * 16);
    // Training paths: every path of the frame keeps its per-level records
    // (64 B each) while the path-record buffer stays within kPathRecordBytes;
    // larger frames record every stride-th path.  A path yields well under
    // one positive-weight record on average (most paths end early or carry
    // no radiance), so sparser recording would leave the Adam batches short of
    // batchSize.  The collection pass samples the pool down to
    // batchSize x (the frame's Adam steps) records; the pool holds the
    // warm-up's or the steady count, whichever is larger.
    const uint64_t poolCap = uint64_t(guide_.batchSize) * std::max(guide_.stepsPerFrame, guide_.warmupStepsPerFrame);
    const uint64_t recBytes = uint64_t(n) * levels * 64;
    guideTrainStride_ = uint32_t(std::max<uint64_t>(1, (recBytes + kPathRecordBytes - 1) / kPathRecordBytes));
    const uint32_t trainPaths = (n + guideTrainStride_ - 1) / guideTrainStride_;
    dPathRec_.allocate(size_t(levels) * trainPaths * 4 * 16);
    dPathRec_.zero(stream_);  // stamp 0 never matches a frame
    dRecMask_.allocate(size_t(trainPaths) * 4);  // reset per frame by ray generation
    dGuidePool_.allocate(size_t(poolCap) * 4 * 16);
    dRrScale_.allocate(size_t(n) * 4);
    dShX_.allocate(queueSlots() * 8);  // per shadow-queue slot
    dNeeStats_.allocate(size_t(kNeeCells) * 16);
    dNeeStats_.zero(stream_);
    MRT_HIP(hipStreamSynchronize(stream_));
    guidePaths_ = n;
    guideLevels_ = levels;
    invalidateGraph();
}

void Engine::ensureGuideNet() {
    if (!guideNet_) {
        guideNet_ = std::make_unique<GuideNetwork>(guide_.batchSize);
        guideNet_->initParams(guide_.seed);
    } else {
        guideNet_->ensureCap(guide_.batchSize);
    }
}

dev::SceneView Engine::sceneView() const {
    dev::SceneView s;
    s.nodes = dNodes_.as<const dev::Node>();
    s.wide = dWide_.as<const float4>();
    s.triVerts = dTriVerts_.as<const float4>();
    s.tris = dTris_.as<const GpuTriangle>();
    s.positions = dPos_.as<const float4>();
    s.normals = dNrm_.as<const float4>();
    s.uvs = dUv_.as<const float2>();
    s.tangents = dTan_.as<const float4>();
    s.instances = dInst_.as<const GpuInstance>();
    s.traceInst = dTraceInst_.as<const TraceInstance>();
    s.materials = dMats_.as<const GpuMaterial>();
    s.lights = dLights_.as<const GpuLight>();
    s.envMap = dEnv_.as<const float4>();
    s.envAlias = dAlias_.as<const AliasEntry>();
    s.envA = dEnvA_.as<const EnvStratumEntry>();
    s.envB = dEnvB_.as<const EnvStratumEntry>();
    s.texTable = dTexTable_.as<const GpuTexture>();
    s.tex8 = dTex8_.as<const uint32_t>();
    s.tex32 = dTex32_.as<const float4>();
    return s;
}

dev::PathView Engine::pathView() const {
    dev::PathView p;
    p.capacity = frameW_ * frameH_;
    p.segCap = dev::queueSegCap(p.capacity);
    p.beta = dBeta_.as<float4>();
    p.radiance = dRad_.as<float4>();
    p.rngCounter = dRng_.as<uint32_t>();
    for (int i = 0; i < 2; ++i) {
        p.rayO[i] = dRayO_[i].as<float4>();
        p.rayD[i] = dRayD_[i].as<float4>();
    }
    p.hit = dHit_.as<float4>();
    p.hitInst = dHitInst_.as<uint32_t>();
    p.shadowO = dShO_.as<float4>();
    p.shadowD = dShD_.as<float4>();
    p.shadowC = dShC_.as<float4>();
    p.counters = dCounters_.as<uint32_t>();
    p.accum = dAccum_.as<float4>();
    p.outRgba8 = dOut8_.as<uint32_t>();
    p.outRgba32 = dOut32_.as<float4>();
    const bool guide = guideActive() && guidePaths_ == p.capacity && guideLevels_ == guideLevels() && guideNet_;
    p.guideIdx = guide ? dGuideIdx_.as<uint32_t>() : nullptr;
    p.guideRec = guide ? dGuideRec_.as<float4>() : nullptr;
    p.guideMix = guide ? dGuideMix_.as<float4>() : nullptr;
    p.pathRec = guide && guide_.enableTraining ? dPathRec_.as<float4>() : nullptr;
    p.recMask = p.pathRec ? dRecMask_.as<uint32_t>() : nullptr;
    const bool nee = guide && guide_.enableSampling && envSplitA_ > 0.0f;
    p.shadowX = nee ? dShX_.as<uint2>() : nullptr;
    p.neeStats = nee ? dNeeStats_.as<float4>() : nullptr;
    p.rrScale = guide && guide_.enableSampling ? dRrScale_.as<float>() : nullptr;
    p.levelRays = static_cast<uint32_t*>(hLevelRays_.get());
    return p;
}

// Reference Renderer.cpp:711-764.
void Engine::fillConsts(FrameConsts& g) const {
    std::memset(&g, 0, sizeof(g));
    float pos[3], fwd[3], up[3];
    float left, right, bottom, top;
    bool ortho = false;
    auto normalize3 = [](float v[3]) {
        const float inv = 1.0f / std::sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
        v[0] *= inv; v[1] *= inv; v[2] *= inv;
    };
    if (scene_.usesCameraEx()) {
        const CameraExData& c = scene_.cameraEx();
        for (int i = 0; i < 3; ++i) { pos[i] = c.position[i]; fwd[i] = c.forward[i]; up[i] = c.up[i]; }
        normalize3(fwd);
        left = c.left; right = c.right; bottom = c.bottom; top = c.top;
        ortho = c.projection == MRT_CAMERA_ORTHOGRAPHIC;
    } else {
        const CameraData& c = scene_.camera();
        for (int i = 0; i < 3; ++i) { pos[i] = c.position[i]; fwd[i] = c.target[i] - c.position[i]; up[i] = c.up[i]; }
        normalize3(fwd);
        const float aspect = float(settings_.width) / float(settings_.height);
        const float tanHalf = std::tan(c.fovYDeg * 0.0174532925199432957692f * 0.5f);
        right = tanHalf * aspect;
        left = -right;
        top = tanHalf;
        bottom = -top;
    }
    float rv[3] = { fwd[1] * up[2] - up[1] * fwd[2], fwd[2] * up[0] - up[2] * fwd[0], fwd[0] * up[1] - up[0] * fwd[1] };
    const float rlen = std::sqrt((rv[0] * rv[0] + rv[1] * rv[1]) + rv[2] * rv[2]);
    if (rlen < 1e-6f) { rv[0] = 1.0f; rv[1] = 0.0f; rv[2] = 0.0f; }
    else normalize3(rv);
    const float uv[3] = { rv[1] * fwd[2] - fwd[1] * rv[2], rv[2] * fwd[0] - fwd[2] * rv[0], rv[0] * fwd[1] - fwd[0] * rv[1] };
    g.camPos = make_float4(pos[0], pos[1], pos[2], 0.0f);
    g.camRight = make_float4(rv[0], rv[1], rv[2], 0.0f);
    g.camUp = make_float4(uv[0], uv[1], uv[2], 0.0f);
    g.camForward = make_float4(fwd[0], fwd[1], fwd[2], 0.0f);
    g.camFrustum = make_float4(left, right, bottom, top);
    g.width = settings_.width;
    g.height = settings_.height;
    g.frameIndex = frameIndex_;
    g.maxBounces = std::min(settings_.maxBounces, 16u);
    g.lightCount = lightCount_;
    g.envWidth = hasEnv_ ? envW_ : 0;
    g.envHeight = hasEnv_ ? envH_ : 0;
    g.flags = (hasEnv_ ? kFlagHasEnv : 0u) | (ortho ? kFlagOrthographic : 0u) |
              (packetCamera_ ? kFlagPacketCamera : 0u);
    g.envRotation = scene_.envRotation();
    g.envIntensity = scene_.envIntensity();
    g.envIntegral = envIntegral_;
    g.exposure = std::exp2(settings_.exposureEV);
    g.fireflyClamp = settings_.fireflyClamp;
    g.tonemap = settings_.tonemap;
    g.tlasRoot = tlasRoot_;
    g.tlasWideRoot = tlasWideRoot_;
    g.envSplitA = hasEnv_ ? envSplitA_ : 0.0f;
    g.envCountA = hasEnv_ ? envCountA_ : 0u;
    g.envCountB = hasEnv_ ? envCountB_ : 0u;
    g.instanceCount = instanceCount_;
    g.seedMix = seed_ ? (seed_ * 0x9E3779B9u) ^ 0x85EBCA6Bu : 0u;
    g.guiding = guide_.enableSampling ? 1u : 0u;
    g.guideSelectProb = guide_.selectProb;
    g.guideTrain = guide_.enableTraining ? 1u : 0u;
    g.guideMinRoughness = guide_.minRoughness;
    g.guideStamp = guideStamp_;
    g.guideTrainStride = guideTrainStride_;
    g.guideTrainPhase = guideStamp_ % guideTrainStride_;
    g.guideBoxMin = make_float4(sceneBoxMin_[0], sceneBoxMin_[1], sceneBoxMin_[2], 0.0f);
    const float ps = sceneBoxScale_ * guide_.positionScale;
    g.guideBoxScale = make_float4(ps, ps, ps, 0.0f);
}

// Depth of a captured plain frame.  Once a level's queue holds at most
// tailPaths_ rays, the tail kernel finishes every path of it, so the later
// per-level launches find empty queues (about 7 us each, 16 of them at 1080p
// with depth 8).  The graph therefore ends with a tail launch that takes
// every path left at the first level the last frame held at most tailPaths_
// rays in.  The tail shades exactly as the per-level kernels do, so the image
// does not depend on the level chosen; a level that holds more rays than the
// threshold in a later frame only makes the tail slower, and the graph is
// captured again (tailLevelFits).  Guided frames keep every launch: their
// tail vertices are shaded unguided.
uint32_t Engine::fittedTailLevel() const {
    if (!levelRaysFresh_ || tailPaths_ == 0 || guideFrame()) return ~0u;
    const uint32_t* n = static_cast<const uint32_t*>(hLevelRays_.get());
    const uint32_t maxB = std::min(settings_.maxBounces, 16u);
    for (uint32_t b = 0; b <= maxB; ++b)
        if (n[b] <= tailPaths_) return b;
    return ~0u;
}

// Whether a graph ending at `level` still suits the last frame: its last
// level within 1/8 above the threshold, the level before it more than 1/8
// below (the margins keep a queue size near the threshold from capturing
// the graph again every frame).
bool Engine::tailLevelFits(uint32_t level) const {
    if (!levelRaysFresh_) return true;
    if (level == ~0u) return fittedTailLevel() == ~0u;
    if (tailPaths_ == 0 || guideFrame()) return false;
    const uint32_t* n = static_cast<const uint32_t*>(hLevelRays_.get());
    const uint64_t tp = tailPaths_;
    return n[level] <= tp + tp / 8 && (level == 0 || n[level - 1] > tp - tp / 8);
}

void Engine::enqueueFrame(hipStream_t s, bool timed, const FrameConsts* g, const dev::BatchEntry* entry,
                          uint32_t offset, uint32_t tailLevel) {
    const uint32_t pixels = frameW_ * frameH_;
    const uint32_t maxB = std::min(settings_.maxBounces, 16u);
    const bool guide = guideFrame();
    dev::BatchView bv;
    bv.entries = entry;
    bv.count = 1;
    bv.pixelSlots = (pixels + 63u) & ~63u;
    bv.cursors = dCursors_.as<uint32_t>();
    auto elapsed = [&](hipEvent_t a, hipEvent_t b) {
        float ms = 0.0f;
        MRT_HIP(hipEventElapsedTime(&ms, a, b));
        return ms;
    };
    if (timed) MRT_HIP(hipEventRecord(ev_[0], s));
    MRT_HIP(dev::launchClearBatch(bv, s));
    enqueueGuideFrameBegin(s, g);
    MRT_HIP(dev::launchGenerateBatch(bv, s));
    if (timed) MRT_HIP(hipEventRecord(ev_[1], s));
    // Per bounce b: one trace launch (closest hits of level b with the shadow
    // rays of level b - 1), the network inference, shade and tail; the
    // deepest level's shadow rays get a trace launch of their own.
    for (uint32_t b = 0; b <= maxB; ++b) {
        // At tailLevel the tail kernel takes every path (threshold ~0u; the
        // requests of a plain frame have no limit of their own).
        const bool last = b == tailLevel;
        const uint32_t tp = last ? ~0u : tailPaths_;
        if (timed) MRT_HIP(hipEventRecord(ev_[2], s));
        MRT_HIP(dev::launchTraceBatch(bv, b, b == 0 ? dev::kNoLevel : b - 1, traceShape(guide), tp, guide, s));
        if (timed) MRT_HIP(hipEventRecord(ev_[3], s));
        enqueueGuideInfer(s, b);
        if (timed) MRT_HIP(hipEventRecord(ev_[9], s));
        if (!last) MRT_HIP(dev::launchShadeBatch(bv, b, tp, guide, shadeGrid(guide), s));
        if (timed) MRT_HIP(hipEventRecord(ev_[4], s));
        MRT_HIP(dev::launchTailBatch(bv, b, tp, tailShape(), s));
        if (timed) {
            MRT_HIP(hipEventRecord(ev_[8], s));
            MRT_HIP(hipEventSynchronize(ev_[8]));
            stats_.stageMs[1] += elapsed(ev_[2], ev_[3]);
            stats_.stageMs[6] += elapsed(ev_[3], ev_[9]);
            stats_.stageMs[2] += elapsed(ev_[9], ev_[4]);
            stats_.stageMs[5] += elapsed(ev_[4], ev_[8]);
        }
        if (last) break;
    }
    if (timed) MRT_HIP(hipEventRecord(ev_[4], s));
    if (tailLevel > maxB) MRT_HIP(dev::launchTraceBatch(bv, dev::kNoLevel, maxB, shadowShape_, tailPaths_, false, s));
    if (timed) {
        MRT_HIP(hipEventRecord(ev_[5], s));
        MRT_HIP(hipEventSynchronize(ev_[5]));
        stats_.stageMs[3] += elapsed(ev_[4], ev_[5]);
    }
    if (timed) MRT_HIP(hipEventRecord(ev_[10], s));
    enqueueGuideTraining(s, g, offset);
    if (timed) MRT_HIP(hipEventRecord(ev_[6], s));
    MRT_HIP(dev::launchFinishBatch(bv, s));  // accumulate + resolve
    if (timed) {
        MRT_HIP(hipEventRecord(ev_[7], s));
        MRT_HIP(hipEventSynchronize(ev_[7]));
        stats_.stageMs[0] = elapsed(ev_[0], ev_[1]);
        stats_.stageMs[7] = elapsed(ev_[10], ev_[6]);
        stats_.stageMs[4] = elapsed(ev_[6], ev_[7]);
    }
}

// Guiding passes of this engine's frame (no-ops for unguided frames), also
// enqueued per request by the merged batch: the per-frame stamp of the
// training records, the network inference of level `bounce`'s query
// records, and the training after the frame.
bool Engine::guideFrame() const { return pathView().guideIdx != nullptr; }

void Engine::enqueueGuideFrameBegin(hipStream_t s, const FrameConsts* g) {
    if (pathView().pathRec) MRT_HIP(dev::launchGuideFrameBegin(guideNet_->deviceWords(), g, s));
}

void Engine::enqueueGuideInfer(hipStream_t s, uint32_t bounce) {
    const dev::PathView pv = pathView();
    if (!pv.guideIdx || !guide_.enableSampling) return;
    dev::RecView rv;
    rv.base = pv.guideRec;
    rv.stride = 3;
    MRT_HIP(dev::launchGuideInfer(rv, pv.counters + dev::cntWord(bounce, dev::kCntGuideRecords, 0),
                                  frameW_ * frameH_, guideNet_->deviceImage(), pv.guideMix, true, s));
}

// Vertices the tail kernel shades are not guided, so a guided request hands
// a level over only once it holds at most a sixteenth of its pixels (besides
// the batch-wide tailPaths bound).  Measured on the 1280x720 room after 256
// training frames (per-sample variance guided / plain, ms per frame): 1/4
// 0.383 at 6.99 ms, 1/16 0.220 at 7.30, 1/64 0.219 at 7.38, never 0.217 at
// 7.38.  The deep levels pay since roulette reads the BSDF-only throughput
// (PathView::rrScale).  With the training warm-up (equal-RMSE time against
// plain): 1/8 0.304, 1/16 0.204, 1/32 0.206.
uint32_t Engine::tailAt() const { return guideFrame() ? frameW_ * frameH_ / 16 : ~0u; }

dev::TraceShape Engine::traceShape(bool guidePrep) const {
    dev::TraceShape t = extendShape_;
    if (guidePrep) t.grid = prepTraceGrid_;  // the query-record variant holds more registers
    return t;
}

uint32_t Engine::trainSteps(uint32_t offset) const {
    if (!pathView().pathRec) return 0;
    return uint64_t(trainedFrames_) + offset < guide_.warmupFrames ? guide_.warmupStepsPerFrame : guide_.stepsPerFrame;
}

// Training records of this frame -> pool -> trainSteps Adam steps, batch t
// taking pool rows t, t + steps, t + 2 steps, ...
void Engine::enqueueGuideTraining(hipStream_t s, const FrameConsts* g, uint32_t offset) {
    const uint32_t steps = trainSteps(offset);
    if (steps == 0) return;
    const uint32_t poolCap = guide_.batchSize * steps;
    lastPoolCap_ = poolCap;
    dev::GuideCollectArgs ca;
    ca.pathRec = dPathRec_.as<const float4>();
    ca.recMask = dRecMask_.as<const uint32_t>();
    ca.radiance = dRad_.as<const float4>();
    ca.paths = guidePaths_;
    ca.trainPaths = (guidePaths_ + guideTrainStride_ - 1) / guideTrainStride_;
    ca.levels = guideLevels_;
    ca.pool = dGuidePool_.as<float4>();
    ca.poolCap = poolCap;
    ca.words = guideNet_->deviceWords();
    ca.g = g;
    MRT_HIP(dev::launchGuideCollect(ca, s));
    const dev::GuideNet net = guideNet_->view();
    dev::GuideAdam h;
    h.lr = guide_.learningRate;
    for (uint32_t t = 0; t < steps; ++t) {
        dev::GuideBatch b;
        b.pool = dGuidePool_.as<const float4>();
        b.total = guideNet_->deviceWords() + dev::kGwPool;
        b.poolCap = poolCap;
        b.phase = t;
        b.period = steps;
        b.cap = guide_.batchSize;
        MRT_HIP(dev::launchGuideStep(b, net, h, s));
    }
}

void Engine::renderFrame(mrtFrameInfo* out) {
    const auto t0 = Clock::now();
    commit(nullptr);
    DeviceScope scope(device_);
    if (frameIndex_ < settings_.sppLimit) {
        FrameConsts* g = static_cast<FrameConsts*>(hConsts_.get());
        if (++guideStamp_ == 0) guideStamp_ = 1;
        fillConsts(*g);
        MRT_HIP(hipMemcpyAsync(dConsts_.get(), g, sizeof(FrameConsts), hipMemcpyHostToDevice, stream_));
        dev::BatchEntry* he = static_cast<dev::BatchEntry*>(hEntry_.get());
        fillEntries(he, dConsts_.as<const FrameConsts>(), 1);
        MRT_HIP(hipMemcpyAsync(dEntry_.get(), he, sizeof(dev::BatchEntry), hipMemcpyHostToDevice, stream_));
        const dev::BatchEntry* entry = dEntry_.as<const dev::BatchEntry>();
        stats_.graphTailLevel = ~0u;
        if (profiling_) {
            std::memset(stats_.stageMs, 0, sizeof(stats_.stageMs));
            enqueueFrame(stream_, true, dConsts_.as<const FrameConsts>(), entry);
            MRT_HIP(hipMemcpyAsync(hCounters_.get(), dCounters_.get(), dev::kCounterWords * 4,
                                   hipMemcpyDeviceToHost, stream_));
            MRT_HIP(hipStreamSynchronize(stream_));
            const uint32_t* c = static_cast<const uint32_t*>(hCounters_.get());
            stats_.guideRows = 0;
            for (uint32_t b = 0; b < dev::kMaxBounceLevels; ++b) {
                stats_.guideRows += c[dev::cntWord(b, dev::kCntGuideRecords, 0)];
                stats_.rays[b] = 0;
                stats_.shadowRays[b] = 0;
                for (uint32_t q = 0; q < dev::kQueueSegs; ++q) {
                    stats_.rays[b] += c[dev::cntWord(b, dev::kCntRays, q)] + c[dev::cntWord(b, dev::kCntTailRays, q)];
                    stats_.shadowRays[b] +=
                        c[dev::cntWord(b, dev::kCntShadows, q)] + c[dev::cntWord(b, dev::kCntTailShadows, q)];
                }
            }
            stats_.status |= c[dev::kStatusWord];
        } else if (useGraph_) {
            if (graphExec_ && graphTrainSteps_ != trainSteps(0)) invalidateGraph();  // warm-up over
            if (graphExec_ && !tailLevelFits(graphTailLevel_)) invalidateGraph();
            if (!graphExec_) {
                graphTrainSteps_ = trainSteps(0);
                graphTailLevel_ = fittedTailLevel();
                MRT_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
                try {
                    enqueueFrame(stream_, false, dConsts_.as<const FrameConsts>(), entry, 0, graphTailLevel_);
                } catch (...) {
                    hipGraph_t g2 = nullptr;
                    (void)hipStreamEndCapture(stream_, &g2);
                    if (g2) (void)hipGraphDestroy(g2);
                    throw;
                }
                MRT_HIP(hipStreamEndCapture(stream_, &graph_));
                MRT_HIP(hipGraphInstantiate(&graphExec_, graph_, nullptr, nullptr, 0));
            }
            MRT_HIP(hipGraphLaunch(graphExec_, stream_));
            MRT_HIP(hipStreamSynchronize(stream_));
            stats_.graphTailLevel = graphTailLevel_;
        } else {
            enqueueFrame(stream_, false, dConsts_.as<const FrameConsts>(), entry);
            MRT_HIP(hipStreamSynchronize(stream_));
        }
        levelRaysFresh_ = true;  // kFinishBatch wrote this frame's queue profile
        if (pathView().pathRec) ++trainedFrames_;
        ++frameIndex_;
    }
    stats_.frameMs = msSince(t0);
    if (out) {
        out->spp = frameIndex_;
        out->converged = frameIndex_ >= settings_.sppLimit ? 1 : 0;
        out->denoised = 0;
        out->frameMs = stats_.frameMs;
    }
}

// Reference Renderer.cpp:892-909.  The RGBA32F division runs on the GPU.
void Engine::readFramebuffer(mrtPixelFormat fmt, void* dst, size_t dstSize) {
    const size_t pixels = size_t(frameW_) * frameH_;
    DeviceScope scope(device_);
    if (fmt == MRT_PIXEL_RGBA8_SRGB) {
        if (dstSize < pixels * 4) throw Error(MRT_ERROR_INVALID_ARGUMENT, "destination too small for RGBA8");
        MRT_HIP(hipMemcpyAsync(dst, dOut8_.get(), pixels * 4, hipMemcpyDeviceToHost, stream_));
        MRT_HIP(hipStreamSynchronize(stream_));
        return;
    }
    if (fmt != MRT_PIXEL_RGBA32F_LINEAR) throw Error(MRT_ERROR_INVALID_ARGUMENT, "unknown pixel format");
    if (dstSize < pixels * 16) throw Error(MRT_ERROR_INVALID_ARGUMENT, "destination too small for RGBA32F");
    MRT_HIP(dev::launchResolve32(pathView(), uint32_t(pixels), frameIndex_, stream_));
    MRT_HIP(hipMemcpyAsync(dst, dOut32_.get(), pixels * 16, hipMemcpyDeviceToHost, stream_));
    MRT_HIP(hipStreamSynchronize(stream_));
}

// -------------------------------------------------------------------- batching --
// Captures the next `frames` frames (clamped to the spp limit) as one graph:
// their constants are uploaded to a device array first, so every frame's
// kernels read their own slot and the chain needs no host step between frames.
// Commits, then uploads the constants of the next `frames` frames (clamped to
// the spp limit) to a device array; returns the frame count.
uint32_t Engine::prepareFrames(uint32_t frames, const FrameConsts** consts) {
    *consts = nullptr;
    commit(nullptr);
    DeviceScope scope(device_);
    MRT_HIP(hipStreamSynchronize(stream_));
    const uint32_t left = frameIndex_ < settings_.sppLimit ? settings_.sppLimit - frameIndex_ : 0;
    frames = std::min(frames, left);
    if (frames == 0) return 0;
    dBatchConsts_.ensure(size_t(frames) * sizeof(FrameConsts));
    hBatchConsts_.ensure(size_t(frames) * sizeof(FrameConsts));
    FrameConsts* h = static_cast<FrameConsts*>(hBatchConsts_.get());
    const uint32_t f0 = frameIndex_, s0 = guideStamp_;
    for (uint32_t i = 0; i < frames; ++i) {
        frameIndex_ = f0 + i;
        if (++guideStamp_ == 0) guideStamp_ = 1;
        fillConsts(h[i]);
    }
    frameIndex_ = f0;
    guideStamp_ = s0;
    MRT_HIP(hipMemcpyAsync(dBatchConsts_.get(), h, size_t(frames) * sizeof(FrameConsts), hipMemcpyHostToDevice,
                           stream_));
    MRT_HIP(hipStreamSynchronize(stream_));
    *consts = dBatchConsts_.as<const FrameConsts>();
    return frames;
}

uint32_t Engine::captureFrames(uint32_t frames, hipGraph_t* out) {
    *out = nullptr;
    const FrameConsts* consts = nullptr;
    frames = prepareFrames(frames, &consts);
    if (frames == 0) return 0;
    DeviceScope scope(device_);
    std::vector<dev::BatchEntry> entries(frames);
    fillEntries(entries.data(), consts, frames);
    dBatchEntries_.ensure(frames * sizeof(dev::BatchEntry));
    dBatchEntries_.upload(entries.data(), frames * sizeof(dev::BatchEntry), 0, stream_);
    const dev::BatchEntry* dEntries = dBatchEntries_.as<const dev::BatchEntry>();
    MRT_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    try {
        for (uint32_t i = 0; i < frames; ++i) enqueueFrame(stream_, false, consts + i, dEntries + i, i);
    } catch (...) {
        hipGraph_t g2 = nullptr;
        (void)hipStreamEndCapture(stream_, &g2);
        if (g2) (void)hipGraphDestroy(g2);
        throw;
    }
    MRT_HIP(hipStreamEndCapture(stream_, out));
    return frames;
}

// The captured frames ran: advance the frame counters as renderFrame would.
void Engine::framesDone(uint32_t frames) {
    for (uint32_t i = 0; i < frames; ++i)
        if (++guideStamp_ == 0) guideStamp_ = 1;
    if (pathView().pathRec) trainedFrames_ += frames;
    frameIndex_ += frames;
}

// Launch-shape autotuner for the path tracer: one profiled frame per
// candidate shape; the one with the least time in the trace and tail
// kernels wins.  Traversal results do not depend on the shape, so neither does
// the image; the trial frames are discarded (accumulation restarts) and do
// not train a guiding network.  Results are cached per (triangle count, pixel
// count, depth, compute units) class.
const uint32_t Engine::kShapeCand[Engine::kShapeCandidates][2] = { { 64, 16 }, { 128, 8 }, { 128, 16 },
                                                                    { 128, 32 }, { 256, 8 }, { 256, 16 } };

bool Engine::shapeTimes(float out[kShapeCandidates]) const {
    if (!shapeTimed_) return false;
    std::copy(shapeMs_, shapeMs_ + kShapeCandidates, out);
    return true;
}

void Engine::autotuneShapes(mrtxPathOptions& chosen) {
    struct Key {
        int tris, pixels, depth, cus;
        bool operator<(const Key& o) const {
            return std::tie(tris, pixels, depth, cus) < std::tie(o.tris, o.pixels, o.depth, o.cus);
        }
    };
    static std::mutex cacheMutex;
    struct Entry {
        dev::TraceShape extend, shadow;
        std::array<float, kShapeCandidates> ms;
    };
    static std::map<Key, Entry> cache;
    commit(nullptr);
    // The class includes the device's compute-unit count: the best shape
    // depends on how many waves the chip holds.
    const Key key{ int(std::lround(std::log2(double(std::max<size_t>(totalTris_, 1))) * 2.0)),
                   int(std::lround(std::log2(double(frameW_) * frameH_) * 2.0)),
                   int(std::min(settings_.maxBounces, 16u)), cuCount_ };
    bool hit = false;
    {
        std::lock_guard<std::mutex> lock(cacheMutex);
        auto it = cache.find(key);
        if (it != cache.end()) {
            extendShape_.block = it->second.extend.block;
            extendShape_.stack = it->second.extend.stack;
            shadowShape_.block = it->second.shadow.block;
            shadowShape_.stack = it->second.shadow.stack;
            std::copy(it->second.ms.begin(), it->second.ms.end(), shapeMs_);
            hit = true;
        }
    }
    if (!hit) {
        // Trial frames run profiled, without an spp limit and without
        // training (so the network, its Adam state and the training-record
        // stamp are as before); the guard restores all of it on any exit.
        struct TrialState {
            Engine& e;
            bool profiling;
            uint32_t sppLimit, training, stamp;
            ~TrialState() {
                e.profiling_ = profiling;
                e.settings_.sppLimit = sppLimit;
                e.guide_.enableTraining = training;
                e.guideStamp_ = stamp;
            }
        } saved{ *this, profiling_, settings_.sppLimit, guide_.enableTraining, guideStamp_ };
        settings_.sppLimit = std::max(saved.sppLimit, 1u << 30);
        profiling_ = true;
        guide_.enableTraining = 0;
        // One shape for both trace launch kinds and the tail kernel: the one
        // with the least time in those kernels over a whole frame.
        float best = 1e30f;
        dev::TraceShape be = extendShape_, bs = shadowShape_;
        for (int k = 0; k < kShapeCandidates; ++k) {
            extendShape_.block = shadowShape_.block = kShapeCand[k][0];
            extendShape_.stack = shadowShape_.stack = kShapeCand[k][1];
            chooseShapes();
            resetAccumulation();
            renderFrame(nullptr);
            const float t = stats_.stageMs[1] + stats_.stageMs[3] + stats_.stageMs[5];
            shapeMs_[k] = t;
            if (t < best) {
                best = t;
                be = extendShape_;
                bs = shadowShape_;
            }
        }
        extendShape_ = be;
        shadowShape_ = bs;
        Entry entry{ be, bs, {} };
        std::copy(shapeMs_, shapeMs_ + kShapeCandidates, entry.ms.begin());
        std::lock_guard<std::mutex> lock(cacheMutex);
        cache[key] = entry;
    }
    shapeTimed_ = true;
    chooseShapes();
    resetAccumulation();
    getPathOptions(chosen);
}

// ------------------------------------------------------------------ guiding --
void Engine::setGuiding(const mrtxGuidingOptions& in) {
    mrtxGuidingOptions o = in;
    o.structSize = sizeof(o);
    if (o.selectProb == 0.0f) o.selectProb = 0.5f;
    if (o.learningRate == 0.0f) o.learningRate = 1e-3f;
    if (o.batchSize == 0) o.batchSize = 65536;
    if (o.stepsPerFrame == 0) o.stepsPerFrame = 4;
    if (o.warmupFrames == 0) o.warmupFrames = 32;
    if (o.warmupStepsPerFrame == 0) o.warmupStepsPerFrame = 16;
    if (o.positionScale == 0.0f) o.positionScale = 1.0f;
    if (!(o.positionScale > 0.0f) || !std::isfinite(o.positionScale))
        throw Error(MRT_ERROR_INVALID_ARGUMENT, "guiding positionScale must be > 0");
    if (!(o.selectProb > 0.0f && o.selectProb <= 1.0f))
        throw Error(MRT_ERROR_INVALID_ARGUMENT, "guiding selectProb must be in (0, 1]");
    if (!(o.learningRate > 0.0f) || !std::isfinite(o.learningRate))
        throw Error(MRT_ERROR_INVALID_ARGUMENT, "guiding learningRate must be > 0");
    if (o.batchSize < 256 || o.batchSize > (1u << 20))
        throw Error(MRT_ERROR_INVALID_ARGUMENT, "guiding batchSize must be in [256, 1048576]");
    if (o.stepsPerFrame > 64) throw Error(MRT_ERROR_INVALID_ARGUMENT, "guiding stepsPerFrame must be <= 64");
    if (o.warmupStepsPerFrame > 64)
        throw Error(MRT_ERROR_INVALID_ARGUMENT, "guiding warmupStepsPerFrame must be <= 64");
    if (!(o.minRoughness >= 0.0f)) throw Error(MRT_ERROR_INVALID_ARGUMENT, "guiding minRoughness must be >= 0");
    DeviceScope scope(device_);
    MRT_HIP(hipStreamSynchronize(stream_));
    const bool realloc = o.batchSize != guide_.batchSize || o.stepsPerFrame != guide_.stepsPerFrame ||
                         o.warmupStepsPerFrame != guide_.warmupStepsPerFrame ||
                         guidePaths_ != frameW_ * frameH_ || guideLevels_ != guideLevels();
    guide_ = o;
    if (guideActive()) {
        ensureGuideNet();
        if (realloc || !dPathRec_.get()) allocGuide();
    }
    invalidateGraph();
}

void Engine::guidingStats(mrtxGuidingStats& o) {
    o = mrtxGuidingStats{};
    o.structSize = sizeof(o);
    if (!guideNet_) return;
    DeviceScope scope(device_);
    MRT_HIP(hipStreamSynchronize(stream_));
    uint32_t w[dev::kGwWords];
    guideNet_->readWords(w, stream_);
    o.steps = w[dev::kGwStep];
    o.lastBatch = w[dev::kGwLastBatch];
    std::memcpy(&o.lastLoss, &w[dev::kGwLastLoss], sizeof(float));
    o.poolRecords = std::min(w[dev::kGwPool], lastPoolCap_);
    o.candidates = w[dev::kGwCand0 + (guideStamp_ & 1u)];
    o.lastFrameRows = stats_.guideRows;
    o.envBrightShare = hasEnv_ ? envSplitA_ : 0.0f;
    if (dNeeStats_.get() && envSplitA_ > 0.0f) {
        // the rule of neeSelect (pt_device.h)
        std::vector<float> c(size_t(kNeeCells) * 4);
        dNeeStats_.download(c.data(), c.size() * 4, 0, stream_);
        MRT_HIP(hipStreamSynchronize(stream_));
        double sum = 0.0;
        for (uint32_t i = 0; i < kNeeCells; ++i) {
            const float* q = &c[size_t(i) * 4];
            if (q[1] < 16.0f || q[3] < 16.0f) continue;
            const float a = std::sqrt(q[0] / q[1]), b = std::sqrt(q[2] / q[3]);
            if (!(a + b > 0.0f)) continue;
            sum += std::clamp(a / (a + b), 0.05f, 0.95f);
            ++o.envCellsLearned;
        }
        o.envMeanSplit = o.envCellsLearned ? float(sum / o.envCellsLearned) : 0.0f;
    }
}

void Engine::getGuideParams(float* dst) {
    DeviceScope scope(device_);
    MRT_HIP(hipStreamSynchronize(stream_));
    ensureGuideNet();
    guideNet_->getParams(dst);
}

void Engine::setGuideParams(const float* src) {
    DeviceScope scope(device_);
    MRT_HIP(hipStreamSynchronize(stream_));
    ensureGuideNet();
    guideNet_->setParams(src);
    trainedFrames_ = 0;
}

void Engine::exportScene(mrtxSceneExport& o) {
    commit(nullptr);
    DeviceScope scope(device_);
    MRT_HIP(hipStreamSynchronize(stream_));
    // the arrays as laid out on the device: after edits they may hold
    // ranges no mesh uses, and meshes behind the TLAS region
    o.blasNodeCount = tlasRoot_;
    o.tlasRoot = tlasRoot_;
    o.triCount = triEnd_;
    o.vertexCount = vertEnd_;
    o.instanceCount = instanceCount_;
    o.materialCount = uint32_t(scene_.materials().size() + 1);
    o.lightCount = lightCount_;
    o.envWidth = hasEnv_ ? envW_ : 0;
    o.envHeight = hasEnv_ ? envH_ : 0;
    o.envIntegral = envIntegral_;
    o.nodeCount = nodeEnd_ > tlasRoot_ + tlasCap_ ? nodeEnd_ : tlasRoot_ + tlasNodeCount_;
    if (o.nodes) dNodes_.download(o.nodes, size_t(o.nodeCount) * 32);
    if (o.triangles) dTris_.download(o.triangles, size_t(triEnd_) * 16);
    if (o.positions) dPos_.download(o.positions, size_t(vertEnd_) * 16);
    if (o.normals) dNrm_.download(o.normals, size_t(vertEnd_) * 16);
    if (o.uvs) dUv_.download(o.uvs, size_t(vertEnd_) * 8);
    if (o.instances) dInst_.download(o.instances, size_t(instanceCount_) * sizeof(GpuInstance));
    if (o.materials) dMats_.download(o.materials, size_t(o.materialCount) * sizeof(GpuMaterial));
    if (o.lights) dLights_.download(o.lights, size_t(lightCount_) * sizeof(GpuLight));
    if (o.envAlias && hasEnv_) dAlias_.download(o.envAlias, size_t(envW_) * envH_ * sizeof(AliasEntry));
    if (o.structSize < sizeof(mrtxSceneExport)) return;
    o.tex8Count = uint32_t(tex8Texels_);
    o.tex32Count = uint32_t(tex32Texels_);
    if (o.tangents) dTan_.download(o.tangents, size_t(vertEnd_) * 16);
    if (o.textureTable) dTexTable_.download(o.textureTable, size_t(kMaxTextureSlots) * sizeof(GpuTexture));
    if (o.tex8 && o.tex8Count) dTex8_.download(o.tex8, size_t(o.tex8Count) * 4);
    if (o.tex32 && o.tex32Count) dTex32_.download(o.tex32, size_t(o.tex32Count) * 16);
}

} // namespace mrt
