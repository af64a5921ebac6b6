// mrtx ray-query scene: one BVH over a triangle set, queried on device ray
// batches.  The BVH is the reference's binned-SAH BVH2 (Bvh.cpp:137-156,
// maxLeaf 4) over triangle bounds accumulated as in Renderer.cpp:348-356.
#include "bvh_build.h"
#include "bvh_gpu_build.h"
#include "bvh_wide.h"
#include "common.h"
#include "mini_raytrace_ext.h"
#include "../kernels/launch.h"

#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <vector>

struct mrtxQueryScene {
    int device = 0;
    std::vector<mrt::BvhNode> nodes;
    std::vector<uint32_t> primOrder;
    mrt::DeviceBuffer dNodes, dWide, dTris, dStatus;
    uint32_t triCount = 0;
    uint32_t leafCount = 0;
    uint32_t maxDepth = 0;
    float buildMs = 0.0f;
    // Default launch shape: the fastest measured on a 10k-triangle soup
    // (the four-wide step walk).
    uint32_t block = 128;
    uint32_t stack = 8;
    uint32_t loop = 2;
};

namespace {

using mrt::Error;

// Deepest stack a depth-d tree needs is d entries; the kernels hold the top
// `stack` entries in LDS and spill up to 64 more to scratch.
constexpr uint32_t kMaxTraversalDepth = 64 + 8;
// Traversal stack entries hold node and triangle indices in 27 bits.
constexpr size_t kMaxTraversalIndex = (size_t(1) << 27) - 1;

mrtxQueryScene* createScene(const float* positions, uint32_t vertexCount, const uint32_t* indices,
                            uint32_t triCount, uint32_t gpuIndex1) {
    if (triCount > 0 && (!positions || !indices))
        throw Error(MRT_ERROR_INVALID_ARGUMENT, "positions and indices are required");
    for (size_t i = 0; i < size_t(triCount) * 3; ++i)
        if (indices[i] >= vertexCount)
            throw Error(MRT_ERROR_INVALID_ARGUMENT, "triangle index out of range");

    auto* sc = new mrtxQueryScene();
    try {
        sc->device = mrt::selectDevice(gpuIndex1);
        mrt::DeviceScope scope(sc->device);

        const auto t0 = std::chrono::steady_clock::now();
        std::vector<mrt::Box3> bounds(triCount);
        for (uint32_t t = 0; t < triCount; ++t)
            for (int k = 0; k < 3; ++k) bounds[t].grow(positions + size_t(indices[size_t(t) * 3 + k]) * 3);
        mrt::BvhBuild b = mrt::buildBvh(bounds, 4);
        const auto t1 = std::chrono::steady_clock::now();
        sc->buildMs = std::chrono::duration<float, std::milli>(t1 - t0).count();
        sc->triCount = triCount;
        sc->leafCount = b.leafCount;
        sc->maxDepth = b.maxDepth;
        if (b.maxDepth > kMaxTraversalDepth)
            throw Error(MRT_ERROR_UNKNOWN, "BVH depth " + std::to_string(b.maxDepth) +
                                               " exceeds the traversal stack");

        // Leaf-ordered, de-indexed vertices: 3 float4 per triangle, v0.w = id bits.
        std::vector<float> tris;
        if (triCount == 0) {
            // Sentinel: one leaf holding one NaN triangle (never hit).
            mrt::BvhNode s{};
            for (int a = 0; a < 3; ++a) {
                s.lo[a] = std::numeric_limits<float>::max();
                s.hi[a] = -std::numeric_limits<float>::max();
            }
            s.leftOrPrim = 0;
            s.packed = 1u << 2;
            b.nodes.push_back(s);
            tris.assign(12, std::numeric_limits<float>::quiet_NaN());
            const uint32_t none = MRTX_NO_HIT;
            std::memcpy(&tris[3], &none, 4);
        } else {
            tris.resize(size_t(triCount) * 12);
            for (uint32_t slot = 0; slot < triCount; ++slot) {
                const uint32_t orig = b.primOrder[slot];
                float* dst = &tris[size_t(slot) * 12];
                for (int k = 0; k < 3; ++k) {
                    const float* p = positions + size_t(indices[size_t(orig) * 3 + k]) * 3;
                    dst[k * 4 + 0] = p[0];
                    dst[k * 4 + 1] = p[1];
                    dst[k * 4 + 2] = p[2];
                    dst[k * 4 + 3] = 0.0f;
                }
                std::memcpy(&dst[3], &orig, 4);
            }
        }
        sc->nodes = std::move(b.nodes);
        sc->primOrder = std::move(b.primOrder);
        if (sc->nodes.size() > kMaxTraversalIndex || triCount > kMaxTraversalIndex)
            throw Error(MRT_ERROR_INVALID_ARGUMENT, "too many triangles for one query scene");
        // Device copy behind one unused leading node: the root then sits at
        // byte 32 and every child pair (indices 2k+1, 2k+2) on a 64-byte
        // boundary, so the two children a traversal step reads share a line.
        std::vector<mrt::BvhNode> dev(sc->nodes.size() + 1);
        std::memcpy(dev.data() + 1, sc->nodes.data(), sc->nodes.size() * sizeof(mrt::BvhNode));
        sc->dNodes.allocate(dev.size() * sizeof(mrt::BvhNode));
        sc->dNodes.upload(dev.data(), dev.size() * sizeof(mrt::BvhNode));
        const std::vector<mrt::WideNode> wide = mrt::collapseWide(sc->nodes);
        if (wide.size() > kMaxTraversalIndex)
            throw Error(MRT_ERROR_INVALID_ARGUMENT, "too many triangles for one query scene");
        sc->dWide.allocate(wide.size() * sizeof(mrt::WideNode));
        sc->dWide.upload(wide.data(), wide.size() * sizeof(mrt::WideNode));
        sc->dTris.allocate(tris.size() * sizeof(float));
        sc->dTris.upload(tris.data(), tris.size() * sizeof(float));
        sc->dStatus.allocate(sizeof(uint32_t));
        sc->dStatus.zero();
        MRT_HIP(hipDeviceSynchronize());
    } catch (...) {
        delete sc;
        throw;
    }
    return sc;
}

mrt::dev::QueryLaunch makeLaunch(mrtxQueryScene* s, const mrtxRay* rays, uint32_t n, void* out) {
    mrt::dev::QueryLaunch q;
    // skip the alignment node
    q.nodes = reinterpret_cast<const mrt::dev::Node*>(s->dNodes.as<const char>() + sizeof(mrt::BvhNode));
    q.wide = s->dWide.as<const float4>();
    q.tris = s->dTris.as<const float4>();
    q.rays = reinterpret_cast<const float4*>(rays);
    q.n = n;
    q.out = out;
    q.status = s->dStatus.as<uint32_t>();
    return q;
}

} // namespace

extern "C" {

MRT_API mrtResult mrtxBuildBvh(const float* bounds, uint32_t count, uint32_t maxLeaf, void* nodes,
                               uint64_t nodeCap, uint32_t* primOrder, uint64_t* outNodeCount) {
    if (count > 0 && (!bounds || !nodes || !primOrder)) return MRT_ERROR_INVALID_ARGUMENT;
    return mrt::guarded([&] {
        std::vector<mrt::Box3> b(count);
        for (uint32_t i = 0; i < count; ++i)
            for (int a = 0; a < 3; ++a) {
                b[i].lo[a] = bounds[size_t(i) * 6 + a];
                b[i].hi[a] = bounds[size_t(i) * 6 + 3 + a];
            }
        const mrt::BvhBuild r = mrt::buildBvh(b, maxLeaf);
        if (outNodeCount) *outNodeCount = r.nodes.size();
        if (r.nodes.size() > nodeCap) throw Error(MRT_ERROR_INVALID_ARGUMENT, "node capacity too small");
        if (!r.nodes.empty()) std::memcpy(nodes, r.nodes.data(), r.nodes.size() * sizeof(mrt::BvhNode));
        if (count) std::memcpy(primOrder, r.primOrder.data(), size_t(count) * sizeof(uint32_t));
        return MRT_SUCCESS;
    });
}

MRT_API mrtResult mrtxBuildBvhDevice(const float* dBounds, uint32_t count, uint32_t maxLeaf, void* dNodes,
                                     uint64_t nodeCap, uint32_t* dPrimOrder, uint64_t* outNodeCount,
                                     void* hipStream) {
    if (count > 0 && (!dBounds || !dNodes || !dPrimOrder)) return MRT_ERROR_INVALID_ARGUMENT;
    if (count > 0 && nodeCap < uint64_t(2) * count - 1) {
        mrt::setLastErrorMessage("mrtxBuildBvhDevice: nodeCap must be at least 2 * count - 1");
        return MRT_ERROR_INVALID_ARGUMENT;
    }
    return mrt::guarded([&] {
        hipStream_t s = static_cast<hipStream_t>(hipStream);
        // One builder per device keeps its workspace (about 100 bytes per
        // primitive of the largest build so far) until the process ends;
        // calls on one device take turns.
        static std::mutex lock;
        static auto* builders = new std::map<int, mrt::GpuBvhBuilder*>;
        int device = 0;
        MRT_HIP(hipGetDevice(&device));
        std::lock_guard<std::mutex> g(lock);
        mrt::GpuBvhBuilder*& builder = (*builders)[device];
        if (!builder) builder = new mrt::GpuBvhBuilder;
        mrt::GpuBvhInfo info;
        uint64_t nodes = 0;
        if (builder->build(dBounds, count, maxLeaf, dNodes, nodeCap, dPrimOrder, info, s)) {
            nodes = info.nodeCount;
        } else {
            // the median split: build on the CPU from the same bounds
            mrt::logMessage(mrt::LogLevel::Info,
                            "mrtxBuildBvhDevice: a node needs the median split; building on the CPU");
            std::vector<mrt::Box3> b(count);
            MRT_HIP(hipMemcpyAsync(b.data(), dBounds, size_t(count) * sizeof(mrt::Box3), hipMemcpyDeviceToHost, s));
            MRT_HIP(hipStreamSynchronize(s));
            const mrt::BvhBuild r = mrt::buildBvh(b, maxLeaf);
            nodes = r.nodes.size();
            MRT_HIP(hipMemcpyAsync(dNodes, r.nodes.data(), nodes * sizeof(mrt::BvhNode), hipMemcpyHostToDevice, s));
            MRT_HIP(hipMemcpyAsync(dPrimOrder, r.primOrder.data(), size_t(count) * 4, hipMemcpyHostToDevice, s));
            MRT_HIP(hipStreamSynchronize(s));
        }
        if (outNodeCount) *outNodeCount = nodes;
        return MRT_SUCCESS;
    });
}

MRT_API mrtResult mrtxCreateQueryScene(const float* positions, uint32_t vertexCount,
                                       const uint32_t* indices, uint32_t triCount,
                                       uint32_t gpuIndex1, mrtxQueryScene** outScene) {
    if (!outScene) return MRT_ERROR_INVALID_ARGUMENT;
    return mrt::guarded([&] {
        *outScene = createScene(positions, vertexCount, indices, triCount, gpuIndex1);
        return MRT_SUCCESS;
    });
}

MRT_API void mrtxDestroyQueryScene(mrtxQueryScene* scene) {
    if (!scene) return;
    {
        mrt::DeviceScope scope(scene->device);
        (void)hipDeviceSynchronize();
    }
    delete scene;
}

MRT_API mrtResult mrtxGetQuerySceneInfo(const mrtxQueryScene* s, mrtxQuerySceneInfo* out) {
    if (!s) return MRT_ERROR_INVALID_HANDLE;
    if (!out || out->structSize < sizeof(mrtxQuerySceneInfo)) return MRT_ERROR_INVALID_ARGUMENT;
    out->triCount = s->triCount;
    out->nodeCount = uint32_t(s->nodes.size());
    out->leafCount = s->leafCount;
    out->maxDepth = s->maxDepth;
    out->buildMs = s->buildMs;
    out->deviceBytes = s->dNodes.size() + s->dWide.size() + s->dTris.size() + s->dStatus.size();
    return MRT_SUCCESS;
}

MRT_API mrtResult mrtxExportQueryBvh(const mrtxQueryScene* s, void* nodes, size_t nodesBytes,
                                     uint32_t* primOrder, size_t primCount) {
    if (!s) return MRT_ERROR_INVALID_HANDLE;
    if (nodes) {
        const size_t need = s->nodes.size() * sizeof(mrt::BvhNode);
        if (nodesBytes < need) return MRT_ERROR_INVALID_ARGUMENT;
        std::memcpy(nodes, s->nodes.data(), need);
    }
    if (primOrder) {
        if (primCount < s->primOrder.size()) return MRT_ERROR_INVALID_ARGUMENT;
        std::memcpy(primOrder, s->primOrder.data(), s->primOrder.size() * sizeof(uint32_t));
    }
    return MRT_SUCCESS;
}

MRT_API mrtResult mrtxSetTraversalConfig(mrtxQueryScene* s, const mrtxTraversalConfig* cfg) {
    if (!s) return MRT_ERROR_INVALID_HANDLE;
    if (!cfg || cfg->structSize < sizeof(mrtxTraversalConfig)) return MRT_ERROR_INVALID_ARGUMENT;
    const uint32_t block = cfg->blockSize ? cfg->blockSize : 128;
    const uint32_t stack = cfg->stackDepth ? cfg->stackDepth : 8;
    if (!mrt::dev::traversalShapeSupported(block, stack, cfg->loopMode)) {
        mrt::setLastErrorMessage("unsupported traversal shape");
        return MRT_ERROR_INVALID_ARGUMENT;
    }
    s->block = block;
    s->stack = stack;
    s->loop = cfg->loopMode;
    return MRT_SUCCESS;
}

MRT_API mrtResult mrtxGetTraversalConfig(const mrtxQueryScene* s, mrtxTraversalConfig* out) {
    if (!s) return MRT_ERROR_INVALID_HANDLE;
    if (!out || out->structSize < sizeof(mrtxTraversalConfig)) return MRT_ERROR_INVALID_ARGUMENT;
    out->blockSize = s->block;
    out->stackDepth = s->stack;
    out->loopMode = s->loop;
    return MRT_SUCCESS;
}

MRT_API mrtResult mrtxTraceClosest(mrtxQueryScene* s, const mrtxRay* dRays, uint32_t n,
                                   mrtxHit* dHits, void* hipStream) {
    if (!s) return MRT_ERROR_INVALID_HANDLE;
    if (n > 0 && (!dRays || !dHits)) return MRT_ERROR_INVALID_ARGUMENT;
    return mrt::guarded([&] {
        mrt::DeviceScope scope(s->device);
        MRT_HIP(mrt::dev::launchTraceClosest(makeLaunch(s, dRays, n, dHits), s->block, s->stack,
                                             s->loop, static_cast<hipStream_t>(hipStream)));
        return MRT_SUCCESS;
    });
}

MRT_API mrtResult mrtxTraceAny(mrtxQueryScene* s, const mrtxRay* dRays, uint32_t n,
                               uint32_t* dOccluded, void* hipStream) {
    if (!s) return MRT_ERROR_INVALID_HANDLE;
    if (n > 0 && (!dRays || !dOccluded)) return MRT_ERROR_INVALID_ARGUMENT;
    return mrt::guarded([&] {
        mrt::DeviceScope scope(s->device);
        MRT_HIP(mrt::dev::launchTraceAny(makeLaunch(s, dRays, n, dOccluded), s->block, s->stack,
                                         s->loop, static_cast<hipStream_t>(hipStream)));
        return MRT_SUCCESS;
    });
}

MRT_API mrtResult mrtxGetQueryStatus(mrtxQueryScene* s, uint32_t* flags) {
    if (!s) return MRT_ERROR_INVALID_HANDLE;
    if (!flags) return MRT_ERROR_INVALID_ARGUMENT;
    return mrt::guarded([&] {
        mrt::DeviceScope scope(s->device);
        MRT_HIP(hipDeviceSynchronize());
        s->dStatus.download(flags, sizeof(uint32_t));
        return MRT_SUCCESS;
    });
}

MRT_API mrtResult mrtxAutotuneTraversal(mrtxQueryScene* s, const mrtxRay* dRays, uint32_t n,
                                        mrtxHit* dHits, mrtxTraversalConfig* chosen) {
    if (!s) return MRT_ERROR_INVALID_HANDLE;
    if (n == 0 || !dRays || !dHits) return MRT_ERROR_INVALID_ARGUMENT;
    return mrt::guarded([&] {
        mrt::DeviceScope scope(s->device);
        hipStream_t st = nullptr;
        MRT_HIP(hipStreamCreate(&st));
        hipEvent_t e0 = nullptr, e1 = nullptr;
        MRT_HIP(hipEventCreate(&e0));
        MRT_HIP(hipEventCreate(&e1));
        float bestMs = 1e30f;
        uint32_t bestBlock = s->block, bestStack = s->stack, bestLoop = s->loop;
        const uint32_t blocks[] = { 64, 128, 256 };
        const uint32_t stacks[] = { 8, 16, 32 };
        for (uint32_t loop = 0; loop < 4; ++loop)
        for (uint32_t blk : blocks)
            for (uint32_t stk : stacks) {
                const auto q = makeLaunch(s, dRays, n, dHits);
                MRT_HIP(mrt::dev::launchTraceClosest(q, blk, stk, loop, st));  // warm-up
                MRT_HIP(hipEventRecord(e0, st));
                for (int rep = 0; rep < 3; ++rep) MRT_HIP(mrt::dev::launchTraceClosest(q, blk, stk, loop, st));
                MRT_HIP(hipEventRecord(e1, st));
                MRT_HIP(hipEventSynchronize(e1));
                float ms = 0.0f;
                MRT_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (ms < bestMs) {
                    bestMs = ms;
                    bestBlock = blk;
                    bestStack = stk;
                    bestLoop = loop;
                }
            }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipStreamDestroy(st);
        s->block = bestBlock;
        s->stack = bestStack;
        s->loop = bestLoop;
        if (chosen && chosen->structSize >= sizeof(mrtxTraversalConfig)) {
            chosen->blockSize = bestBlock;
            chosen->stackDepth = bestStack;
            chosen->loopMode = bestLoop;
        }
        return MRT_SUCCESS;
    });
}

} // extern "C"
