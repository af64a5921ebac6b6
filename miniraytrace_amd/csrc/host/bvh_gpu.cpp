// Host side of the GPU binned-SAH builder (kernels/bvh_gpu.hip): the level
// loop over large nodes, the small-subtree tasks and the final depth-first
// placement.  The result equals buildBvh's (bvh_build.cpp) node for node.
#include "bvh_gpu_build.h"

#include "common.h"
#include "kernels/bvh_gpu.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <cstring>
#include <functional>

namespace mrt {

namespace {

using dev::LargeChunk;
using dev::LargeNode;
using dev::OutNode;
using dev::SmallPlace;
using dev::SmallResult;
using dev::SmallTask;

static_assert(sizeof(OutNode) == sizeof(BvhNode), "device node record is the 32-byte node");

struct ChildRef {
    bool small;
    uint32_t id;
};

struct HostLarge {
    uint32_t begin, count, depth;
    LargeNode dec{};       // decision of the level pass
    bool leaf = false;
    ChildRef kid[2]{};
};

} // namespace

struct GpuBvhBuilder::Work {
    DeviceBuffer idx, pos, acc, counts, totals, level, tasks, results, tmp, placed, flags;
    PinnedBuffer up, down;  // staging of the per-level copies

    // Host arrays -> one pinned block -> one copy into dst; returns the
    // offsets of the parts (16-byte aligned).
    template <size_t N>
    std::array<size_t, N> upload(DeviceBuffer& dst, const std::array<std::pair<const void*, size_t>, N>& parts,
                                 hipStream_t s) {
        std::array<size_t, N> off{};
        size_t total = 0;
        for (size_t i = 0; i < N; ++i) {
            off[i] = total;
            total += (parts[i].second + 15) & ~size_t(15);
        }
        up.ensure(total);
        for (size_t i = 0; i < N; ++i)
            if (parts[i].second) std::memcpy(static_cast<char*>(up.get()) + off[i], parts[i].first, parts[i].second);
        dst.ensure(total);
        if (total) MRT_HIP(hipMemcpyAsync(dst.get(), up.get(), total, hipMemcpyHostToDevice, s));
        return off;
    }
};

GpuBvhBuilder::GpuBvhBuilder() : w_(new Work) {}
GpuBvhBuilder::~GpuBvhBuilder() { delete w_; }

bool GpuBvhBuilder::build(const float* dBounds, uint32_t n, uint32_t maxLeaf, void* dNodes, uint64_t nodeCap,
                          uint32_t* dOrder, GpuBvhInfo& info, hipStream_t s) {
    info = GpuBvhInfo{};
    if (n == 0) return true;
    if (n >= (1u << 31)) throw Error(MRT_ERROR_INVALID_ARGUMENT, "GPU BVH build: more than 2^31 - 1 primitives");
    const auto t0 = std::chrono::steady_clock::now();
    auto msSince = [](std::chrono::steady_clock::time_point t) {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
    };
    maxLeaf = std::max(1u, maxLeaf);
    Work& w = *w_;
    w.idx.ensure(size_t(n) * 4);
    w.pos.ensure(size_t(n) * 4);
    w.tmp.ensure(size_t(n) * 2 * sizeof(OutNode));
    w.flags.ensure(16);
    uint32_t* idx = w.idx.as<uint32_t>();
    uint32_t* flags = w.flags.as<uint32_t>();
    MRT_HIP(dev::launchBvhIota(idx, n, flags, s));

    std::vector<HostLarge> large;
    std::vector<SmallTask> tasks;
    std::vector<uint32_t> level, next;
    auto child = [&](uint32_t begin, uint32_t count, uint32_t depth) -> ChildRef {
        if (count > dev::kSmallMax) {
            large.push_back(HostLarge{ begin, count, depth });
            next.push_back(uint32_t(large.size() - 1));
            return ChildRef{ false, uint32_t(large.size() - 1) };
        }
        tasks.push_back(SmallTask{ begin, count, depth, 2 * begin });
        return ChildRef{ true, uint32_t(tasks.size() - 1) };
    };
    const ChildRef root = child(0, n, 0);
    level.swap(next);

    std::vector<LargeNode> ln;
    std::vector<LargeChunk> ch;
    uint32_t hostFlags = 0;
    while (!level.empty()) {
        ln.assign(level.size(), LargeNode{});
        ch.clear();
        for (size_t i = 0; i < level.size(); ++i) {
            const HostLarge& h = large[level[i]];
            LargeNode& d = ln[i];
            d.begin = h.begin;
            d.count = h.count;
            d.chunk0 = uint32_t(ch.size());
            for (uint32_t b = h.begin; b < h.begin + h.count; b += dev::kChunk)
                ch.push_back(LargeChunk{ uint32_t(i), b, std::min(b + dev::kChunk, h.begin + h.count) });
            d.chunks = uint32_t(ch.size()) - d.chunk0;
        }
        const uint32_t nn = uint32_t(ln.size()), nc = uint32_t(ch.size());
        w.acc.ensure(size_t(nn) * dev::kLargeAccWords * sizeof(int));
        w.counts.ensure(size_t(nc) * 2 * 4);
        w.totals.ensure(size_t(nn) * 2 * 4);
        const auto off = w.upload<2>(w.level, { { { ln.data(), nn * sizeof(LargeNode) },
                                                  { ch.data(), nc * sizeof(LargeChunk) } } }, s);
        LargeNode* dLevel = reinterpret_cast<LargeNode*>(static_cast<char*>(w.level.get()) + off[0]);
        dev::LargeLevelArgs a;
        a.bounds = dBounds;
        a.idx = idx;
        a.nodes = dLevel;
        a.nodeCount = nn;
        a.chunks = reinterpret_cast<const LargeChunk*>(static_cast<char*>(w.level.get()) + off[1]);
        a.chunkCount = nc;
        a.acc = w.acc.as<int>();
        a.counts = w.counts.as<uint32_t>();
        a.totals = w.totals.as<uint32_t>();
        a.pos = w.pos.as<uint32_t>();
        a.maxLeaf = maxLeaf;
        a.flags = flags;
        MRT_HIP(dev::launchBvhLargeLevel(a, s));
        w.down.ensure(nn * sizeof(LargeNode) + 16);
        MRT_HIP(hipMemcpyAsync(w.down.get(), dLevel, nn * sizeof(LargeNode), hipMemcpyDeviceToHost, s));
        MRT_HIP(hipMemcpyAsync(static_cast<char*>(w.down.get()) + nn * sizeof(LargeNode), flags, 4,
                               hipMemcpyDeviceToHost, s));
        MRT_HIP(hipStreamSynchronize(s));
        std::memcpy(ln.data(), w.down.get(), nn * sizeof(LargeNode));
        std::memcpy(&hostFlags, static_cast<char*>(w.down.get()) + nn * sizeof(LargeNode), 4);
        if (hostFlags & dev::kBuildNeedsMedian) return false;
        next.clear();
        for (size_t i = 0; i < level.size(); ++i) {
            const uint32_t id = level[i];
            const LargeNode& d = ln[i];
            large[id].dec = d;
            if (d.split == dev::kSplitLeaf) {
                large[id].leaf = true;
                continue;
            }
            const uint32_t begin = large[id].begin, count = large[id].count, depth = large[id].depth;
            const uint32_t mid = (d.mid > 0 && d.mid < count) ? d.mid : count / 2;
            const ChildRef l = child(begin, mid, depth + 1);
            const ChildRef r = child(begin + mid, count - mid, depth + 1);
            large[id].kid[0] = l;
            large[id].kid[1] = r;
        }
        level.swap(next);
        ++info.levels;
    }
    info.msLevels = msSince(t0);
    const auto t1 = std::chrono::steady_clock::now();

    const uint32_t nt = uint32_t(tasks.size());
    std::vector<SmallResult> res(nt);
    if (nt) {
        w.upload<1>(w.tasks, { { { tasks.data(), nt * sizeof(SmallTask) } } }, s);
        w.results.ensure(nt * sizeof(SmallResult));
        dev::SmallArgs a;
        a.bounds = dBounds;
        a.idx = idx;
        a.tasks = w.tasks.as<const SmallTask>();
        a.taskCount = nt;
        a.tmp = w.tmp.as<OutNode>();
        a.results = w.results.as<SmallResult>();
        a.maxLeaf = maxLeaf;
        a.flags = flags;
        MRT_HIP(dev::launchBvhSmall(a, s));
        w.down.ensure(nt * sizeof(SmallResult) + 16);
        MRT_HIP(hipMemcpyAsync(w.down.get(), w.results.get(), nt * sizeof(SmallResult), hipMemcpyDeviceToHost, s));
        MRT_HIP(hipMemcpyAsync(static_cast<char*>(w.down.get()) + nt * sizeof(SmallResult), flags, 4,
                               hipMemcpyDeviceToHost, s));
        MRT_HIP(hipStreamSynchronize(s));
        std::memcpy(res.data(), w.down.get(), nt * sizeof(SmallResult));
        std::memcpy(&hostFlags, static_cast<char*>(w.down.get()) + nt * sizeof(SmallResult), 4);
        if (hostFlags & dev::kBuildNeedsMedian) return false;
    }
    info.msSubtrees = msSince(t1);
    const auto t2 = std::chrono::steady_clock::now();

    // Depth-first placement: a node's child pair is allocated when the node
    // is reached, then its left subtree, then its right subtree.
    std::function<uint32_t(ChildRef)> sub = [&](ChildRef c) -> uint32_t {
        if (c.small) return res[c.id].nodes - 1;
        const HostLarge& h = large[c.id];
        if (h.leaf) return 0;
        return 2 + sub(h.kid[0]) + sub(h.kid[1]);
    };
    std::vector<OutNode> largeOut;
    std::vector<uint32_t> where;
    std::vector<SmallPlace> place(nt);
    uint32_t maxDepth = 0, leaves = 0;
    std::function<void(ChildRef, uint32_t, uint32_t)> assign = [&](ChildRef c, uint32_t at, uint32_t alloc) {
        if (c.small) {
            place[c.id] = SmallPlace{ at, alloc };
            maxDepth = std::max(maxDepth, res[c.id].maxDepth);
            leaves += res[c.id].leaves;
            return;
        }
        const HostLarge& h = large[c.id];
        OutNode o;
        std::memcpy(o.lo, h.dec.lo, sizeof(o.lo));
        std::memcpy(o.hi, h.dec.hi, sizeof(o.hi));
        if (h.leaf) {
            o.leftOrPrim = h.begin;
            o.packed = (h.count << 2) | h.dec.axis;
            maxDepth = std::max(maxDepth, h.depth);
            ++leaves;
        } else {
            o.leftOrPrim = alloc;
            o.packed = h.dec.axis;
        }
        largeOut.push_back(o);
        where.push_back(at);
        if (h.leaf) return;
        const uint32_t left = sub(h.kid[0]);
        assign(h.kid[0], alloc, alloc + 2);
        assign(h.kid[1], alloc + 1, alloc + 2 + left);
    };
    const uint32_t total = 1 + sub(root);
    if (total > nodeCap) throw Error(MRT_ERROR_INVALID_ARGUMENT, "node capacity too small");
    assign(root, 0, 1);

    dev::PlaceArgs p{};
    const uint32_t nl = uint32_t(largeOut.size());
    const auto po = w.upload<3>(w.placed, { { { largeOut.data(), nl * sizeof(OutNode) },
                                              { where.data(), nl * 4 },
                                              { place.data(), nt * sizeof(SmallPlace) } } }, s);
    const char* placed = static_cast<const char*>(w.placed.get());
    p.large = reinterpret_cast<const OutNode*>(placed + po[0]);
    p.largeWhere = reinterpret_cast<const uint32_t*>(placed + po[1]);
    p.largeCount = nl;
    p.tmp = w.tmp.as<const OutNode>();
    p.tasks = w.tasks.as<const SmallTask>();
    p.results = w.results.as<const SmallResult>();
    p.place = reinterpret_cast<const SmallPlace*>(placed + po[2]);
    p.taskCount = nt;
    p.dst = static_cast<OutNode*>(dNodes);
    MRT_HIP(dev::launchBvhPlace(p, s));
    MRT_HIP(hipMemcpyAsync(dOrder, idx, size_t(n) * 4, hipMemcpyDeviceToDevice, s));
    MRT_HIP(hipStreamSynchronize(s));
    info.msPlace = msSince(t2);
    info.nodeCount = total;
    info.maxDepth = maxDepth;
    info.leafCount = leaves;
    info.largeNodes = uint32_t(large.size());
    info.subtrees = nt;
    return true;
}

BvhBuild buildBvhOnDevice(GpuBvhBuilder& builder, const std::vector<Box3>& bounds, uint32_t maxLeaf, hipStream_t s,
                          bool* onGpu) {
    const uint32_t n = uint32_t(bounds.size());
    if (onGpu) *onGpu = false;
    if (n == 0) return BvhBuild{};
    static_assert(sizeof(Box3) == 24, "Box3 is lo xyz, hi xyz");
    DeviceBuffer dB(size_t(n) * 24), dNodes(size_t(2) * n * sizeof(BvhNode)), dOrder(size_t(n) * 4);
    dB.upload(bounds.data(), size_t(n) * 24, 0, s);
    GpuBvhInfo info;
    if (!builder.build(dB.as<const float>(), n, maxLeaf, dNodes.get(), uint64_t(2) * n, dOrder.as<uint32_t>(), info,
                       s))
        return buildBvh(bounds, maxLeaf);
    BvhBuild out;
    out.nodes.resize(info.nodeCount);
    out.primOrder.resize(n);
    dNodes.download(out.nodes.data(), size_t(info.nodeCount) * sizeof(BvhNode), 0, s);
    dOrder.download(out.primOrder.data(), size_t(n) * 4, 0, s);
    MRT_HIP(hipStreamSynchronize(s));
    for (int a = 0; a < 3; ++a) {
        out.rootBounds.lo[a] = out.nodes[0].lo[a];
        out.rootBounds.hi[a] = out.nodes[0].hi[a];
    }
    out.maxDepth = info.maxDepth;
    out.leafCount = info.leafCount;
    if (onGpu) *onGpu = true;
    return out;
}

} // namespace mrt
