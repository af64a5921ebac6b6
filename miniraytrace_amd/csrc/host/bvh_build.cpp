// Binned-SAH BVH2 builder; see bvh_build.h for the contract.
#include "bvh_build.h"

#include <algorithm>
#include <future>
#include <memory>
#include <thread>

namespace mrt {

float Box3::area() const {
    if (!valid()) return 0.0f;
    const float ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
    const float s = (ex * ey + ey * ez) + ez * ex;
    return 2.0f * s;
}

namespace {

constexpr uint32_t kBins = 32;
constexpr float kTraversalCost = 1.0f;
constexpr float kIntersectCost = 1.5f;
constexpr uint32_t kParallelMinPrims = 16384;

struct Ref {
    Box3 box;
    float c[3];
    uint32_t index;
};

// Temporary tree: subtrees are built independently (possibly on different
// threads) and linearised afterwards in depth-first pair order.
struct TNode {
    Box3 box;
    uint32_t begin = 0, count = 0, axis = 0;
    std::unique_ptr<TNode> kid[2];
    bool leaf() const { return !kid[0]; }
};

struct Ctx {
    std::vector<Ref>& refs;
    uint32_t maxLeaf;
    int parDepth;  // spawn threads only above this recursion depth
};

void build(Ctx& ctx, TNode& n, uint32_t begin, uint32_t end, int depth) {
    auto& refs = ctx.refs;
    const uint32_t count = end - begin;
    Box3 bounds, cbox;
    for (uint32_t i = begin; i < end; ++i) {
        bounds.grow(refs[i].box);
        cbox.grow(refs[i].c);
    }
    n.box = bounds;
    n.begin = begin;
    n.count = count;
    const float ce[3] = { cbox.hi[0] - cbox.lo[0], cbox.hi[1] - cbox.lo[1], cbox.hi[2] - cbox.lo[2] };
    uint32_t axis = 0;
    if (ce[1] > ce[0]) axis = 1;
    if (ce[2] > ce[axis]) axis = 2;
    n.axis = axis;
    if (count <= ctx.maxLeaf || ce[axis] < 1e-8f) return;

    Box3 binBox[kBins];
    uint32_t binCount[kBins] = {};
    const float k = float(kBins) * (1.0f - 1e-6f) / ce[axis];
    const float origin = cbox.lo[axis];
    auto binOf = [&](const Ref& r) { return static_cast<uint32_t>(k * (r.c[axis] - origin)); };
    for (uint32_t i = begin; i < end; ++i) {
        const uint32_t b = binOf(refs[i]);
        binBox[b].grow(refs[i].box);
        binCount[b]++;
    }

    float rightArea[kBins];
    uint32_t rightCount[kBins];
    {
        Box3 acc;
        uint32_t cnt = 0;
        for (int i = int(kBins) - 1; i > 0; --i) {
            acc.grow(binBox[i]);
            cnt += binCount[i];
            rightArea[i] = acc.area();
            rightCount[i] = cnt;
        }
    }
    float bestCost = kIntersectCost * float(count);
    int bestSplit = -1;
    {
        Box3 acc;
        uint32_t cnt = 0;
        const float invRoot = 1.0f / std::max(bounds.area(), 1e-12f);
        const float scale = kIntersectCost * invRoot;
        for (uint32_t i = 0; i + 1 < kBins; ++i) {
            acc.grow(binBox[i]);
            cnt += binCount[i];
            if (cnt == 0 || rightCount[i + 1] == 0) continue;
            const float l = acc.area() * float(cnt);
            const float r = rightArea[i + 1] * float(rightCount[i + 1]);
            const float cost = kTraversalCost + scale * (l + r);
            if (cost < bestCost) {
                bestCost = cost;
                bestSplit = int(i);
            }
        }
    }

    uint32_t mid;
    if (bestSplit < 0) {
        if (count <= ctx.maxLeaf * 2) return;
        mid = begin + count / 2;
        std::nth_element(refs.begin() + begin, refs.begin() + mid, refs.begin() + end,
                         [axis](const Ref& a, const Ref& b) { return a.c[axis] < b.c[axis]; });
    } else {
        const uint32_t split = uint32_t(bestSplit);
        auto it = std::partition(refs.begin() + begin, refs.begin() + end,
                                 [&](const Ref& r) { return binOf(r) <= split; });
        mid = static_cast<uint32_t>(it - refs.begin());
        if (mid == begin || mid == end) mid = begin + count / 2;
    }

    n.kid[0] = std::make_unique<TNode>();
    n.kid[1] = std::make_unique<TNode>();
    if (depth < ctx.parDepth && count >= kParallelMinPrims) {
        auto left = std::async(std::launch::async,
                               [&] { build(ctx, *n.kid[0], begin, mid, depth + 1); });
        build(ctx, *n.kid[1], mid, end, depth + 1);
        left.get();
    } else {
        build(ctx, *n.kid[0], begin, mid, depth + 1);
        build(ctx, *n.kid[1], mid, end, depth + 1);
    }
}

void emit(const TNode& t, uint32_t slot, uint32_t depth, BvhBuild& out) {
    BvhNode& n = out.nodes[slot];
    for (int a = 0; a < 3; ++a) {
        n.lo[a] = t.box.lo[a];
        n.hi[a] = t.box.hi[a];
    }
    out.maxDepth = std::max(out.maxDepth, depth);
    if (t.leaf()) {
        n.leftOrPrim = t.begin;
        n.packed = (t.count << 2) | t.axis;
        out.leafCount++;
        return;
    }
    const uint32_t left = static_cast<uint32_t>(out.nodes.size());
    out.nodes.emplace_back();
    out.nodes.emplace_back();
    out.nodes[slot].leftOrPrim = left;  // re-index: emplace_back may reallocate
    out.nodes[slot].packed = t.axis;
    emit(*t.kid[0], left, depth + 1, out);
    emit(*t.kid[1], left + 1, depth + 1, out);
}

} // namespace

BvhBuild buildBvh(const std::vector<Box3>& bounds, uint32_t maxLeaf, int threads) {
    BvhBuild out;
    const uint32_t count = static_cast<uint32_t>(bounds.size());
    if (count == 0) return out;
    if (threads <= 0) threads = int(std::max(1u, std::thread::hardware_concurrency()));

    std::vector<Ref> refs(count);
    for (uint32_t i = 0; i < count; ++i) {
        refs[i].box = bounds[i];
        for (int a = 0; a < 3; ++a) refs[i].c[a] = 0.5f * (bounds[i].lo[a] + bounds[i].hi[a]);
        refs[i].index = i;
    }
    int parDepth = 0;
    while ((1 << parDepth) < threads && parDepth < 8) ++parDepth;
    Ctx ctx{ refs, std::max(1u, maxLeaf), parDepth + 1 };
    TNode root;
    build(ctx, root, 0, count, 0);

    out.nodes.reserve(size_t(2) * count);
    out.nodes.emplace_back();
    emit(root, 0, 0, out);
    out.primOrder.resize(count);
    for (uint32_t i = 0; i < count; ++i) out.primOrder[i] = refs[i].index;
    out.rootBounds = root.box;
    return out;
}

} // namespace mrt
