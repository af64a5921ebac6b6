// Binned-SAH BVH construction on the GPU with the reference builder's output,
// node for node (core/src/accel/Bvh.cpp:41-156): 32 bins over the widest
// centroid axis, traversal cost 1.0, intersection cost 1.5, the partition
// order of libstdc++'s std::partition, child pairs allocated in depth-first
// order.
//
// std::partition (the bidirectional-iterator algorithm libstdc++ uses for
// random-access ranges) leaves the elements that already sit on their side
// where they are and swaps the i-th left-side element that fails the
// predicate (counted from the left) with the i-th right-side element that
// passes it (counted from the right); the split position is the number of
// passing elements, known from the bin counts.  Both orders come from
// prefix counts, so the partition runs in parallel and still produces the
// serial permutation.  Every floating-point expression is the builder's,
// evaluated in the same order without contraction, so boxes, bins and SAH
// costs agree bit for bit.
//
// Two phases:
//   large nodes (more than kSmallMax primitives): one level at a time, each
//     node's range split into chunks of kChunk slots; per-chunk reductions
//     and bin histograms are combined with integer atomics on the
//     order-preserving integer image of the floats (min / max / add of
//     counts: exact whatever the order);
//   small nodes: one workgroup builds the whole subtree of a node in LDS,
//     depth first, with its nodes numbered locally in the reference's pair
//     order; the host then places every subtree at its final index.
// The median split the reference falls back to when SAH prefers an
// oversized leaf (std::nth_element, introselect) is not restated: a node
// that needs it sets a flag and the caller builds on the CPU instead.
#pragma clang fp contract(off)
#include "bvh_gpu.h"

#include <hip/hip_runtime.h>

namespace mrt {
namespace dev {

namespace {

constexpr uint32_t kBins = 32;
constexpr float kTraversalCost = 1.0f;
constexpr float kIntersectCost = 1.5f;
constexpr float kBinScale = float(kBins) * (1.0f - 1e-6f);
constexpr float kFltMax = 3.402823466e+38f;

// Order-preserving map of floats (no NaN) to signed integers and back.
__device__ __forceinline__ int fkey(float f) {
    const int i = __float_as_int(f);
    return i >= 0 ? i : (i ^ 0x7FFFFFFF);
}
__device__ __forceinline__ float funkey(int k) { return __int_as_float(k >= 0 ? k : (k ^ 0x7FFFFFFF)); }

struct Box {
    float lo[3], hi[3];
};
__device__ __forceinline__ void boxEmpty(Box& b) {
    for (int a = 0; a < 3; ++a) {
        b.lo[a] = kFltMax;
        b.hi[a] = -kFltMax;
    }
}
// Box3::grow of bvh_build.h
__device__ __forceinline__ void growBox(Box& b, const float lo[3], const float hi[3]) {
    for (int a = 0; a < 3; ++a) {
        b.lo[a] = (lo[a] < b.lo[a]) ? lo[a] : b.lo[a];
        b.hi[a] = (b.hi[a] < hi[a]) ? hi[a] : b.hi[a];
    }
}
__device__ __forceinline__ float boxArea(const Box& b) {
    if (!(b.lo[0] <= b.hi[0])) return 0.0f;
    const float ex = b.hi[0] - b.lo[0], ey = b.hi[1] - b.lo[1], ez = b.hi[2] - b.lo[2];
    const float s = (ex * ey + ey * ez) + ez * ex;
    return 2.0f * s;
}

__device__ __forceinline__ void loadPrim(const float* __restrict__ bounds, uint32_t p, float lo[3], float hi[3]) {
    const float* b = bounds + size_t(p) * 6;
    for (int a = 0; a < 3; ++a) {
        lo[a] = b[a];
        hi[a] = b[3 + a];
    }
}
__device__ __forceinline__ float centroid(const float* __restrict__ bounds, uint32_t p, int axis) {
    const float* b = bounds + size_t(p) * 6;
    return 0.5f * (b[axis] + b[3 + axis]);
}

// Node decision from its box, centroid box and count.
struct Decision {
    int axis;
    bool leaf;
    float k, origin;
};
__device__ __forceinline__ Decision decide(const Box& cb, uint32_t count, uint32_t maxLeaf) {
    const float ce[3] = { cb.hi[0] - cb.lo[0], cb.hi[1] - cb.lo[1], cb.hi[2] - cb.lo[2] };
    Decision d;
    d.axis = 0;
    if (ce[1] > ce[0]) d.axis = 1;
    if (ce[2] > ce[d.axis]) d.axis = 2;
    d.leaf = count <= maxLeaf || ce[d.axis] < 1e-8f;
    d.k = d.leaf ? 0.0f : kBinScale / ce[d.axis];
    d.origin = cb.lo[d.axis];
    return d;
}
// The builder's bin index; it is below kBins for every finite centroid (the
// clamp only keeps a NaN centroid inside the bin array).
__device__ __forceinline__ uint32_t binOf(const Decision& d, float c) {
    return min(static_cast<uint32_t>(d.k * (c - d.origin)), kBins - 1);
}

// SAH sweep over the bins (one thread); returns the best split bin or -1.
__device__ int sahSweep(const Box* binBox, const uint32_t* binCount, const Box& bounds, uint32_t count) {
    float rightArea[kBins];
    uint32_t rightCount[kBins];
    {
        Box acc;
        boxEmpty(acc);
        uint32_t cnt = 0;
        for (int i = int(kBins) - 1; i > 0; --i) {
            growBox(acc, binBox[i].lo, binBox[i].hi);
            cnt += binCount[i];
            rightArea[i] = boxArea(acc);
            rightCount[i] = cnt;
        }
    }
    float bestCost = kIntersectCost * float(count);
    int bestSplit = -1;
    Box acc;
    boxEmpty(acc);
    uint32_t cnt = 0;
    const float invRoot = 1.0f / fmaxf(boxArea(bounds), 1e-12f);
    const float scale = kIntersectCost * invRoot;
    for (uint32_t i = 0; i + 1 < kBins; ++i) {
        growBox(acc, binBox[i].lo, binBox[i].hi);
        cnt += binCount[i];
        if (cnt == 0 || rightCount[i + 1] == 0) continue;
        const float l = boxArea(acc) * float(cnt);
        const float r = rightArea[i + 1] * float(rightCount[i + 1]);
        const float cost = kTraversalCost + scale * (l + r);
        if (cost < bestCost) {
            bestCost = cost;
            bestSplit = int(i);
        }
    }
    return bestSplit;
}

// ------------------------------------------------------------------ large nodes --
// Per-node accumulators (integer images): box lo/hi, centroid box lo/hi,
// then per bin box lo/hi and count.
constexpr int kAccBox = 0, kAccCBox = 6, kAccBins = 12;
constexpr int kAccPerBin = 7;
constexpr int kAccWords = kAccBins + int(kBins) * kAccPerBin;  // 236

__device__ __forceinline__ void accMin(int* w, float f) { atomicMin(w, fkey(f)); }
__device__ __forceinline__ void accMax(int* w, float f) { atomicMax(w, fkey(f)); }

__global__ __launch_bounds__(256) void kLargeClear(int* __restrict__ acc, uint32_t nodes) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nodes * kAccWords) return;
    const int w = int(i % kAccWords);
    int v;
    if (w < kAccBins) v = (w % 6) < 3 ? fkey(kFltMax) : fkey(-kFltMax);
    else {
        const int j = (w - kAccBins) % kAccPerBin;
        v = j == 6 ? 0 : (j < 3 ? fkey(kFltMax) : fkey(-kFltMax));
    }
    acc[i] = v;
}

// Box and centroid box of each chunk, combined per node.
__global__ __launch_bounds__(256) void kLargeBounds(const float* __restrict__ bounds, const uint32_t* __restrict__ idx,
                                                    const LargeChunk* __restrict__ chunks, int* __restrict__ acc) {
    __shared__ int red[12][256];
    const LargeChunk c = chunks[blockIdx.x];
    Box b, cb;
    boxEmpty(b);
    boxEmpty(cb);
    for (uint32_t s = c.begin + threadIdx.x; s < c.end; s += 256) {
        float lo[3], hi[3], cc[3];
        loadPrim(bounds, idx[s], lo, hi);
        for (int a = 0; a < 3; ++a) cc[a] = 0.5f * (lo[a] + hi[a]);
        growBox(b, lo, hi);
        growBox(cb, cc, cc);
    }
    for (int a = 0; a < 3; ++a) {
        red[a][threadIdx.x] = fkey(b.lo[a]);
        red[3 + a][threadIdx.x] = fkey(b.hi[a]);
        red[6 + a][threadIdx.x] = fkey(cb.lo[a]);
        red[9 + a][threadIdx.x] = fkey(cb.hi[a]);
    }
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h)
            for (int v = 0; v < 12; ++v) {
                const int o = red[v][threadIdx.x + h];
                const bool isMin = (v % 6) < 3;
                red[v][threadIdx.x] = isMin ? min(red[v][threadIdx.x], o) : max(red[v][threadIdx.x], o);
            }
        __syncthreads();
    }
    if (threadIdx.x < 12) {
        int* w = acc + size_t(c.node) * kAccWords + threadIdx.x;  // box 0..5, centroid box 6..11
        if ((threadIdx.x % 6) < 3) atomicMin(w, red[threadIdx.x][0]);
        else atomicMax(w, red[threadIdx.x][0]);
    }
}

__device__ __forceinline__ void accBoxes(const int* a, Box& b, Box& cb) {
    for (int k = 0; k < 3; ++k) {
        b.lo[k] = funkey(a[kAccBox + k]);
        b.hi[k] = funkey(a[kAccBox + 3 + k]);
        cb.lo[k] = funkey(a[kAccCBox + k]);
        cb.hi[k] = funkey(a[kAccCBox + 3 + k]);
    }
}

__global__ __launch_bounds__(256) void kLargeDecide(LargeNode* __restrict__ nodes, uint32_t count,
                                                    const int* __restrict__ acc, uint32_t maxLeaf) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= count) return;
    LargeNode& n = nodes[i];
    Box b, cb;
    accBoxes(acc + size_t(i) * kAccWords, b, cb);
    const Decision d = decide(cb, n.count, maxLeaf);
    for (int a = 0; a < 3; ++a) {
        n.lo[a] = b.lo[a];
        n.hi[a] = b.hi[a];
    }
    n.axis = uint32_t(d.axis);
    n.k = d.k;
    n.origin = d.origin;
    n.split = d.leaf ? kSplitLeaf : kSplitPending;
}

// Bin histograms of each chunk (LDS), combined per node.
__global__ __launch_bounds__(256) void kLargeBins(const float* __restrict__ bounds, const uint32_t* __restrict__ idx,
                                                  const LargeChunk* __restrict__ chunks,
                                                  const LargeNode* __restrict__ nodes, int* __restrict__ acc) {
    __shared__ int bins[kBins * kAccPerBin];
    const LargeChunk c = chunks[blockIdx.x];
    const LargeNode& n = nodes[c.node];
    if (n.split == kSplitLeaf) return;  // uniform per block
    for (uint32_t i = threadIdx.x; i < kBins * kAccPerBin; i += 256) {
        const uint32_t j = i % kAccPerBin;
        bins[i] = j == 6 ? 0 : (j < 3 ? fkey(kFltMax) : fkey(-kFltMax));
    }
    __syncthreads();
    Decision d;
    d.axis = int(n.axis);
    d.k = n.k;
    d.origin = n.origin;
    for (uint32_t s = c.begin + threadIdx.x; s < c.end; s += 256) {
        float lo[3], hi[3];
        loadPrim(bounds, idx[s], lo, hi);
        const uint32_t bi = binOf(d, 0.5f * (lo[d.axis] + hi[d.axis]));
        int* w = bins + bi * kAccPerBin;
        for (int a = 0; a < 3; ++a) {
            atomicMin(w + a, fkey(lo[a]));
            atomicMax(w + 3 + a, fkey(hi[a]));
        }
        atomicAdd(w + 6, 1);
    }
    __syncthreads();
    int* dst = acc + size_t(c.node) * kAccWords + kAccBins;
    for (uint32_t i = threadIdx.x; i < kBins * kAccPerBin; i += 256) {
        const uint32_t j = i % kAccPerBin;
        if (j == 6) {
            if (bins[i]) atomicAdd(dst + i, bins[i]);
        } else if (j < 3) {
            atomicMin(dst + i, bins[i]);
        } else {
            atomicMax(dst + i, bins[i]);
        }
    }
}

__global__ __launch_bounds__(256) void kLargeSah(LargeNode* __restrict__ nodes, uint32_t count,
                                                 const int* __restrict__ acc, uint32_t maxLeaf,
                                                 uint32_t* __restrict__ flags) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= count) return;
    LargeNode& n = nodes[i];
    if (n.split == kSplitLeaf) return;
    const int* a = acc + size_t(i) * kAccWords + kAccBins;
    Box binBox[kBins];
    uint32_t binCount[kBins];
    for (uint32_t b = 0; b < kBins; ++b) {
        for (int k = 0; k < 3; ++k) {
            binBox[b].lo[k] = funkey(a[b * kAccPerBin + k]);
            binBox[b].hi[k] = funkey(a[b * kAccPerBin + 3 + k]);
        }
        binCount[b] = uint32_t(a[b * kAccPerBin + 6]);
    }
    Box bounds;
    for (int k = 0; k < 3; ++k) {
        bounds.lo[k] = n.lo[k];
        bounds.hi[k] = n.hi[k];
    }
    const int split = sahSweep(binBox, binCount, bounds, n.count);
    if (split < 0) {
        if (n.count <= 2 * maxLeaf) {
            n.split = kSplitLeaf;
            return;
        }
        atomicOr(flags, kBuildNeedsMedian);  // the reference's nth_element split
        n.split = kSplitLeaf;
        return;
    }
    uint32_t m = 0;
    for (int b = 0; b <= split; ++b) m += binCount[b];
    n.split = uint32_t(split);
    n.mid = m;  // elements passing the predicate = the partition point (relative)
}

// Prefix counts of the slots that change side: left-side elements failing
// the predicate and right-side elements passing it.
__device__ __forceinline__ bool misplaced(const LargeNode& n, uint32_t rel, float c, const Decision& d) {
    const bool pass = binOf(d, c) <= n.split;
    return rel < n.mid ? !pass : pass;
}

__global__ __launch_bounds__(256) void kLargeCount(const float* __restrict__ bounds, const uint32_t* __restrict__ idx,
                                                   const LargeChunk* __restrict__ chunks,
                                                   const LargeNode* __restrict__ nodes, uint32_t* __restrict__ counts) {
    __shared__ uint32_t part[2][256];
    const LargeChunk c = chunks[blockIdx.x];
    const LargeNode& n = nodes[c.node];
    uint32_t left = 0, right = 0;
    if (n.split < kBins && n.mid > 0 && n.mid < n.count) {
        Decision d;
        d.axis = int(n.axis);
        d.k = n.k;
        d.origin = n.origin;
        for (uint32_t s = c.begin + threadIdx.x; s < c.end; s += 256) {
            const uint32_t rel = s - n.begin;
            if (misplaced(n, rel, centroid(bounds, idx[s], d.axis), d)) {
                if (rel < n.mid) ++left;
                else ++right;
            }
        }
    }
    part[0][threadIdx.x] = left;
    part[1][threadIdx.x] = right;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            part[0][threadIdx.x] += part[0][threadIdx.x + h];
            part[1][threadIdx.x] += part[1][threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        counts[2 * blockIdx.x] = part[0][0];
        counts[2 * blockIdx.x + 1] = part[1][0];
    }
}

// Exclusive scans of the per-chunk counts within each node (chunks of a node
// are consecutive), one thread per node; totals[node] = right-side slots
// that change side (as many as left-side ones).
__global__ __launch_bounds__(256) void kLargeScan(const LargeNode* __restrict__ nodes, uint32_t nodeCount,
                                                  uint32_t* __restrict__ counts, uint32_t* __restrict__ totals) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nodeCount) return;
    const LargeNode& n = nodes[i];
    uint32_t l = 0, r = 0;
    for (uint32_t c = n.chunk0; c < n.chunk0 + n.chunks; ++c) {
        const uint32_t cl = counts[2 * c], cr = counts[2 * c + 1];
        counts[2 * c] = l;
        counts[2 * c + 1] = r;
        l += cl;
        r += cr;
    }
    totals[i] = r;
    totals[nodeCount + i] = l;
}

// Rank of each misplaced slot within its chunk (block scan over the chunk
// in rounds of 256 slots), plus the chunk's offset.
template <typename F>
__device__ __forceinline__ void chunkRanks(const LargeChunk& c, F&& visit) {
    __shared__ uint32_t scan[2][256];
    __shared__ uint32_t carry[2];
    if (threadIdx.x < 2) carry[threadIdx.x] = 0;
    __syncthreads();
    for (uint32_t s0 = c.begin; s0 < c.end; s0 += 256) {
        const uint32_t s = s0 + threadIdx.x;
        int side = -1;  // 0: left misplaced, 1: right misplaced
        visit(s, side, 0u, false);
        scan[0][threadIdx.x] = side == 0;
        scan[1][threadIdx.x] = side == 1;
        __syncthreads();
        for (int o = 1; o < 256; o <<= 1) {  // inclusive Hillis-Steele scan
            uint32_t a = 0, b = 0;
            if (threadIdx.x >= uint32_t(o)) {
                a = scan[0][threadIdx.x - o];
                b = scan[1][threadIdx.x - o];
            }
            __syncthreads();
            scan[0][threadIdx.x] += a;
            scan[1][threadIdx.x] += b;
            __syncthreads();
        }
        if (side >= 0) visit(s, side, carry[side] + scan[side][threadIdx.x] - 1, true);
        __syncthreads();
        if (threadIdx.x == 0) {
            carry[0] += scan[0][255];
            carry[1] += scan[1][255];
        }
        __syncthreads();
    }
}

// Right-side misplaced slots record their positions by rank.
__global__ __launch_bounds__(256) void kLargeRight(const float* __restrict__ bounds, const uint32_t* __restrict__ idx,
                                                   const LargeChunk* __restrict__ chunks,
                                                   const LargeNode* __restrict__ nodes,
                                                   const uint32_t* __restrict__ counts, uint32_t* __restrict__ pos) {
    const LargeChunk c = chunks[blockIdx.x];
    const LargeNode& n = nodes[c.node];
    if (!(n.split < kBins && n.mid > 0 && n.mid < n.count)) return;  // uniform per block
    Decision d;
    d.axis = int(n.axis);
    d.k = n.k;
    d.origin = n.origin;
    const uint32_t off = counts[2 * blockIdx.x + 1];
    chunkRanks(c, [&](uint32_t s, int& side, uint32_t rank, bool emit) {
        if (!emit) {
            if (s < c.end) {
                const uint32_t rel = s - n.begin;
                if (rel >= n.mid && misplaced(n, rel, centroid(bounds, idx[s], d.axis), d)) side = 1;
            }
            return;
        }
        pos[n.begin + off + rank] = s;
    });
}

// Left-side misplaced slot i (from the left) swaps with the right-side
// misplaced slot i counted from the right.
__global__ __launch_bounds__(256) void kLargeSwap(const float* __restrict__ bounds, uint32_t* __restrict__ idx,
                                                  const LargeChunk* __restrict__ chunks,
                                                  const LargeNode* __restrict__ nodes,
                                                  const uint32_t* __restrict__ counts, const uint32_t* __restrict__ pos,
                                                  const uint32_t* __restrict__ misplacedTotal) {
    const LargeChunk c = chunks[blockIdx.x];
    const LargeNode& n = nodes[c.node];
    if (!(n.split < kBins && n.mid > 0 && n.mid < n.count)) return;
    Decision d;
    d.axis = int(n.axis);
    d.k = n.k;
    d.origin = n.origin;
    const uint32_t off = counts[2 * blockIdx.x];
    const uint32_t total = misplacedTotal[c.node];
    chunkRanks(c, [&](uint32_t s, int& side, uint32_t rank, bool emit) {
        if (!emit) {
            if (s < c.end) {
                const uint32_t rel = s - n.begin;
                if (rel < n.mid && misplaced(n, rel, centroid(bounds, idx[s], d.axis), d)) side = 0;
            }
            return;
        }
        if (off + rank >= total) return;  // cannot happen: both sides hold `total` such slots
        const uint32_t q = pos[n.begin + (total - 1 - (off + rank))];
        const uint32_t a = idx[s], b = idx[q];
        idx[s] = b;
        idx[q] = a;
    });
}

// ------------------------------------------------------------------ small nodes --
constexpr uint32_t kSmallThreads = 256;

struct SmallStackEntry {
    uint32_t begin, count, node, depth;
};

__device__ __forceinline__ void blockMinMax(Box& b, Box& cb, int (*red)[kSmallThreads]) {
    for (int a = 0; a < 3; ++a) {
        red[a][threadIdx.x] = fkey(b.lo[a]);
        red[3 + a][threadIdx.x] = fkey(b.hi[a]);
        red[6 + a][threadIdx.x] = fkey(cb.lo[a]);
        red[9 + a][threadIdx.x] = fkey(cb.hi[a]);
    }
    __syncthreads();
    for (int h = kSmallThreads / 2; h > 0; h >>= 1) {
        if (threadIdx.x < uint32_t(h))
            for (int v = 0; v < 12; ++v) {
                const int o = red[v][threadIdx.x + h];
                red[v][threadIdx.x] = (v % 6) < 3 ? min(red[v][threadIdx.x], o) : max(red[v][threadIdx.x], o);
            }
        __syncthreads();
    }
    for (int a = 0; a < 3; ++a) {
        b.lo[a] = funkey(red[a][0]);
        b.hi[a] = funkey(red[3 + a][0]);
        cb.lo[a] = funkey(red[6 + a][0]);
        cb.hi[a] = funkey(red[9 + a][0]);
    }
    __syncthreads();
}

// One workgroup per small node: the node's whole subtree, depth first, in
// LDS.  Subtree nodes are written to out + task.outBase with local indices
// (root 0, child pairs allocated in depth-first order from 1); leaves hold
// absolute slot indices.
__global__ __launch_bounds__(kSmallThreads) void kBuildSmall(const float* __restrict__ bounds,
                                                             uint32_t* __restrict__ idx,
                                                             const SmallTask* __restrict__ tasks,
                                                             OutNode* __restrict__ out,
                                                             SmallResult* __restrict__ results, uint32_t maxLeaf,
                                                             uint32_t* __restrict__ flags) {
    __shared__ uint32_t sIdx[kSmallMax];
    __shared__ float sC[3][kSmallMax];
    __shared__ uint32_t sPos[kSmallMax];
    __shared__ SmallStackEntry stack[kSmallMax];
    __shared__ int red[12][kSmallThreads];
    __shared__ int bins[kBins * kAccPerBin];
    __shared__ uint32_t scan[kSmallThreads];
    __shared__ uint32_t shared[8];  // sp, alloc, split, mid, begin, count, node, depth
    __shared__ float sDec[2];       // k, origin
    __shared__ uint32_t sAxis, sCarry, sTotal, sMaxDepth, sLeaves;

    const SmallTask t = tasks[blockIdx.x];
    for (uint32_t i = threadIdx.x; i < t.count; i += kSmallThreads) {
        const uint32_t p = idx[t.begin + i];
        sIdx[i] = p;
        for (int a = 0; a < 3; ++a) sC[a][i] = centroid(bounds, p, a);
    }
    if (threadIdx.x == 0) {
        stack[0] = SmallStackEntry{ 0u, t.count, 0u, t.depth };
        shared[0] = 1;  // stack size
        shared[1] = 1;  // next free local node (root is 0)
        sMaxDepth = 0;
        sLeaves = 0;
    }
    __syncthreads();

    while (shared[0] > 0) {
        if (threadIdx.x == 0) {
            const SmallStackEntry e = stack[--shared[0]];
            shared[4] = e.begin;
            shared[5] = e.count;
            shared[6] = e.node;
            shared[7] = e.depth;
        }
        __syncthreads();
        const uint32_t b0 = shared[4], cnt = shared[5], node = shared[6], depth = shared[7];
        Box b, cb;
        boxEmpty(b);
        boxEmpty(cb);
        for (uint32_t i = threadIdx.x; i < cnt; i += kSmallThreads) {
            float lo[3], hi[3], cc[3];
            loadPrim(bounds, sIdx[b0 + i], lo, hi);
            for (int a = 0; a < 3; ++a) cc[a] = sC[a][b0 + i];
            growBox(b, lo, hi);
            growBox(cb, cc, cc);
        }
        blockMinMax(b, cb, red);
        const Decision d0 = decide(cb, cnt, maxLeaf);
        bool leaf = d0.leaf;
        if (!leaf) {
            for (uint32_t i = threadIdx.x; i < kBins * kAccPerBin; i += kSmallThreads) {
                const uint32_t j = i % kAccPerBin;
                bins[i] = j == 6 ? 0 : (j < 3 ? fkey(kFltMax) : fkey(-kFltMax));
            }
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < cnt; i += kSmallThreads) {
                float lo[3], hi[3];
                loadPrim(bounds, sIdx[b0 + i], lo, hi);
                const uint32_t bi = binOf(d0, sC[d0.axis][b0 + i]);
                int* w = bins + bi * kAccPerBin;
                for (int a = 0; a < 3; ++a) {
                    atomicMin(w + a, fkey(lo[a]));
                    atomicMax(w + 3 + a, fkey(hi[a]));
                }
                atomicAdd(w + 6, 1);
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                Box binBox[kBins];
                uint32_t binCount[kBins];
                for (uint32_t k = 0; k < kBins; ++k) {
                    for (int a = 0; a < 3; ++a) {
                        binBox[k].lo[a] = funkey(bins[k * kAccPerBin + a]);
                        binBox[k].hi[a] = funkey(bins[k * kAccPerBin + 3 + a]);
                    }
                    binCount[k] = uint32_t(bins[k * kAccPerBin + 6]);
                }
                const int split = sahSweep(binBox, binCount, b, cnt);
                uint32_t m = 0;
                if (split >= 0)
                    for (int k = 0; k <= split; ++k) m += binCount[k];
                else if (cnt > 2 * maxLeaf)
                    atomicOr(flags, kBuildNeedsMedian);
                shared[2] = split < 0 ? kSplitLeaf : uint32_t(split);
                shared[3] = m;
            }
            __syncthreads();
            leaf = shared[2] == kSplitLeaf;
        }
        if (leaf) {
            if (threadIdx.x == 0) {
                OutNode& o = out[t.outBase + node];
                for (int a = 0; a < 3; ++a) {
                    o.lo[a] = b.lo[a];
                    o.hi[a] = b.hi[a];
                }
                o.leftOrPrim = t.begin + b0;
                o.packed = (cnt << 2) | uint32_t(d0.axis);
                sMaxDepth = max(sMaxDepth, depth);
                ++sLeaves;
            }
            __syncthreads();
            continue;
        }
        const uint32_t split = shared[2];
        uint32_t mid = shared[3];
        if (mid > 0 && mid < cnt) {
            // right-side slots passing the predicate, ranked from the left
            if (threadIdx.x == 0) sCarry = 0;
            __syncthreads();
            for (int pass = 0; pass < 2; ++pass) {
                if (threadIdx.x == 0) sCarry = 0;
                __syncthreads();
                for (uint32_t r0 = 0; r0 < cnt; r0 += kSmallThreads) {
                    const uint32_t i = r0 + threadIdx.x;
                    bool mis = false;
                    if (i < cnt) {
                        const bool passP = binOf(d0, sC[d0.axis][b0 + i]) <= split;
                        mis = pass == 0 ? (i >= mid && passP) : (i < mid && !passP);
                    }
                    scan[threadIdx.x] = mis ? 1u : 0u;
                    __syncthreads();
                    for (int o = 1; o < int(kSmallThreads); o <<= 1) {
                        const uint32_t v = threadIdx.x >= uint32_t(o) ? scan[threadIdx.x - o] : 0u;
                        __syncthreads();
                        scan[threadIdx.x] += v;
                        __syncthreads();
                    }
                    const uint32_t rank = sCarry + scan[threadIdx.x] - 1;
                    uint32_t a = 0, c = 0;
                    uint32_t q = 0;
                    if (mis && pass == 0) sPos[rank] = i;
                    if (mis && pass == 1 && rank >= sTotal) mis = false;  // cannot happen (equal counts)
                    if (mis && pass == 1) {
                        q = sPos[sTotal - 1 - rank];
                        a = sIdx[b0 + i];
                        c = sIdx[b0 + q];
                    }
                    float ca[3], cq[3];
                    if (mis && pass == 1)
                        for (int k = 0; k < 3; ++k) {
                            ca[k] = sC[k][b0 + i];
                            cq[k] = sC[k][b0 + q];
                        }
                    __syncthreads();
                    if (mis && pass == 1) {
                        sIdx[b0 + i] = c;
                        sIdx[b0 + q] = a;
                        for (int k = 0; k < 3; ++k) {
                            sC[k][b0 + i] = cq[k];
                            sC[k][b0 + q] = ca[k];
                        }
                    }
                    __syncthreads();
                    if (threadIdx.x == 0) sCarry += scan[kSmallThreads - 1];
                    __syncthreads();
                }
                if (threadIdx.x == 0 && pass == 0) sTotal = sCarry;
                __syncthreads();
            }
        } else {
            mid = cnt / 2;
        }
        if (threadIdx.x == 0) {
            const uint32_t left = shared[1];
            shared[1] = left + 2;
            OutNode& o = out[t.outBase + node];
            for (int a = 0; a < 3; ++a) {
                o.lo[a] = b.lo[a];
                o.hi[a] = b.hi[a];
            }
            o.leftOrPrim = left;  // local index; placed by kPlaceSmall
            o.packed = uint32_t(d0.axis);
            uint32_t sp = shared[0];
            stack[sp++] = SmallStackEntry{ b0 + mid, cnt - mid, left + 1, depth + 1 };
            stack[sp++] = SmallStackEntry{ b0, mid, left, depth + 1 };
            shared[0] = sp;
        }
        __syncthreads();
    }
    for (uint32_t i = threadIdx.x; i < t.count; i += kSmallThreads) idx[t.begin + i] = sIdx[i];
    if (threadIdx.x == 0) {
        results[blockIdx.x].nodes = shared[1];
        results[blockIdx.x].maxDepth = sMaxDepth;
        results[blockIdx.x].leaves = sLeaves;
    }
}

// Copies each subtree to its final place: local node j >= 1 goes to
// base + j - 1 and the root to rootIndex; inner nodes' local child
// indices become base + j - 1.  One block per subtree.
__global__ __launch_bounds__(256) void kPlaceSmall(const OutNode* __restrict__ tmp, const SmallTask* __restrict__ tasks,
                                                   const SmallResult* __restrict__ results,
                                                   const SmallPlace* __restrict__ place, OutNode* __restrict__ dst) {
    const SmallTask t = tasks[blockIdx.x];
    const SmallPlace p = place[blockIdx.x];
    const uint32_t n = results[blockIdx.x].nodes;
    for (uint32_t j = threadIdx.x; j < n; j += 256) {
        OutNode o = tmp[t.outBase + j];
        if ((o.packed >> 2) == 0) o.leftOrPrim = p.base + o.leftOrPrim - 1;
        dst[j == 0 ? p.root : p.base + j - 1] = o;
    }
}

__global__ __launch_bounds__(256) void kPlaceLarge(const OutNode* __restrict__ src, const uint32_t* __restrict__ where,
                                                   uint32_t count, OutNode* __restrict__ dst) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < count) dst[where[i]] = src[i];
}

__global__ __launch_bounds__(256) void kIota(uint32_t* __restrict__ idx, uint32_t n) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) idx[i] = i;
}

__global__ void kClearFlags(uint32_t* __restrict__ flags) { *flags = 0; }

} // namespace

uint32_t gridFor(uint32_t n) { return (n + 255) / 256; }

hipError_t launchBvhIota(uint32_t* idx, uint32_t n, uint32_t* flags, hipStream_t s) {
    hipLaunchKernelGGL(kClearFlags, dim3(1), dim3(1), 0, s, flags);
    if (n) hipLaunchKernelGGL(kIota, dim3(gridFor(n)), dim3(256), 0, s, idx, n);
    return hipGetLastError();
}

hipError_t launchBvhLargeLevel(const LargeLevelArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(kLargeClear, dim3(gridFor(a.nodeCount * kAccWords)), dim3(256), 0, s, a.acc, a.nodeCount);
    hipLaunchKernelGGL(kLargeBounds, dim3(a.chunkCount), dim3(256), 0, s, a.bounds, a.idx, a.chunks, a.acc);
    hipLaunchKernelGGL(kLargeDecide, dim3(gridFor(a.nodeCount)), dim3(256), 0, s, a.nodes, a.nodeCount, a.acc,
                       a.maxLeaf);
    hipLaunchKernelGGL(kLargeBins, dim3(a.chunkCount), dim3(256), 0, s, a.bounds, a.idx, a.chunks, a.nodes, a.acc);
    hipLaunchKernelGGL(kLargeSah, dim3(gridFor(a.nodeCount)), dim3(256), 0, s, a.nodes, a.nodeCount, a.acc,
                       a.maxLeaf, a.flags);
    hipLaunchKernelGGL(kLargeCount, dim3(a.chunkCount), dim3(256), 0, s, a.bounds, a.idx, a.chunks, a.nodes,
                       a.counts);
    hipLaunchKernelGGL(kLargeScan, dim3(gridFor(a.nodeCount)), dim3(256), 0, s, a.nodes, a.nodeCount, a.counts,
                       a.totals);
    hipLaunchKernelGGL(kLargeRight, dim3(a.chunkCount), dim3(256), 0, s, a.bounds, a.idx, a.chunks, a.nodes,
                       a.counts, a.pos);
    hipLaunchKernelGGL(kLargeSwap, dim3(a.chunkCount), dim3(256), 0, s, a.bounds, a.idx, a.chunks, a.nodes, a.counts,
                       a.pos, a.totals);
    return hipGetLastError();
}

hipError_t launchBvhSmall(const SmallArgs& a, hipStream_t s) {
    if (a.taskCount == 0) return hipSuccess;
    hipLaunchKernelGGL(kBuildSmall, dim3(a.taskCount), dim3(kSmallThreads), 0, s, a.bounds, a.idx, a.tasks, a.tmp,
                       a.results, a.maxLeaf, a.flags);
    return hipGetLastError();
}

hipError_t launchBvhPlace(const PlaceArgs& a, hipStream_t s) {
    if (a.largeCount)
        hipLaunchKernelGGL(kPlaceLarge, dim3(gridFor(a.largeCount)), dim3(256), 0, s, a.large, a.largeWhere,
                           a.largeCount, a.dst);
    if (a.taskCount)
        hipLaunchKernelGGL(kPlaceSmall, dim3(a.taskCount), dim3(256), 0, s, a.tmp, a.tasks, a.results, a.place,
                           a.dst);
    return hipGetLastError();
}

} // namespace dev
} // namespace mrt
