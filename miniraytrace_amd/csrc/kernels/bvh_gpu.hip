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
//   small nodes: one wavefront builds the whole subtree of a node in LDS,
//     depth first, with its nodes numbered locally in the reference's pair
//     order; the host then places every subtree at its final index.
// A box bound where -0.0 and +0.0 tie takes the sign of the first primitive
// in slot order holding a zero there, as Box3::grow's serial loop does.
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
// The same map with -0.0 taken as +0.0: a min / max over these keys finds the
// numeric extreme; which zero the builder keeps is settled apart (zeroCode).
__device__ __forceinline__ int fkeyNum(float f) { return __float_as_uint(f) == 0x80000000u ? 0 : fkey(f); }
__device__ __forceinline__ bool isZero(float f) { return (__float_as_uint(f) & 0x7FFFFFFFu) == 0; }

// Box3::grow keeps the value it already holds on a tie, so when a box bound is
// zero and its node holds both -0.0 and +0.0 there, the sign is that of the
// first primitive in slot order whose bound is zero.  zeroCode ranks such
// bounds by slot (twice the slot, plus the sign bit); the least code wins.
constexpr uint32_t kNoZero = 0xFFFFFFFFu;
__device__ __forceinline__ uint32_t zeroCode(float f, uint32_t slot) {
    return isZero(f) ? 2 * slot + (__float_as_uint(f) >> 31) : kNoZero;
}
__device__ __forceinline__ float zeroFromCode(float v, uint32_t code) {
    return (v == 0.0f && code != kNoZero) ? ((code & 1) ? -0.0f : 0.0f) : v;
}

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
// v[a] by selects: a run-time index into a register array would put the
// array in scratch memory.
__device__ __forceinline__ float pick(const float v[3], int a) { return a == 0 ? v[0] : (a == 1 ? v[1] : v[2]); }
__device__ __forceinline__ Decision decide(const Box& cb, uint32_t count, uint32_t maxLeaf) {
    const float ce[3] = { cb.hi[0] - cb.lo[0], cb.hi[1] - cb.lo[1], cb.hi[2] - cb.lo[2] };
    Decision d;
    d.axis = 0;
    if (ce[1] > ce[0]) d.axis = 1;
    if (ce[2] > pick(ce, d.axis)) d.axis = 2;
    const float e = pick(ce, d.axis);
    d.leaf = count <= maxLeaf || e < 1e-8f;
    d.k = d.leaf ? 0.0f : kBinScale / e;
    d.origin = pick(cb.lo, d.axis);
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
// the zero codes of the box bounds, then per bin box lo/hi and count.
constexpr int kAccBox = 0, kAccCBox = 6, kAccZero = 12, kAccBins = 18;
constexpr int kAccPerBin = 7;
constexpr int kAccWords = kAccBins + int(kBins) * kAccPerBin;
static_assert(kAccWords == int(kLargeAccWords), "accumulator words per large node");

__device__ __forceinline__ void accMin(int* w, float f) { atomicMin(w, fkey(f)); }
__device__ __forceinline__ void accMax(int* w, float f) { atomicMax(w, fkey(f)); }

__global__ __launch_bounds__(256) void kLargeClear(int* __restrict__ acc, uint32_t nodes) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nodes * kAccWords) return;
    const int w = int(i % kAccWords);
    int v;
    if (w >= kAccZero && w < kAccBins) v = int(kNoZero);
    else if (w < kAccBins) v = (w % 6) < 3 ? fkey(kFltMax) : fkey(-kFltMax);
    else {
        const int j = (w - kAccBins) % kAccPerBin;
        v = j == 6 ? 0 : (j < 3 ? fkey(kFltMax) : fkey(-kFltMax));
    }
    acc[i] = v;
}

// Box, centroid box and zero codes of each chunk, combined per node.
__global__ __launch_bounds__(256) void kLargeBounds(const float* __restrict__ bounds, const uint32_t* __restrict__ idx,
                                                    const LargeChunk* __restrict__ chunks, int* __restrict__ acc) {
    __shared__ int red[18][256];
    const LargeChunk c = chunks[blockIdx.x];
    Box b, cb;
    boxEmpty(b);
    boxEmpty(cb);
    uint32_t z[6] = { kNoZero, kNoZero, kNoZero, kNoZero, kNoZero, kNoZero };
    for (uint32_t s = c.begin + threadIdx.x; s < c.end; s += 256) {
        float lo[3], hi[3], cc[3];
        loadPrim(bounds, idx[s], lo, hi);
        for (int a = 0; a < 3; ++a) {
            cc[a] = 0.5f * (lo[a] + hi[a]);
            z[a] = min(z[a], zeroCode(lo[a], s));
            z[3 + a] = min(z[3 + a], zeroCode(hi[a], s));
        }
        growBox(b, lo, hi);
        growBox(cb, cc, cc);
    }
    for (int a = 0; a < 3; ++a) {
        red[a][threadIdx.x] = fkeyNum(b.lo[a]);
        red[3 + a][threadIdx.x] = fkeyNum(b.hi[a]);
        red[6 + a][threadIdx.x] = fkey(cb.lo[a]);
        red[9 + a][threadIdx.x] = fkey(cb.hi[a]);
    }
    for (int v = 0; v < 6; ++v) red[12 + v][threadIdx.x] = int(z[v]);
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            for (int v = 0; v < 12; ++v) {
                const int o = red[v][threadIdx.x + h];
                const bool isMin = (v % 6) < 3;
                red[v][threadIdx.x] = isMin ? min(red[v][threadIdx.x], o) : max(red[v][threadIdx.x], o);
            }
            for (int v = 12; v < 18; ++v)
                red[v][threadIdx.x] = int(min(uint32_t(red[v][threadIdx.x]), uint32_t(red[v][threadIdx.x + h])));
        }
        __syncthreads();
    }
    if (threadIdx.x < 18) {
        int* w = acc + size_t(c.node) * kAccWords + threadIdx.x;  // box 0..5, centroid box 6..11, zero codes 12..17
        const int v = red[threadIdx.x][0];
        if (threadIdx.x >= 12) {
            if (uint32_t(v) != kNoZero) atomicMin(reinterpret_cast<unsigned int*>(w), uint32_t(v));
        } else if ((threadIdx.x % 6) < 3) {
            atomicMin(w, v);
        } else {
            atomicMax(w, v);
        }
    }
}

__device__ __forceinline__ void accBoxes(const int* a, Box& b, Box& cb) {
    for (int k = 0; k < 3; ++k) {
        b.lo[k] = zeroFromCode(funkey(a[kAccBox + k]), uint32_t(a[kAccZero + k]));
        b.hi[k] = zeroFromCode(funkey(a[kAccBox + 3 + k]), uint32_t(a[kAccZero + 3 + k]));
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
        const uint32_t bi = binOf(d, 0.5f * (pick(lo, d.axis) + pick(hi, d.axis)));
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
// are consecutive), one workgroup per node in rounds of 256 chunks;
// totals[node] = right-side slots that change side (as many as left-side
// ones), totals[nodeCount + node] = left-side ones.
__global__ __launch_bounds__(256) void kLargeScan(const LargeNode* __restrict__ nodes, uint32_t nodeCount,
                                                  uint32_t* __restrict__ counts, uint32_t* __restrict__ totals) {
    __shared__ uint32_t scan[2][256];
    __shared__ uint32_t carry[2];
    const LargeNode& n = nodes[blockIdx.x];
    if (threadIdx.x < 2) carry[threadIdx.x] = 0;
    __syncthreads();
    for (uint32_t c0 = 0; c0 < n.chunks; c0 += 256) {
        const uint32_t c = n.chunk0 + c0 + threadIdx.x;
        const bool in = c0 + threadIdx.x < n.chunks;
        const uint32_t l = in ? counts[2 * c] : 0u, r = in ? counts[2 * c + 1] : 0u;
        scan[0][threadIdx.x] = l;
        scan[1][threadIdx.x] = r;
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
        if (in) {
            counts[2 * c] = carry[0] + scan[0][threadIdx.x] - l;
            counts[2 * c + 1] = carry[1] + scan[1][threadIdx.x] - r;
        }
        __syncthreads();
        if (threadIdx.x < 2) carry[threadIdx.x] += scan[threadIdx.x][255];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        totals[blockIdx.x] = carry[1];
        totals[nodeCount + blockIdx.x] = carry[0];
    }
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
// One wavefront per subtree: every node decision is a wave-wide reduction,
// scan or ballot done with DPP row operations and lane reads, so the
// depth-first loop runs without workgroup barriers (with a one-wave
// workgroup __syncthreads orders the LDS accesses and emits no s_barrier).
constexpr uint32_t kWave = 64;

// DPP controls (within rows of 16 lanes)
constexpr int kQuadXor1 = 0xB1;       // quad_perm [1,0,3,2]
constexpr int kQuadXor2 = 0x4E;       // quad_perm [2,3,0,1]
constexpr int kRowHalfMirror = 0x141; // lane i <-> 7 - i within each half row
constexpr int kRowMirror = 0x140;     // lane i <-> 15 - i within each row
constexpr int kRowShr1 = 0x111, kRowShr2 = 0x112, kRowShr4 = 0x114, kRowShr8 = 0x118;
constexpr int kRowBcast15 = 0x142;    // lane 15 of each row to the next row

template <int Ctrl>
__device__ __forceinline__ int dppPerm(int v) {
    return __builtin_amdgcn_mov_dpp(v, Ctrl, 0xF, 0xF, false);
}
// Lanes whose source is outside the row (or whose row is masked off) take
// `id`.
template <int Ctrl, int RowMask>
__device__ __forceinline__ int dppShift(int v, int id) {
    return __builtin_amdgcn_update_dpp(id, v, Ctrl, RowMask, 0xF, false);
}

struct OpMin {
    static constexpr int id = 0x7FFFFFFF;
    __device__ static int f(int a, int b) { return min(a, b); }
};
struct OpMax {
    static constexpr int id = int(0x80000000);
    __device__ static int f(int a, int b) { return max(a, b); }
};
struct OpAdd {
    static constexpr int id = 0;
    __device__ static int f(int a, int b) { return a + b; }
};

// Reduction over all 64 lanes (every lane active), uniform result.
template <typename Op>
__device__ __forceinline__ int waveReduce(int v) {
    v = Op::f(v, dppPerm<kQuadXor1>(v));
    v = Op::f(v, dppPerm<kQuadXor2>(v));
    v = Op::f(v, dppPerm<kRowHalfMirror>(v));
    v = Op::f(v, dppPerm<kRowMirror>(v));
    const int a = Op::f(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16));
    const int b = Op::f(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48));
    return Op::f(a, b);
}
// Inclusive scan over lanes 0-31 and, separately, over lanes 32-63.
template <typename Op>
__device__ __forceinline__ int halfScan(int v) {
    v = Op::f(v, dppShift<kRowShr1, 0xF>(v, Op::id));
    v = Op::f(v, dppShift<kRowShr2, 0xF>(v, Op::id));
    v = Op::f(v, dppShift<kRowShr4, 0xF>(v, Op::id));
    v = Op::f(v, dppShift<kRowShr8, 0xF>(v, Op::id));
    return Op::f(v, dppShift<kRowBcast15, 0xA>(v, Op::id));
}

__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
// Lanes below this one with their bit set in m.
__device__ __forceinline__ uint32_t ranksBelow(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
}

// One subtree in LDS: primitive bounds and indices by slot, and the stack of
// nodes still to visit (begin | count << 16, node | depth << 16).
struct SmallLds {
    float b[6][kSmallMax];
    uint32_t idx[kSmallMax];
    uint2 stack[kSmallMax];
    uint32_t pos[kSmallMax / 2];
    int bins[kBins * kAccPerBin];
};

// Sign of the first slot of [b0, b0 + cnt) whose bound `comp` is zero.
__device__ float firstZero(const SmallLds& l, int comp, uint32_t b0, uint32_t cnt) {
    for (uint32_t r = 0; r < cnt; r += kWave) {
        const uint32_t i = r + threadIdx.x;
        const float v = i < cnt ? l.b[comp][b0 + i] : 1.0f;
        const uint64_t m = __ballot(isZero(v));
        if (m) return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), int(__builtin_ctzll(m))));
    }
    return 0.0f;
}

// A node of cnt <= kSmallMax primitives: its whole subtree, depth first.
// Subtree nodes are written to out + task.outBase with local indices (root
// 0, child pairs allocated in depth-first order from 1); leaves hold absolute
// slot indices.
__global__ __launch_bounds__(kWave) void kBuildSmall(const float* __restrict__ bounds, uint32_t* __restrict__ idx,
                                                    const SmallTask* __restrict__ tasks, OutNode* __restrict__ out,
                                                    SmallResult* __restrict__ results, uint32_t maxLeaf,
                                                    uint32_t* __restrict__ flags) {
    __shared__ SmallLds l;
    const uint32_t lane = threadIdx.x;
    const SmallTask t = tasks[blockIdx.x];
    for (uint32_t i = lane; i < t.count; i += kWave) {
        const uint32_t p = idx[t.begin + i];
        float lo[3], hi[3];
        loadPrim(bounds, p, lo, hi);
        l.idx[i] = p;
        for (int a = 0; a < 3; ++a) {
            l.b[a][i] = lo[a];
            l.b[3 + a][i] = hi[a];
        }
    }
    __syncthreads();

    // the node being visited; after a split, its left child
    uint32_t b0 = 0, cnt = t.count, node = 0, depth = t.depth;
    uint32_t sp = 0, alloc = 1, maxDepth = 0, leaves = 0;
    for (;;) {
        // box and centroid box
        Box b, cb;
        boxEmpty(b);
        boxEmpty(cb);
        bool zero = false;
        for (uint32_t i = lane; i < cnt; i += kWave) {
            float lo[3], hi[3], cc[3];
            for (int a = 0; a < 3; ++a) {
                lo[a] = l.b[a][b0 + i];
                hi[a] = l.b[3 + a][b0 + i];
                cc[a] = 0.5f * (lo[a] + hi[a]);
                zero = zero || isZero(lo[a]) || isZero(hi[a]);
            }
            growBox(b, lo, hi);
            growBox(cb, cc, cc);
        }
        for (int a = 0; a < 3; ++a) {
            b.lo[a] = funkey(waveReduce<OpMin>(fkeyNum(b.lo[a])));
            b.hi[a] = funkey(waveReduce<OpMax>(fkeyNum(b.hi[a])));
            cb.lo[a] = funkey(waveReduce<OpMin>(fkey(cb.lo[a])));
            cb.hi[a] = funkey(waveReduce<OpMax>(fkey(cb.hi[a])));
        }
        if (__ballot(zero))
            for (int a = 0; a < 3; ++a) {
                if (b.lo[a] == 0.0f) b.lo[a] = firstZero(l, a, b0, cnt);
                if (b.hi[a] == 0.0f) b.hi[a] = firstZero(l, 3 + a, b0, cnt);
            }
        const Decision d = decide(cb, cnt, maxLeaf);

        uint32_t split = kSplitLeaf, mid = 0;
        if (!d.leaf) {
            for (uint32_t i = lane; i < kBins * kAccPerBin; i += kWave) {
                const uint32_t j = i % kAccPerBin;
                l.bins[i] = j == 6 ? 0 : (j < 3 ? fkey(kFltMax) : fkey(-kFltMax));
            }
            __syncthreads();
            for (uint32_t i = lane; i < cnt; i += kWave) {
                const uint32_t bi = binOf(d, 0.5f * (l.b[d.axis][b0 + i] + l.b[3 + d.axis][b0 + i]));
                int* w = l.bins + bi * kAccPerBin;
                for (int a = 0; a < 3; ++a) {
                    atomicMin(w + a, fkey(l.b[a][b0 + i]));
                    atomicMax(w + 3 + a, fkey(l.b[3 + a][b0 + i]));
                }
                atomicAdd(w + 6, 1);
            }
            __syncthreads();
            // SAH over the 31 split planes (sahSweep as scans): lanes 0-31 hold
            // bins 0..31 and scan the left unions, lanes 32-63 hold bins
            // 31..0 and scan the right unions; the cost of split plane i is
            // the same expression on the same values.
            const uint32_t bin = lane < kBins ? lane : 2 * kBins - 1 - lane;
            Box u;
            int k[7];
            for (int a = 0; a < 3; ++a) {
                k[a] = halfScan<OpMin>(l.bins[bin * kAccPerBin + a]);
                k[3 + a] = halfScan<OpMax>(l.bins[bin * kAccPerBin + 3 + a]);
            }
            const uint32_t c = uint32_t(halfScan<OpAdd>(l.bins[bin * kAccPerBin + 6]));
            for (int a = 0; a < 3; ++a) {
                u.lo[a] = funkey(k[a]);
                u.hi[a] = funkey(k[3 + a]);
            }
            const float area = boxArea(u) * float(c);  // l of sahSweep (lanes < 32) or r (lanes >= 32)
            // the right side of plane i starts at bin i + 1: lane 62 - i
            const int from = int(2 * kBins - 2) - int(lane);
            const float rArea = __shfl(area, from & 63);
            const uint32_t rc = __shfl(c, from & 63);
            const float invRoot = 1.0f / fmaxf(boxArea(b), 1e-12f);
            const float scale = kIntersectCost * invRoot;
            const float cost = kTraversalCost + scale * (area + rArea);
            const bool ok = lane + 1 < kBins && c != 0 && rc != 0 && cost < kIntersectCost * float(cnt);
            const int best = waveReduce<OpMin>(ok ? __float_as_int(cost) : OpMin::id);
            if (best != OpMin::id) {
                split = uint32_t(__builtin_ctzll(__ballot(ok && __float_as_int(cost) == best)));
                mid = uint32_t(__builtin_amdgcn_readlane(int(c), int(split)));
            } else if (cnt > 2 * maxLeaf && lane == 0) {
                atomicOr(flags, kBuildNeedsMedian);  // the reference's nth_element split
            }
        }

        if (split == kSplitLeaf) {
            if (lane == 0) {
                uint4* o = reinterpret_cast<uint4*>(out + t.outBase + node);
                o[0] = make_uint4(__float_as_uint(b.lo[0]), __float_as_uint(b.lo[1]), __float_as_uint(b.lo[2]),
                                  t.begin + b0);
                o[1] = make_uint4(__float_as_uint(b.hi[0]), __float_as_uint(b.hi[1]), __float_as_uint(b.hi[2]),
                                  (cnt << 2) | uint32_t(d.axis));
            }
            maxDepth = max(maxDepth, depth);
            ++leaves;
            if (sp == 0) break;
            --sp;
            const uint32_t e0 = uniform(l.stack[sp].x), e1 = uniform(l.stack[sp].y);
            b0 = e0 & 0xFFFFu;
            cnt = e0 >> 16;
            node = e1 & 0xFFFFu;
            depth = e1 >> 16;
            continue;
        }
        if (mid > 0 && mid < cnt) {
            // right-side slots passing the predicate, ranked from the left
            uint32_t total = 0;
            for (uint32_t r = mid; r < cnt; r += kWave) {
                const uint32_t i = r + lane;
                const bool mis = i < cnt && binOf(d, 0.5f * (l.b[d.axis][b0 + i] + l.b[3 + d.axis][b0 + i])) <= split;
                const uint64_t m = __ballot(mis);
                if (mis) l.pos[total + ranksBelow(m)] = i;
                total += uint32_t(__popcll(m));
            }
            __syncthreads();
            // left-side slots failing it swap with them counted from the right
            uint32_t seen = 0;
            for (uint32_t r = 0; r < mid; r += kWave) {
                const uint32_t i = r + lane;
                const bool mis = i < mid && binOf(d, 0.5f * (l.b[d.axis][b0 + i] + l.b[3 + d.axis][b0 + i])) > split;
                const uint64_t m = __ballot(mis);
                const uint32_t rank = seen + ranksBelow(m);
                seen += uint32_t(__popcll(m));
                if (mis && rank < total) {  // rank < total always: both sides hold `total` such slots
                    const uint32_t q = l.pos[total - 1 - rank];
                    const uint32_t a = l.idx[b0 + i], e = l.idx[b0 + q];
                    l.idx[b0 + i] = e;
                    l.idx[b0 + q] = a;
                    for (int j = 0; j < 6; ++j) {
                        const float va = l.b[j][b0 + i], ve = l.b[j][b0 + q];
                        l.b[j][b0 + i] = ve;
                        l.b[j][b0 + q] = va;
                    }
                }
            }
        } else {
            mid = cnt / 2;
        }
        const uint32_t left = alloc;
        alloc += 2;
        if (lane == 0) {
            uint4* o = reinterpret_cast<uint4*>(out + t.outBase + node);
            o[0] = make_uint4(__float_as_uint(b.lo[0]), __float_as_uint(b.lo[1]), __float_as_uint(b.lo[2]), left);
            o[1] = make_uint4(__float_as_uint(b.hi[0]), __float_as_uint(b.hi[1]), __float_as_uint(b.hi[2]),
                              uint32_t(d.axis));  // local child index; placed by kPlaceSmall
            l.stack[sp] = make_uint2((b0 + mid) | ((cnt - mid) << 16), (left + 1) | ((depth + 1) << 16));
        }
        ++sp;
        cnt = mid;
        node = left;
        ++depth;
        __syncthreads();
    }
    for (uint32_t i = lane; i < t.count; i += kWave) idx[t.begin + i] = l.idx[i];
    if (lane == 0) {
        results[blockIdx.x].nodes = alloc;
        results[blockIdx.x].maxDepth = maxDepth;
        results[blockIdx.x].leaves = leaves;
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
    hipLaunchKernelGGL(kLargeScan, dim3(a.nodeCount), dim3(256), 0, s, a.nodes, a.nodeCount, a.counts, a.totals);
    hipLaunchKernelGGL(kLargeRight, dim3(a.chunkCount), dim3(256), 0, s, a.bounds, a.idx, a.chunks, a.nodes,
                       a.counts, a.pos);
    hipLaunchKernelGGL(kLargeSwap, dim3(a.chunkCount), dim3(256), 0, s, a.bounds, a.idx, a.chunks, a.nodes, a.counts,
                       a.pos, a.totals);
    return hipGetLastError();
}

hipError_t launchBvhSmall(const SmallArgs& a, hipStream_t s) {
    if (a.taskCount == 0) return hipSuccess;
    hipLaunchKernelGGL(kBuildSmall, dim3(a.taskCount), dim3(kWave), 0, s, a.bounds, a.idx, a.tasks, a.tmp,
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
