// Ray-query kernels over one BVH (closest hit / any hit) for gfx950.
//
// One lane per ray.  A lane keeps the current node in registers; at an
// internal node it fetches both children (adjacent 32-byte records), tests
// both boxes, continues into the nearer hit child and pushes the farther one
// with its entry distance; a popped entry is skipped without a fetch when its
// entry distance is beyond the current limit.  Closest-hit ties in t resolve
// to the lowest triangle id.
#include "launch.h"
#include "trace_core.h"

namespace mrt {
namespace dev {

template <int BLOCK, int S>
__global__ __launch_bounds__(BLOCK) void kTraceClosest(const Node* __restrict__ nodes,
                                                       const float4* __restrict__ tris,
                                                       const float4* __restrict__ rays,
                                                       uint32_t n, uint2* __restrict__ hits,
                                                       uint32_t* __restrict__ status) {
    __shared__ uint2 ldsStack[S * BLOCK];
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const float4 r0 = rays[2 * size_t(i)];
    const float4 r1 = rays[2 * size_t(i) + 1];
    const RayK r = setupRay(r0.x, r0.y, r0.z, r1.x, r1.y, r1.z, r0.w);

    float best = r1.w;
    uint32_t bestId = kNoHit;
    bool hit = false;

    ShortStack<S, BLOCK> stack(ldsStack + threadIdx.x);
    Node cur = loadNode(nodes, 0);
    float te;
    bool have = slab(cur, r, best * kLimitWiden, te);
    while (true) {
        if (have) {
            const uint32_t lp = nodeLeft(cur);
            const uint32_t cnt = nodeCount(cur);
            if (cnt == 0) {
                const Node c0 = loadNode(nodes, lp);
                const Node c1 = loadNode(nodes, lp + 1);
                const float lim = best * kLimitWiden;
                float t0, t1;
                const bool h0 = slab(c0, r, lim, t0);
                const bool h1 = slab(c1, r, lim, t1);
                if (h0 && h1) {
                    const bool firstNear = t0 <= t1;
                    stack.push(make_uint2(firstNear ? lp + 1 : lp, __float_as_uint(firstNear ? t1 : t0)));
                    cur = firstNear ? c0 : c1;
                    continue;
                }
                if (h0 || h1) {
                    cur = h0 ? c0 : c1;
                    continue;
                }
            } else {
                for (uint32_t k = lp; k < lp + cnt; ++k) {
                    const float4 v0 = tris[3 * size_t(k)];
                    const float4 v1 = tris[3 * size_t(k) + 1];
                    const float4 v2 = tris[3 * size_t(k) + 2];
                    float t, b1, b2;
                    if (!isectTri(r, v0, v1, v2, best, t, b1, b2)) continue;
                    const uint32_t id = __float_as_uint(v0.w);
                    if (t < best || (hit && t == best && id < bestId)) {
                        best = t;
                        bestId = id;
                        hit = true;
                    }
                }
            }
        }
        have = false;
        while (!stack.empty()) {
            const uint2 e = stack.pop();
            if (__uint_as_float(e.y) <= best * kLimitWiden) {
                cur = loadNode(nodes, e.x);
                have = true;
                break;
            }
        }
        if (!have) break;
    }
    hits[i] = make_uint2(bestId, __float_as_uint(best));
    if (stack.overflow) atomicOr(status, 1u);
}

template <int BLOCK, int S>
__global__ __launch_bounds__(BLOCK) void kTraceAny(const Node* __restrict__ nodes,
                                                   const float4* __restrict__ tris,
                                                   const float4* __restrict__ rays, uint32_t n,
                                                   uint32_t* __restrict__ occluded,
                                                   uint32_t* __restrict__ status) {
    __shared__ uint2 ldsStack[S * BLOCK];
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const float4 r0 = rays[2 * size_t(i)];
    const float4 r1 = rays[2 * size_t(i) + 1];
    const RayK r = setupRay(r0.x, r0.y, r0.z, r1.x, r1.y, r1.z, r0.w);
    const float tmax = r1.w;
    const float lim = tmax * kLimitWiden;

    ShortStack<S, BLOCK> stack(ldsStack + threadIdx.x);
    uint32_t occ = 0;
    Node cur = loadNode(nodes, 0);
    float te;
    bool have = slab(cur, r, lim, te);
    while (!occ) {
        if (have) {
            const uint32_t lp = nodeLeft(cur);
            const uint32_t cnt = nodeCount(cur);
            if (cnt == 0) {
                const Node c0 = loadNode(nodes, lp);
                const Node c1 = loadNode(nodes, lp + 1);
                float t0, t1;
                const bool h0 = slab(c0, r, lim, t0);
                const bool h1 = slab(c1, r, lim, t1);
                if (h0 && h1) {
                    stack.push(make_uint2(lp + 1, __float_as_uint(t1)));
                    cur = c0;
                    continue;
                }
                if (h0 || h1) {
                    cur = h0 ? c0 : c1;
                    continue;
                }
            } else {
                for (uint32_t k = lp; k < lp + cnt; ++k) {
                    float t, b1, b2;
                    if (isectTri(r, tris[3 * size_t(k)], tris[3 * size_t(k) + 1],
                                 tris[3 * size_t(k) + 2], tmax, t, b1, b2) && t < tmax) {
                        occ = 1;
                        break;
                    }
                }
                if (occ) break;
            }
        }
        if (stack.empty()) break;
        cur = loadNode(nodes, stack.pop().x);
        have = true;
    }
    occluded[i] = occ;
    if (stack.overflow) atomicOr(status, 1u);
}

template <int BLOCK, int S>
static hipError_t launchClosestT(const QueryLaunch& q, hipStream_t s) {
    const uint32_t grid = (q.n + BLOCK - 1) / BLOCK;
    hipLaunchKernelGGL((kTraceClosest<BLOCK, S>), dim3(grid), dim3(BLOCK), 0, s, q.nodes, q.tris,
                       q.rays, q.n, reinterpret_cast<uint2*>(q.out), q.status);
    return hipGetLastError();
}
template <int BLOCK, int S>
static hipError_t launchAnyT(const QueryLaunch& q, hipStream_t s) {
    const uint32_t grid = (q.n + BLOCK - 1) / BLOCK;
    hipLaunchKernelGGL((kTraceAny<BLOCK, S>), dim3(grid), dim3(BLOCK), 0, s, q.nodes, q.tris,
                       q.rays, q.n, reinterpret_cast<uint32_t*>(q.out), q.status);
    return hipGetLastError();
}

#define MRT_QUERY_SHAPES(X) \
    X(64, 8) X(64, 16) X(64, 32) X(128, 8) X(128, 16) X(128, 32) X(256, 8) X(256, 16) X(256, 32)

hipError_t launchTraceClosest(const QueryLaunch& q, uint32_t block, uint32_t stack, hipStream_t s) {
    if (q.n == 0) return hipSuccess;
#define X(B, S) if (block == B && stack == S) return launchClosestT<B, S>(q, s);
    MRT_QUERY_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launchTraceAny(const QueryLaunch& q, uint32_t block, uint32_t stack, hipStream_t s) {
    if (q.n == 0) return hipSuccess;
#define X(B, S) if (block == B && stack == S) return launchAnyT<B, S>(q, s);
    MRT_QUERY_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

bool traversalShapeSupported(uint32_t block, uint32_t stack) {
#define X(B, S) if (block == B && stack == S) return true;
    MRT_QUERY_SHAPES(X)
#undef X
    return false;
}

} // namespace dev
} // namespace mrt
