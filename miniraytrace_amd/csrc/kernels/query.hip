// Ray-query kernels over one BVH (closest hit / any hit) for gfx950.
//
// One lane per ray.  A lane keeps the cursor of the current node in
// registers; at an internal node it fetches both children (adjacent 32-byte
// records, one 64-byte line), tests both boxes, continues into the nearer hit
// child and pushes the farther one's cursor with its entry distance; a popped
// entry is skipped when its entry distance is beyond the current limit, and
// otherwise continues without a node fetch.  Closest-hit ties in t resolve
// to the lowest triangle id.
//
// Loop structures (chosen per scene by the autotuner):
//   LOOP 0  one step per iteration: an inner-node step or a leaf step (a
//           whole leaf: one triangle per step measured 30 % slower here);
//   LOOP 1  "while-while" (Aila & Laine 2009): descend until the lane holds
//           a leaf, so the lanes of a wave test their leaf triangles together;
//   LOOP 2  as 0 over the four-wide tree collapsed from the binary one
//           (trace_core.h, host/bvh_wide.h): one step tests four child boxes
//           of one 128-byte node, loading each axis's near and far planes of
//           all four from the offsets the ray's direction signs pick, visits
//           the nearest hit child and pushes the other hit children farthest
//           first;
//   LOOP 3  as 1 over the four-wide tree.
// All return the same hits: box tests are conservative and ties in t go to
// the lowest triangle id, so no result depends on the tree or visiting order.
#include "launch.h"
#include "trace_core.h"

namespace mrt {
namespace dev {

struct ClosestState {
    float best;
    uint32_t bestId;
    bool hit;
};

__device__ __forceinline__ void leafClosest(const float4* __restrict__ tris, const RayK& r, uint32_t lp,
                                            uint32_t cnt, ClosestState& st) {
    for (uint32_t k = lp; k < lp + cnt; ++k) {
        const float4 v0 = tris[3 * size_t(k)];
        const float4 v1 = tris[3 * size_t(k) + 1];
        const float4 v2 = tris[3 * size_t(k) + 2];
        float t, b1, b2;
        if (!isectTri(r, v0, v1, v2, st.best, t, b1, b2)) continue;
        const uint32_t id = __float_as_uint(v0.w);
        if (t < st.best || (st.hit && t == st.best && id < st.bestId)) {
            st.best = t;
            st.bestId = id;
            st.hit = true;
        }
    }
}

// Pops until an entry within the limit is found; returns false when empty.
template <int S, int BLOCK>
__device__ __forceinline__ bool popNext(ShortStack<S, BLOCK>& stack, const Node* __restrict__ nodes,
                                        float limit, Cursor& cur) {
    while (!stack.empty()) {
        const uint2 e = stack.pop();
        if (__uint_as_float(e.y) <= limit) {
            cur = entryCursor(nodes, e);
            return true;
        }
    }
    return false;
}

// Inner-node step: fetches both children of `cur`, returns true and sets cur
// to the next node to visit (pushing the farther child when both are hit);
// false when neither child is hit.
template <int S, int BLOCK>
__device__ __forceinline__ bool innerStep(const Node* __restrict__ nodes, const RayK& r, float lim,
                                          ShortStack<S, BLOCK>& stack, Cursor& cur) {
    const uint32_t lp = cur.index;
    const Node c0 = loadNode(nodes, lp);
    const Node c1 = loadNode(nodes, lp + 1);
    float t0, t1;
    const bool h0 = slab(c0, r, lim, t0);
    const bool h1 = slab(c1, r, lim, t1);
    if (h0 && h1) {
        const bool firstNear = t0 <= t1;
        stack.push(firstNear ? nodeEntry(c1, lp + 1, t1) : nodeEntry(c0, lp, t0));
        cur = cursorOf(firstNear ? c0 : c1);
        return true;
    }
    if (h0 || h1) {
        cur = cursorOf(h0 ? c0 : c1);
        return true;
    }
    return false;
}

// Wide inner step: tests the four children of wide node cur.index, visits
// the nearest hit child and pushes the other hit children farthest first;
// false when none is hit.
template <int S, int BLOCK>
__device__ __forceinline__ bool wideInnerStep(const float4* __restrict__ wide, const Node* __restrict__ nodes,
                                              const RayK& r, const WideRows& rows, float lim,
                                              ShortStack<S, BLOCK>& stack, Cursor& cur) {
    WideStep ws = wideTest(wide, cur.index, r, rows, lim);
    const int hits = wideSort(ws);
    if (hits == 0) return false;
    if (hits > 3) stack.push(make_uint2(ws.entry[3], ws.key[3]));
    if (hits > 2) stack.push(make_uint2(ws.entry[2], ws.key[2]));
    if (hits > 1) stack.push(make_uint2(ws.entry[1], ws.key[1]));
    cur = wideEntryCursor(nodes, make_uint2(ws.entry[0], ws.key[0]));
    return true;
}

template <int S, int BLOCK>
__device__ __forceinline__ bool widePopNext(ShortStack<S, BLOCK>& stack, const Node* __restrict__ nodes,
                                            float limit, Cursor& cur) {
    while (!stack.empty()) {
        const uint2 e = stack.pop();
        if (__uint_as_float(e.y) <= limit) {
            cur = wideEntryCursor(nodes, e);
            return true;
        }
    }
    return false;
}

template <int BLOCK, int S, int LOOP>
__global__ __launch_bounds__(BLOCK) void kTraceClosestWide(const float4* __restrict__ wide,
                                                           const Node* __restrict__ nodes,
                                                           const float4* __restrict__ tris,
                                                           const float4* __restrict__ rays, uint32_t n,
                                                           uint2* __restrict__ hits, uint32_t* __restrict__ status) {
    __shared__ uint2 ldsStack[S * BLOCK];
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const float4 r0 = rays[2 * size_t(i)];
    const float4 r1 = rays[2 * size_t(i) + 1];
    const RayK r = setupRay(r0.x, r0.y, r0.z, r1.x, r1.y, r1.z, r0.w);
    const WideRows rows = wideRows(r);
    ClosestState st{ r1.w, kNoHit, false };
    uint2 spill[kSpill];
    ShortStack<S, BLOCK> stack(ldsStack + threadIdx.x, spill);
    Cursor cur{ 0u, 0u };  // the root wide node
    bool have = true;
    if (LOOP == 2) {
        while (have) {
            if (cur.count == 0) {
                if (wideInnerStep(wide, nodes, r, rows, st.best * kLimitWiden, stack, cur)) continue;
            } else {
                leafClosest(tris, r, cur.index, cur.count, st);
            }
            have = widePopNext(stack, nodes, st.best * kLimitWiden, cur);
        }
    } else {
        while (have) {
            while (have && cur.count == 0) {
                if (!wideInnerStep(wide, nodes, r, rows, st.best * kLimitWiden, stack, cur))
                    have = widePopNext(stack, nodes, st.best * kLimitWiden, cur);
            }
            if (!have) break;
            leafClosest(tris, r, cur.index, cur.count, st);
            have = widePopNext(stack, nodes, st.best * kLimitWiden, cur);
        }
    }
    hits[i] = make_uint2(st.bestId, __float_as_uint(st.best));
    if (stack.overflow) atomicOr(status, 1u);
}

template <int BLOCK, int S, int LOOP>
__global__ __launch_bounds__(BLOCK) void kTraceAnyWide(const float4* __restrict__ wide,
                                                       const Node* __restrict__ nodes,
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
    const WideRows rows = wideRows(r);
    const float tmax = r1.w;
    const float lim = tmax * kLimitWiden;
    uint2 spill[kSpill];
    ShortStack<S, BLOCK> stack(ldsStack + threadIdx.x, spill);
    uint32_t occ = 0;
    Cursor cur{ 0u, 0u };
    bool have = true;
    while (have) {
        if (LOOP == 3) {
            while (have && cur.count == 0) {
                if (!wideInnerStep(wide, nodes, r, rows, lim, stack, cur)) have = widePopNext(stack, nodes, lim, cur);
            }
            if (!have) break;
        }
        if (cur.count == 0) {
            if (wideInnerStep(wide, nodes, r, rows, lim, stack, cur)) continue;
        } else {
            const uint32_t lp = cur.index;
            for (uint32_t k = lp; k < lp + cur.count; ++k) {
                float t, b1, b2;
                if (isectTri(r, tris[3 * size_t(k)], tris[3 * size_t(k) + 1], tris[3 * size_t(k) + 2], tmax, t,
                             b1, b2) &&
                    t < tmax) {
                    occ = 1;
                    break;
                }
            }
            if (occ) break;
        }
        have = widePopNext(stack, nodes, lim, cur);
    }
    occluded[i] = occ;
    if (stack.overflow) atomicOr(status, 1u);
}

template <int BLOCK, int S, int LOOP>
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

    ClosestState st{ r1.w, kNoHit, false };
    uint2 spill[kSpill];
    ShortStack<S, BLOCK> stack(ldsStack + threadIdx.x, spill);
    const Node root = loadNode(nodes, 0);
    Cursor cur = cursorOf(root);
    float te;
    bool have = slab(root, r, st.best * kLimitWiden, te);
    if (LOOP == 0) {
        while (have) {
            if (cur.count == 0) {
                if (innerStep(nodes, r, st.best * kLimitWiden, stack, cur)) continue;
            } else {
                leafClosest(tris, r, cur.index, cur.count, st);
            }
            have = popNext(stack, nodes, st.best * kLimitWiden, cur);
        }
    } else {
        while (have) {
            // descend to a leaf
            while (have && cur.count == 0) {
                if (!innerStep(nodes, r, st.best * kLimitWiden, stack, cur))
                    have = popNext(stack, nodes, st.best * kLimitWiden, cur);
            }
            if (!have) break;
            leafClosest(tris, r, cur.index, cur.count, st);
            have = popNext(stack, nodes, st.best * kLimitWiden, cur);
        }
    }
    hits[i] = make_uint2(st.bestId, __float_as_uint(st.best));
    if (stack.overflow) atomicOr(status, 1u);
}

template <int BLOCK, int S, int LOOP>
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

    uint2 spill[kSpill];
    ShortStack<S, BLOCK> stack(ldsStack + threadIdx.x, spill);
    uint32_t occ = 0;
    const Node root = loadNode(nodes, 0);
    Cursor cur = cursorOf(root);
    float te;
    bool have = slab(root, r, lim, te);
    while (have) {
        if (LOOP != 0) {
            while (have && cur.count == 0) {
                if (!innerStep(nodes, r, lim, stack, cur)) have = popNext(stack, nodes, lim, cur);
            }
            if (!have) break;
        }
        const uint32_t cnt = cur.count;
        if (cnt == 0) {
            if (innerStep(nodes, r, lim, stack, cur)) continue;
        } else {
            const uint32_t lp = cur.index;
            for (uint32_t k = lp; k < lp + cnt; ++k) {
                float t, b1, b2;
                if (isectTri(r, tris[3 * size_t(k)], tris[3 * size_t(k) + 1], tris[3 * size_t(k) + 2], tmax, t,
                             b1, b2) &&
                    t < tmax) {
                    occ = 1;
                    break;
                }
            }
            if (occ) break;
        }
        have = popNext(stack, nodes, lim, cur);
    }
    occluded[i] = occ;
    if (stack.overflow) atomicOr(status, 1u);
}

template <int BLOCK, int S, int LOOP>
static hipError_t launchClosestT(const QueryLaunch& q, hipStream_t s) {
    const uint32_t grid = (q.n + BLOCK - 1) / BLOCK;
    if constexpr (LOOP >= 2)
        hipLaunchKernelGGL((kTraceClosestWide<BLOCK, S, LOOP>), dim3(grid), dim3(BLOCK), 0, s, q.wide, q.nodes, q.tris,
                           q.rays, q.n, reinterpret_cast<uint2*>(q.out), q.status);
    else
        hipLaunchKernelGGL((kTraceClosest<BLOCK, S, LOOP>), dim3(grid), dim3(BLOCK), 0, s, q.nodes, q.tris,
                           q.rays, q.n, reinterpret_cast<uint2*>(q.out), q.status);
    return hipGetLastError();
}
template <int BLOCK, int S, int LOOP>
static hipError_t launchAnyT(const QueryLaunch& q, hipStream_t s) {
    const uint32_t grid = (q.n + BLOCK - 1) / BLOCK;
    if constexpr (LOOP >= 2)
        hipLaunchKernelGGL((kTraceAnyWide<BLOCK, S, LOOP>), dim3(grid), dim3(BLOCK), 0, s, q.wide, q.nodes, q.tris,
                           q.rays, q.n, reinterpret_cast<uint32_t*>(q.out), q.status);
    else
        hipLaunchKernelGGL((kTraceAny<BLOCK, S, LOOP>), dim3(grid), dim3(BLOCK), 0, s, q.nodes, q.tris,
                           q.rays, q.n, reinterpret_cast<uint32_t*>(q.out), q.status);
    return hipGetLastError();
}

#define MRT_QUERY_SHAPES(X)                                                                          \
    X(64, 8, 0) X(64, 16, 0) X(64, 32, 0) X(128, 8, 0) X(128, 16, 0) X(128, 32, 0) X(256, 8, 0)       \
    X(256, 16, 0) X(256, 32, 0) X(64, 8, 1) X(64, 16, 1) X(64, 32, 1) X(128, 8, 1) X(128, 16, 1)     \
    X(128, 32, 1) X(256, 8, 1) X(256, 16, 1) X(256, 32, 1)                                            \
    X(64, 8, 2) X(64, 16, 2) X(64, 32, 2) X(128, 8, 2) X(128, 16, 2) X(128, 32, 2) X(256, 8, 2)       \
    X(256, 16, 2) X(256, 32, 2) X(64, 8, 3) X(64, 16, 3) X(64, 32, 3) X(128, 8, 3) X(128, 16, 3)     \
    X(128, 32, 3) X(256, 8, 3) X(256, 16, 3) X(256, 32, 3)

hipError_t launchTraceClosest(const QueryLaunch& q, uint32_t block, uint32_t stack, uint32_t loop,
                              hipStream_t s) {
    if (q.n == 0) return hipSuccess;
#define X(B, S, L) if (block == B && stack == S && loop == L) return launchClosestT<B, S, L>(q, s);
    MRT_QUERY_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launchTraceAny(const QueryLaunch& q, uint32_t block, uint32_t stack, uint32_t loop,
                          hipStream_t s) {
    if (q.n == 0) return hipSuccess;
#define X(B, S, L) if (block == B && stack == S && loop == L) return launchAnyT<B, S, L>(q, s);
    MRT_QUERY_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

bool traversalShapeSupported(uint32_t block, uint32_t stack, uint32_t loop) {
#define X(B, S, L) if (block == B && stack == S && loop == L) return true;
    MRT_QUERY_SHAPES(X)
#undef X
    return false;
}

} // namespace dev
} // namespace mrt
