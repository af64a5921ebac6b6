// Two-level traversal (TLAS over instances, BLAS over triangles) for the
// path tracer's extend (closest hit) and shadow (any hit) stages.
//
// Semantics follow core/shaders/intersect.glsl:88-178: the world ray is
// transformed into each instance's object space with an unnormalised
// direction, so hit distances are shared between spaces and one limit prunes
// both levels.  Differences from the reference, all in the direction of
// exactness: watertight triangle test, conservative box test, TLAS leaves
// holding several instances (coincident centroids) are all visited, ties in
// t resolve to the lowest (instance, triangle) pair, and the stack spills
// instead of overflowing.
#pragma once

#include "common/gpu_types.h"
#include "pt_params.h"
#include "trace_core.h"

namespace mrt {
namespace dev {

struct SceneHit {
    float t, b1, b2;
    uint32_t tri, inst;
};

__device__ __forceinline__ void toObject(const TraceInstance& ti, float ox, float oy, float oz,
                                         float dx, float dy, float dz, float3& o, float3& d) {
    o.x = ((ti.row0.x * ox + ti.row0.y * oy) + ti.row0.z * oz) + ti.row0.w;
    o.y = ((ti.row1.x * ox + ti.row1.y * oy) + ti.row1.z * oz) + ti.row1.w;
    o.z = ((ti.row2.x * ox + ti.row2.y * oy) + ti.row2.z * oz) + ti.row2.w;
    d.x = (ti.row0.x * dx + ti.row0.y * dy) + ti.row0.z * dz;
    d.y = (ti.row1.x * dx + ti.row1.y * dy) + ti.row1.z * dz;
    d.z = (ti.row2.x * dx + ti.row2.y * dy) + ti.row2.z * dz;
}

// Tests triangle k (vertices v0..v2) of the current instance.  ANY: returns
// true when it occludes (t < tmax).  Closest hit: keeps it if it is nearer
// than `best`, ties going to the lowest (instance, triangle) pair; returns
// false.
template <bool ANY>
__device__ __forceinline__ bool testTriangle(const RayK& r, const float4& v0, const float4& v1, const float4& v2,
                                             uint32_t k, uint32_t curInst, float tmax, float& best, bool& found,
                                             uint32_t& bestTri, uint32_t& bestInst, float& bestB1, float& bestB2,
                                             bool anyLane = false) {
    float t, b1, b2;
    if (!isectTri(r, v0, v1, v2, best, t, b1, b2)) return false;
    if (ANY || anyLane) return t < tmax;
    const bool better = t < best ||
                        (found && t == best && (curInst < bestInst || (curInst == bestInst && k < bestTri)));
    if (better) {
        best = t;
        found = true;
        bestTri = k;
        bestInst = curInst;
        bestB1 = b1;
        bestB2 = b2;
    }
    return false;
}

template <bool ANY>
__device__ __forceinline__ bool blasTriangle(const SceneView& sc, const RayK& r, uint32_t k, uint32_t curInst,
                                             float tmax, float& best, bool& found, uint32_t& bestTri,
                                             uint32_t& bestInst, float& bestB1, float& bestB2, bool anyLane = false) {
    const float4* tv = sc.triVerts + 3 * size_t(k);
    return testTriangle<ANY>(r, ldGlobal4(tv), ldGlobal4(tv + 1), ldGlobal4(tv + 2), k, curInst, tmax, best, found,
                             bestTri, bestInst, bestB1, bestB2, anyLane);
}

// ANY = false: closest hit in (tmin, tmax) -> returns true and fills `out`.
// ANY = true: returns true as soon as some triangle has tmin < t < tmax.
// TRI_STEP = true: a traversal step tests one triangle of a leaf and the lane
// goes on with the rest of its leaf in the following steps.  In the batched
// trace kernel almost every step has some lane of the wave at a leaf, so the
// wave's leaf branch then costs one test instead of the longest leaf's count.
// TRI_STEP = false tests the whole leaf in one step (the tail kernel, whose
// waves hold few live lanes, is faster that way).
template <bool ANY, int S, int BLOCK, bool TRI_STEP = true>
__device__ bool traceScene(const SceneView& sc, uint32_t tlasRoot, uint32_t instanceCount,
                           float3 o, float3 d, float tmin, float tmax, uint2* ldsLane,
                           SceneHit& out, bool& overflow) {
    if (instanceCount == 0) return false;
    // The world ray's constants are rebuilt when traversal returns from an
    // instance rather than kept, which keeps register pressure down.  The top
    // level only tests boxes (its leaves are instances), so the world ray
    // needs no triangle-test shear.
    RayK r = setupRayBox(o.x, o.y, o.z, d.x, d.y, d.z, tmin);
    uint2 spill[kSpill];
    ShortStack<S, BLOCK> stack(ldsLane, spill);
    bool inBlas = false;
    int blasBase = 0;
    uint32_t curInst = 0;
    float best = tmax;
    bool found = false;
    uint32_t bestTri = kNoHit, bestInst = kNoHit;
    float bestB1 = 0.0f, bestB2 = 0.0f;

    Node root = loadNode(sc.nodes, tlasRoot);
    Cursor cur = cursorOf(root);
    float te;
    bool have = slab(root, r, best * kLimitWiden, te);
    bool pendingInst = false;
    uint32_t pendingIdx = 0;

    while (true) {
        if (pendingInst) {
            // Enter an instance: object-space ray, BLAS root.
            pendingInst = false;
            const TraceInstance* tp = sc.traceInst + pendingIdx;
            TraceInstance ti;
            ti.row0 = ldGlobal4(&tp->row0);
            ti.row1 = ldGlobal4(&tp->row1);
            ti.row2 = ldGlobal4(&tp->row2);
            ti.blasRoot = __float_as_uint(ldGlobal4(reinterpret_cast<const float4*>(&tp->blasRoot)).x);
            float3 oo, dd;
            toObject(ti, o.x, o.y, o.z, d.x, d.y, d.z, oo, dd);
            r = setupRay(oo.x, oo.y, oo.z, dd.x, dd.y, dd.z, tmin);
            inBlas = true;
            blasBase = stack.sp;
            curInst = pendingIdx;
            root = loadNode(sc.nodes, ti.blasRoot);
            cur = cursorOf(root);
            have = slab(root, r, best * kLimitWiden, te);
        }
        if (have) {
            const uint32_t lp = cur.index;
            const uint32_t cnt = cur.count;
            if (cnt == 0) {
                const Node c0 = loadNode(sc.nodes, lp);
                const Node c1 = loadNode(sc.nodes, lp + 1);
                const float lim = best * kLimitWiden;
                float t0, t1;
                const bool h0 = slab(c0, r, lim, t0);
                const bool h1 = slab(c1, r, lim, t1);
                if (h0 && h1) {
                    const bool firstNear = t0 <= t1;
                    stack.push(firstNear ? nodeEntry(c1, lp + 1, t1) : nodeEntry(c0, lp, t0));
                    cur = cursorOf(firstNear ? c0 : c1);
                    continue;
                }
                if (h0 || h1) {
                    cur = cursorOf(h0 ? c0 : c1);
                    continue;
                }
            } else if (inBlas) {
                if constexpr (TRI_STEP) {
                    if (blasTriangle<ANY>(sc, r, lp, curInst, tmax, best, found, bestTri, bestInst, bestB1, bestB2))
                        return true;
                    if (cnt > 1) {
                        cur = Cursor{ lp + 1, cnt - 1 };
                        continue;
                    }
                } else {
                    for (uint32_t k = lp; k < lp + cnt; ++k)
                        if (blasTriangle<ANY>(sc, r, k, curInst, tmax, best, found, bestTri, bestInst, bestB1, bestB2))
                            return true;
                }
            } else {
                // TLAS leaf: one or more instances [lp, lp + cnt).
                for (uint32_t j = cnt - 1; j >= 1; --j) stack.push(instanceEntry(lp + j));
                pendingInst = true;
                pendingIdx = lp;
                continue;
            }
        }
        have = false;
        while (true) {
            if (inBlas && stack.sp == blasBase) {
                inBlas = false;
                r = setupRayBox(o.x, o.y, o.z, d.x, d.y, d.z, tmin);
            }
            if (stack.empty()) break;
            const uint2 e = stack.pop();
            if (entryCode(e) == kEntryInstance) {
                pendingInst = true;
                pendingIdx = e.x & kEntryIndexMask;
                break;
            }
            if (__uint_as_float(e.y) <= best * kLimitWiden) {
                cur = entryCursor(sc.nodes, e);
                have = true;
                break;
            }
        }
        if (!have && !pendingInst) break;
    }
    if (stack.overflow) overflow = true;
    if (ANY || !found) return false;
    out.t = best;
    out.b1 = bestB1;
    out.b2 = bestB2;
    out.tri = bestTri;
    out.inst = bestInst;
    return true;
}

// The same walk over the four-wide trees (host/bvh_wide.h): the TLAS and
// every BLAS collapsed from their binary trees, whose leaves they share.  A
// step tests the four children of the current wide node (trace_core.h
// wideTest), goes on with the nearest hit child and pushes the other hit
// children farthest first; a leaf step is as in traceScene.  The hits are
// those of traceScene: the box tests are conservative and the tie rule is
// the same, so neither tree shape nor visiting order changes a result.
// Camera-ray packets keep the binary walk (traceScenePacket).  anyLane:
// this lane's ray is an any-hit query whatever ANY says (lanes of one call
// may mix closest-hit and any-hit rays).
template <bool ANY, int S, int BLOCK, bool TRI_STEP = true>
__device__ bool traceSceneWide(const SceneView& sc, uint32_t tlasWideRoot, uint32_t instanceCount,
                               float3 o, float3 d, float tmin, float tmax, uint2* ldsLane,
                               SceneHit& out, bool& overflow, bool anyLane = false) {
    if (instanceCount == 0) return false;
    RayK r = setupRayBox(o.x, o.y, o.z, d.x, d.y, d.z, tmin);
    WideRows rows = wideRows(r);
    uint2 spill[kSpill];
    ShortStack<S, BLOCK> stack(ldsLane, spill);
    bool inBlas = false;
    int blasBase = 0;
    uint32_t curInst = 0;
    float best = tmax;
    bool found = false;
    uint32_t bestTri = kNoHit, bestInst = kNoHit;
    float bestB1 = 0.0f, bestB2 = 0.0f;

    Cursor cur{ tlasWideRoot, 0u };
    bool have = true;
    bool pendingInst = false;
    uint32_t pendingIdx = 0;

    while (true) {
        if (pendingInst) {
            // Enter an instance: object-space ray, BLAS root (a wide node).
            pendingInst = false;
            const TraceInstance* tp = sc.traceInst + pendingIdx;
            TraceInstance ti;
            ti.row0 = ldGlobal4(&tp->row0);
            ti.row1 = ldGlobal4(&tp->row1);
            ti.row2 = ldGlobal4(&tp->row2);
            const float4 tail = ldGlobal4(reinterpret_cast<const float4*>(&tp->blasRoot));
            float3 oo, dd;
            toObject(ti, o.x, o.y, o.z, d.x, d.y, d.z, oo, dd);
            r = setupRay(oo.x, oo.y, oo.z, dd.x, dd.y, dd.z, tmin);
            rows = wideRows(r);
            inBlas = true;
            blasBase = stack.sp;
            curInst = pendingIdx;
            cur = Cursor{ __float_as_uint(tail.z), 0u };  // TraceInstance::wideRoot
            have = true;
        }
        if (have) {
            const uint32_t lp = cur.index;
            const uint32_t cnt = cur.count;
            if (cnt == 0) {
                WideStep ws = wideTest<false>(sc.wide, lp, r, rows, best * kLimitWiden);
                const int hits = wideSort(ws);
                if (hits > 0) {
                    if (hits > 3) stack.push(make_uint2(ws.entry[3], ws.key[3]));
                    if (hits > 2) stack.push(make_uint2(ws.entry[2], ws.key[2]));
                    if (hits > 1) stack.push(make_uint2(ws.entry[1], ws.key[1]));
                    cur = wideEntryCursor(sc.nodes, make_uint2(ws.entry[0], ws.key[0]));
                    continue;
                }
            } else if (inBlas) {
                if constexpr (TRI_STEP) {
                    if (blasTriangle<ANY>(sc, r, lp, curInst, tmax, best, found, bestTri, bestInst, bestB1, bestB2, anyLane))
                        return true;
                    if (cnt > 1) {
                        cur = Cursor{ lp + 1, cnt - 1 };
                        continue;
                    }
                } else {
                    for (uint32_t k = lp; k < lp + cnt; ++k)
                        if (blasTriangle<ANY>(sc, r, k, curInst, tmax, best, found, bestTri, bestInst, bestB1, bestB2, anyLane))
                            return true;
                }
            } else {
                // TLAS leaf: one or more instances [lp, lp + cnt).
                for (uint32_t j = cnt - 1; j >= 1; --j) stack.push(instanceEntry(lp + j));
                pendingInst = true;
                pendingIdx = lp;
                continue;
            }
        }
        have = false;
        while (true) {
            if (inBlas && stack.sp == blasBase) {
                inBlas = false;
                r = setupRayBox(o.x, o.y, o.z, d.x, d.y, d.z, tmin);
                rows = wideRows(r);
            }
            if (stack.empty()) break;
            const uint2 e = stack.pop();
            if (entryCode(e) == kEntryInstance) {
                pendingInst = true;
                pendingIdx = e.x & kEntryIndexMask;
                break;
            }
            if (__uint_as_float(e.y) <= best * kLimitWiden) {
                cur = wideEntryCursor(sc.nodes, e);
                have = true;
                break;
            }
        }
        if (!have && !pendingInst) break;
    }
    if (stack.overflow) overflow = true;
    if (ANY || anyLane || !found) return false;
    out.t = best;
    out.b1 = bestB1;
    out.b2 = bestB2;
    out.tri = bestTri;
    out.inst = bestInst;
    return true;
}

// ------------------------------------------------------------ wave packets --
// Closest hit for rays that arrive in coherent groups (camera rays: a wave
// holds an 8 x 8 pixel tile).  The wave walks the trees together: it enters a
// node when the box test passes for some lane, so every node that a lane's own
// traversal would enter is entered; culling stays per lane (each lane tests
// every box against its own limit).  Node, triangle and instance reads are
// wave-uniform and go through the scalar data cache; each lane keeps its own
// closest hit with the same tie rule (lowest instance, then triangle), so the
// hits equal those of traceScene whatever order the nodes are visited in.
// The stack is one wave-uniform list of node indices (and instances to
// enter) in LDS; a popped node is loaded again and tested against every
// lane's current limit.  Must be called by every lane of the wave that is
// still active in the caller (the control flow inside is wave-uniform).
constexpr int kPacketStack = kSpill + 8;  // the scene's stack depth is bounded by this on the host
constexpr uint32_t kPacketInstance = 0x80000000u;

__device__ __forceinline__ uint32_t waveUniform(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }

// 16-byte read of scene data that no launch writes, through the scalar data
// cache when the address is wave-uniform.
__device__ __forceinline__ float4 ldScene4(const float4* p) {
    const nativeF32x4 v = *(const __attribute__((address_space(4))) nativeF32x4*)p;
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ Node loadNodeUniform(const Node* __restrict__ nodes, uint32_t i) {
    return Node{ ldScene4(&nodes[i].a), ldScene4(&nodes[i].b) };
}

__device__ bool traceScenePacket(const SceneView& sc, uint32_t tlasRoot, uint32_t instanceCount, float3 o,
                                 float3 d, float tmin, float tmax, uint32_t* waveStack, SceneHit& out,
                                 bool& overflow) {
    if (instanceCount == 0) return false;
    RayK r = setupRayBox(o.x, o.y, o.z, d.x, d.y, d.z, tmin);
    float best = tmax;
    bool found = false;
    uint32_t bestTri = kNoHit, bestInst = kNoHit;
    float bestB1 = 0.0f, bestB2 = 0.0f;
    uint32_t sp = 0, blasBase = 0, curInst = 0;
    bool inBlas = false;
    float te;
    Node cn = loadNodeUniform(sc.nodes, waveUniform(tlasRoot));  // node the wave stands at
    bool have = __ballot(slab(cn, r, best * kLimitWiden, te)) != 0;
    uint32_t pendingInst = kNoHit;

    while (true) {
        if (pendingInst != kNoHit) {
            // Enter an instance: object-space rays, BLAS root.
            const TraceInstance* tp = sc.traceInst + pendingInst;
            TraceInstance ti;
            ti.row0 = ldScene4(&tp->row0);
            ti.row1 = ldScene4(&tp->row1);
            ti.row2 = ldScene4(&tp->row2);
            ti.blasRoot = waveUniform(__float_as_uint(ldScene4(reinterpret_cast<const float4*>(&tp->blasRoot)).x));
            float3 oo, dd;
            toObject(ti, o.x, o.y, o.z, d.x, d.y, d.z, oo, dd);
            r = setupRay(oo.x, oo.y, oo.z, dd.x, dd.y, dd.z, tmin);
            inBlas = true;
            blasBase = sp;
            curInst = pendingInst;
            pendingInst = kNoHit;
            cn = loadNodeUniform(sc.nodes, ti.blasRoot);
            have = __ballot(slab(cn, r, best * kLimitWiden, te)) != 0;
        }
        if (have) {
            const uint32_t lp = waveUniform(nodeLeft(cn));
            const uint32_t cnt = waveUniform(nodeCount(cn));
            if (cnt == 0) {
                const Node c0 = loadNodeUniform(sc.nodes, lp);
                const Node c1 = loadNodeUniform(sc.nodes, lp + 1);
                const float lim = best * kLimitWiden;
                float t0, t1;
                const bool h0 = slab(c0, r, lim, t0);
                const bool h1 = slab(c1, r, lim, t1);
                const uint64_t m0 = __ballot(h0), m1 = __ballot(h1);
                if (m0 != 0 && m1 != 0) {
                    // Both children: the one nearer for more lanes first.
                    const uint64_t near0 = __ballot(h0 && (!h1 || t0 <= t1));
                    const uint64_t near1 = __ballot(h1 && (!h0 || t1 < t0));
                    const bool first0 = __popcll(near0) >= __popcll(near1);
                    if (sp < uint32_t(kPacketStack)) {
                        waveStack[sp] = first0 ? lp + 1 : lp;
                        ++sp;
                    } else {
                        overflow = true;
                    }
                    cn = first0 ? c0 : c1;
                    continue;
                }
                if (m0 != 0 || m1 != 0) {
                    cn = m0 != 0 ? c0 : c1;
                    continue;
                }
            } else if (inBlas) {
                for (uint32_t k = lp; k < lp + cnt; ++k) {
                    const float4* tv = sc.triVerts + 3 * size_t(k);
                    testTriangle<false>(r, ldScene4(tv), ldScene4(tv + 1), ldScene4(tv + 2), k, curInst, tmax, best,
                                        found, bestTri, bestInst, bestB1, bestB2);
                }
            } else {
                // TLAS leaf: instances [lp, lp + cnt), the first entered now.
                for (uint32_t j = cnt - 1; j >= 1; --j) {
                    if (sp < uint32_t(kPacketStack)) {
                        waveStack[sp] = kPacketInstance | (lp + j);
                        ++sp;
                    } else {
                        overflow = true;
                    }
                }
                pendingInst = lp;
                continue;
            }
        }
        have = false;
        while (true) {
            if (inBlas && sp == blasBase) {
                inBlas = false;
                r = setupRayBox(o.x, o.y, o.z, d.x, d.y, d.z, tmin);
            }
            if (sp == 0) break;
            --sp;
            const uint32_t e = waveUniform(waveStack[sp]);
            if (e & kPacketInstance) {
                pendingInst = e & ~kPacketInstance;
                break;
            }
            cn = loadNodeUniform(sc.nodes, e);
            if (__ballot(slab(cn, r, best * kLimitWiden, te)) != 0) {
                have = true;
                break;
            }
        }
        if (!have && pendingInst == kNoHit) break;
    }
    if (!found) return false;
    out.t = best;
    out.b1 = bestB1;
    out.b2 = bestB2;
    out.tri = bestTri;
    out.inst = bestInst;
    return true;
}

} // namespace dev
} // namespace mrt
