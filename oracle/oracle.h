// Oracle: CPU restatement of the reference path-tracing hot path.
//
// TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load this library, and only as the checker
// (or as the timed CPU baseline).  The product library (libmrt.so) never links
// or calls it.
//
// Build: `make -C oracle` -> oracle/_build/liboracle.so (g++ -O3,
// -ffp-contract=off, no fast-math; every float operation is an IEEE-754
// single-precision operation in source order so results are bit-reproducible).
//
// Each function cites the reference file:line it restates.  Pinning status of
// every function (what checks it beyond this restatement) is listed in the
// header comment of oracle/oracle.cpp.
#pragma once

#include <cstddef>
#include <cstdint>

extern "C" {

// ---- 32-byte BVH node, reference core/src/accel/Bvh.hpp:15-26 -------------
struct OrNode {
    float    lo[3];
    uint32_t leftOrPrim;  // internal: left child (right = left + 1); leaf: first prim
    float    hi[3];
    uint32_t packed;      // (count << 2) | split axis; count == 0 -> internal
};

// ---- RNG: pcg4d, reference core/shaders/random.glsl:6-31 -------------------
void or_pcg4d(uint32_t v[4]);
// Advance the (px, py, frame, counter) state and draw; mirrors rng1/rng2.
float or_rng1(uint32_t state[4]);
void  or_rng2(uint32_t state[4], float out[2]);

// ---- BVH build: binned SAH, reference core/src/accel/Bvh.cpp:13-156 -------
// bounds: 6 floats per prim (lo xyz, hi xyz).  Returns the node count; writes
// at most nodeCap nodes into `nodes` and `count` prim indices into primOrder.
// Returns -1 when nodeCap is too small.
int64_t or_bvh_build(const float* bounds, uint32_t count, uint32_t maxLeaf,
                     OrNode* nodes, uint64_t nodeCap, uint32_t* primOrder);

// Triangle bounds in the reference's Aabb::expand order (Renderer.cpp:348-356).
void or_tri_bounds(const float* verts9, uint32_t triCount, float* bounds6);

// ---- intersection ----------------------------------------------------------
// Moeller-Trumbore exactly as tests/bvh_tests.cpp:19-34 (t in (1e-4, tMax)).
float or_isect_mt(const float* v0, const float* v1, const float* v2,
                  const float* o, const float* d, float tMax);
// Watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2(1) 2013).
// Returns 1 on a hit with tMin < t <= tLimit (t <= for ties); writes t, b1, b2.
int or_isect_wt(const float* v0, const float* v1, const float* v2,
                const float* o, const float* d, float tMin, float tLimit,
                float* tOut, float* b1Out, float* b2Out);

// ---- ray queries over a triangle soup -------------------------------------
// rays: n x 8 floats (ox oy oz tmin dx dy dz tmax); hits: n x (uint32 id, float t).
// id = original triangle index, 0xFFFFFFFF = miss.
//
// Brute force over all triangles in original order with the watertight test:
// the lowest id wins an exact tie in t.
void or_brute_closest(const float* verts9, uint32_t triCount, const float* rays,
                      uint64_t n, uint32_t* hits, int threads);
// Traversal of a built BVH with the same watertight test and a conservative
// slab test; tie-break on (t, original id).  Must equal brute force exactly.
void or_bvh_closest(const OrNode* nodes, const uint32_t* primOrder,
                    const float* verts9, const float* rays, uint64_t n,
                    uint32_t* hits, int threads);
// Reference traversal, tests/bvh_tests.cpp:49-73 (stack 64, right child first,
// Moeller-Trumbore, strict t < tMax).  hits as above (original ids).
void or_ref_traverse(const OrNode* nodes, const uint32_t* primOrder,
                     const float* verts9, const float* rays, uint64_t n,
                     uint32_t* hits, int threads);
// Reference GLSL order (core/shaders/intersect.glsl:10-86: near child first by
// the split axis, cull when tEnter >= tMax): counts nodes popped and triangles
// tested per ray (the algorithmic fetch count).  counts: n x 2 uint32 (may be
// NULL); hits may be NULL.  Returns total (nodes, tris) in sums[2].
void or_glsl_traverse(const OrNode* nodes, const uint32_t* primOrder,
                      const float* verts9, const float* rays, uint64_t n,
                      uint32_t* hits, uint32_t* counts, uint64_t sums[2],
                      int threads);
// Any-hit occlusion with the watertight test (t in (tMin, tMax)).
void or_brute_any(const float* verts9, uint32_t triCount, const float* rays,
                  uint64_t n, uint8_t* occluded, int threads);

int or_hardware_threads(void);
}
