// Device-side BVH traversal building blocks for gfx950 (wave64).
//
// Node layout: the reference's 32-byte BvhNode (core/src/accel/Bvh.hpp:15-26)
// read as two float4: a = (lo.xyz, leftOrPrim bits), b = (hi.xyz, packed bits),
// packed = (count << 2) | axis, count == 0 -> internal node whose children sit
// at leftOrPrim and leftOrPrim + 1.  The host places node arrays so that every
// child pair starts on a 64-byte boundary: the two children a step fetches
// then share one cache line.
//
// Triangles are read from a de-indexed, leaf-ordered array of 3 float4 per
// triangle (v0.xyz + id bits, v1.xyz, v2.xyz): one 48-byte contiguous read per
// test instead of the reference's index fetch followed by three dependent
// position fetches (core/shaders/intersect.glsl:21-25).
//
// Intersection is the watertight test of Woop, Benthin and Wald (JCGT 2013).
// The box test uses IEEE minNum/maxNum (a NaN from 0*inf drops out) and widens
// far distances by 2^-20 and the hit limit by 2^-16, so rounding never culls a
// box that holds a hit at t <= limit; together with a (t, id) tie-break this
// makes traversal results equal brute force exactly.  The box test's
// reciprocal direction comes from v_rcp_f32 (1 ulp): with the subtraction
// and product roundings a slab distance is off by at most ~2^-22 relative,
// well inside the widening.  The triangle test's constants use IEEE
// division, and all its expressions are evaluated without FMA contraction,
// so that the CPU oracle (g++ -ffp-contract=off) reproduces t bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mrt {
namespace dev {

constexpr float kFarWiden = 1.0f + 1.0f / 1048576.0f;
constexpr float kLimitWiden = 1.0f + 1.0f / 65536.0f;
constexpr uint32_t kNoHit = 0xFFFFFFFFu;

struct Node {
    float4 a;  // lo.xyz, leftOrPrim
    float4 b;  // hi.xyz, packed
};

// 16-byte load through a pointer marked as global memory.  Pointers the
// compiler cannot place (members of structures read from memory, such as a
// batch entry's scene view) otherwise compile to FLAT instructions, which
// also count against the LDS wait counter: every wait on the LDS traversal
// stack would then wait on the node and triangle fetches in flight as well.
// (The load goes through the native vector type, which has no constructor
// to bind to an address-space-qualified object.)
typedef float nativeF32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 ldGlobal4(const float4* p) {
    const nativeF32x4 v = *(const __attribute__((address_space(1))) nativeF32x4*)p;
    return make_float4(v.x, v.y, v.z, v.w);
}

__device__ __forceinline__ Node loadNode(const Node* __restrict__ nodes, uint32_t i) {
    return Node{ ldGlobal4(&nodes[i].a), ldGlobal4(&nodes[i].b) };
}
__device__ __forceinline__ uint32_t nodeLeft(const Node& n) { return __float_as_uint(n.a.w); }
__device__ __forceinline__ uint32_t nodeCount(const Node& n) { return __float_as_uint(n.b.w) >> 2; }

// What traversal needs of a node whose box it has entered: the first child
// (inner node, count 0) or the first primitive and the primitive count.
struct Cursor {
    uint32_t index, count;
};
__device__ __forceinline__ Cursor cursorOf(const Node& n) { return Cursor{ nodeLeft(n), nodeCount(n) }; }

// Stack entries carry the cursor of the node they stand for, so a pop needs
// no node fetch: x = (code << 27) | index, y = entry distance bits, with
//   code 0       inner node, index = first child (children at index, index + 1)
//   code 1..29   leaf of `code` primitives starting at index
//   code 30      instance to enter (two-level traversal), index = instance
//   code 31      leaf of 30 or more primitives, index = the node (fetched again)
// Node, primitive and instance indices stay below 2^27 (checked on the host).
constexpr uint32_t kEntryIndexBits = 27;
constexpr uint32_t kEntryIndexMask = (1u << kEntryIndexBits) - 1u;
constexpr uint32_t kEntryInstance = 30;
constexpr uint32_t kEntryRefetch = 31;
constexpr uint32_t kEntryMaxLeaf = 29;

__device__ __forceinline__ uint2 nodeEntry(const Node& n, uint32_t nodeIdx, float tEnter) {
    const uint32_t cnt = nodeCount(n);
    const uint32_t x = cnt <= kEntryMaxLeaf ? (cnt << kEntryIndexBits) | nodeLeft(n)
                                            : (kEntryRefetch << kEntryIndexBits) | nodeIdx;
    return make_uint2(x, __float_as_uint(tEnter));
}
__device__ __forceinline__ uint2 instanceEntry(uint32_t inst) {
    return make_uint2((kEntryInstance << kEntryIndexBits) | inst, 0u);
}
__device__ __forceinline__ uint32_t entryCode(uint2 e) { return e.x >> kEntryIndexBits; }
// Cursor of a node entry (code != kEntryInstance).
__device__ __forceinline__ Cursor entryCursor(const Node* __restrict__ nodes, uint2 e) {
    const uint32_t code = entryCode(e), idx = e.x & kEntryIndexMask;
    if (code == kEntryRefetch) return cursorOf(loadNode(nodes, idx));
    return Cursor{ idx, code };
}

// Per-ray constants: origin, reciprocal direction and the shear of the
// watertight test.
struct RayK {
    float ox, oy, oz;
    float ix, iy, iz;
    float sx, sy, sz;
    int kx, ky, kz;
    float tmin;
};

__device__ __forceinline__ float pick3(float x, float y, float z, int k) {
    return k == 0 ? x : (k == 1 ? y : z);
}

// Box-test constants only (origin, reciprocal direction, tmin): all that the
// top level of a two-level traversal needs, whose leaves are instances; the
// shear of the triangle test is left at a dummy identity.
__device__ __forceinline__ RayK setupRayBox(float ox, float oy, float oz, float dx, float dy,
                                            float dz, float tmin) {
    RayK r;
    r.ox = ox; r.oy = oy; r.oz = oz;
    r.ix = __builtin_amdgcn_rcpf(dx);
    r.iy = __builtin_amdgcn_rcpf(dy);
    r.iz = __builtin_amdgcn_rcpf(dz);
    r.sx = 0.0f; r.sy = 0.0f; r.sz = 1.0f;
    r.kx = 0; r.ky = 1; r.kz = 2;
    r.tmin = tmin;
    return r;
}

__device__ __forceinline__ RayK setupRay(float ox, float oy, float oz, float dx, float dy,
                                         float dz, float tmin) {
#pragma clang fp contract(off)
    RayK r = setupRayBox(ox, oy, oz, dx, dy, dz, tmin);
    const float ax = fabsf(dx), ay = fabsf(dy), az = fabsf(dz);
    int kz = 0;
    if (ay > ax) kz = 1;
    if (az > (kz == 0 ? ax : ay)) kz = 2;
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    const float dkz = pick3(dx, dy, dz, kz);
    if (dkz < 0.0f) { const int t = kx; kx = ky; ky = t; }
    r.kx = kx; r.ky = ky; r.kz = kz;
    r.sx = pick3(dx, dy, dz, kx) / dkz;
    r.sy = pick3(dx, dy, dz, ky) / dkz;
    r.sz = 1.0f / dkz;
    return r;
}

// Conservative slab test.  Returns true when the box may hold a hit with
// t <= limit; tEnter receives the (clamped at 0) entry distance.
//
// The near and far plane of each axis are chosen from the sign of the
// reciprocal direction, so the only NaN (0 * inf, a ray parallel to and lying
// on a box face) lands in the max over near distances or the min over far
// distances, where IEEE maxNum/minNum drop it: a vacuous constraint, which is
// the right answer for a ray inside that slab.
__device__ __forceinline__ bool slab(const Node& n, const RayK& r, float limit, float& tEnter) {
#pragma clang fp contract(off)
    const bool px = r.ix >= 0.0f, py = r.iy >= 0.0f, pz = r.iz >= 0.0f;
    const float nx = ((px ? n.a.x : n.b.x) - r.ox) * r.ix, fx = ((px ? n.b.x : n.a.x) - r.ox) * r.ix;
    const float ny = ((py ? n.a.y : n.b.y) - r.oy) * r.iy, fy = ((py ? n.b.y : n.a.y) - r.oy) * r.iy;
    const float nz = ((pz ? n.a.z : n.b.z) - r.oz) * r.iz, fz = ((pz ? n.b.z : n.a.z) - r.oz) * r.iz;
    const float enter = fmaxf(fmaxf(nx, ny), fmaxf(nz, 0.0f));
    const float far = fminf(fminf(fx, fy), fz) * kFarWiden;
    tEnter = enter;
    return enter <= fminf(far, limit);
}

// Watertight ray/triangle test.  Hit iff tmin < t <= tLimit; writes t and the
// barycentric weights of v1 and v2.
__device__ __forceinline__ bool isectTri(const RayK& r, const float4& v0, const float4& v1,
                                         const float4& v2, float tLimit, float& tOut,
                                         float& b1, float& b2) {
#pragma clang fp contract(off)
    const float Ax0 = v0.x - r.ox, Ay0 = v0.y - r.oy, Az0 = v0.z - r.oz;
    const float Bx0 = v1.x - r.ox, By0 = v1.y - r.oy, Bz0 = v1.z - r.oz;
    const float Cx0 = v2.x - r.ox, Cy0 = v2.y - r.oy, Cz0 = v2.z - r.oz;
    const float Akz = pick3(Ax0, Ay0, Az0, r.kz);
    const float Bkz = pick3(Bx0, By0, Bz0, r.kz);
    const float Ckz = pick3(Cx0, Cy0, Cz0, r.kz);
    const float Ax = pick3(Ax0, Ay0, Az0, r.kx) - r.sx * Akz;
    const float Ay = pick3(Ax0, Ay0, Az0, r.ky) - r.sy * Akz;
    const float Bx = pick3(Bx0, By0, Bz0, r.kx) - r.sx * Bkz;
    const float By = pick3(Bx0, By0, Bz0, r.ky) - r.sy * Bkz;
    const float Cx = pick3(Cx0, Cy0, Cz0, r.kx) - r.sx * Ckz;
    const float Cy = pick3(Cx0, Cy0, Cz0, r.ky) - r.sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = float(double(Cx) * double(By) - double(Cy) * double(Bx));
        V = float(double(Ax) * double(Cy) - double(Ay) * double(Cx));
        W = float(double(Bx) * double(Ay) - double(By) * double(Ax));
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = (U + V) + W;
    if (det == 0.0f) return false;
    const float Az = r.sz * Akz, Bz = r.sz * Bkz, Cz = r.sz * Ckz;
    const float T = (U * Az + V * Bz) + W * Cz;
    const float rcp = 1.0f / det;
    const float t = T * rcp;
    if (!(t > r.tmin && t <= tLimit)) return false;
    tOut = t;
    b1 = V * rcp;
    b2 = W * rcp;
    return true;
}

// Short traversal stack: the top S entries of each lane live in LDS (entry s
// of lane `tid` at lds[s * BLOCK + tid], so a wave's accesses to one slot are
// conflict free); deeper entries spill to a private scratch array.  Entries
// are encoded as described at nodeEntry.
constexpr int kSpill = 64;

//
// The spill array is a separate private object from the stack's scalar state
// so that the scalars stay in registers (an aggregate holding a dynamically
// indexed array lives in scratch as a whole).
template <int S, int BLOCK>
struct ShortStack {
    static_assert((S & (S - 1)) == 0, "LDS stack depth must be a power of two");
    uint2* lds;
    uint2* spill;  // kSpill entries of private (scratch) memory
    int sp = 0;
    int base = 0;
    bool overflow = false;

    __device__ __forceinline__ ShortStack(uint2* ldsLane, uint2* spillArray) : lds(ldsLane), spill(spillArray) {}
    __device__ __forceinline__ bool empty() const { return sp == 0; }
    __device__ __forceinline__ void push(uint2 e) {
        if (sp - base == S) {
            if (base >= kSpill) {  // deeper than the spill area: drop (reported)
                overflow = true;
                return;
            }
            spill[base] = lds[(base & (S - 1)) * BLOCK];
            ++base;
        }
        lds[(sp & (S - 1)) * BLOCK] = e;
        ++sp;
    }
    // An empty LDS part is first refilled with one entry from the spill
    // area, so the pop itself is always one 64-bit LDS read.
    __device__ __forceinline__ uint2 pop() {
        if (sp == base) {
            --base;
            lds[(base & (S - 1)) * BLOCK] = spill[base];
        }
        --sp;
        return lds[(sp & (S - 1)) * BLOCK];
    }
};

} // namespace dev
} // namespace mrt
