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
__device__ __forceinline__ bool slabBox(float lx, float ly, float lz, float hx, float hy, float hz, const RayK& r,
                                        float limit, float& tEnter) {
#pragma clang fp contract(off)
    const bool px = r.ix >= 0.0f, py = r.iy >= 0.0f, pz = r.iz >= 0.0f;
    const float nx = ((px ? lx : hx) - r.ox) * r.ix, fx = ((px ? hx : lx) - r.ox) * r.ix;
    const float ny = ((py ? ly : hy) - r.oy) * r.iy, fy = ((py ? hy : ly) - r.oy) * r.iy;
    const float nz = ((pz ? lz : hz) - r.oz) * r.iz, fz = ((pz ? hz : lz) - r.oz) * r.iz;
    const float enter = fmaxf(fmaxf(nx, ny), fmaxf(nz, 0.0f));
    const float far = fminf(fminf(fx, fy), fz) * kFarWiden;
    tEnter = enter;
    return enter <= fminf(far, limit);
}
__device__ __forceinline__ bool slab(const Node& n, const RayK& r, float limit, float& tEnter) {
    return slabBox(n.a.x, n.a.y, n.a.z, n.b.x, n.b.y, n.b.z, r, limit, tEnter);
}

// ------------------------------------------------------------ wide nodes --
// The four-wide traversal tree collapsed from the binary BVH on the host
// (host/bvh_wide.h): 128 bytes per node, eight float4 rows = the four
// children's lo.x, lo.y, lo.z, hi.x, hi.y, hi.z, first node or primitive,
// code.  Codes: 0 inner (child = wide node), 1..29 a leaf of that many
// primitives (child = first primitive), kWideCodeBigLeaf a larger leaf
// (child = its binary node), kWideCodeEmpty an unused slot.  Stack entries
// use the encoding of nodeEntry: inner children with index = wide node, big
// leaves as kEntryRefetch with index = binary node.
//
// Per ray, the row offsets of each axis's near and far planes follow the
// direction's sign (the choice slab() makes per box), so a step loads the
// four children's near and far planes directly, without a select per plane.
//
// A plane's distance is one fused multiply-add, fma(plane, inv, c), with c
// per ray and axis: c = -o inv - eps for near planes and -o inv + eps for far
// planes, eps = 2^-22 |o inv|.  The one error this adds to slab()'s
// (plane - o) * inv, the absolute rounding of o * inv, is at most half an ulp
// of it, below eps (two ulps) even after c's own rounding, so near distances
// only come out smaller and far distances larger than with exact o * inv; the
// single rounding of the fma and the reciprocal's ulp are covered by the far
// widening as in slab().  The test stays conservative, which is all the
// exactness of the hits needs.  (Rays parallel to an axis can turn a plane
// distance into NaN, which maxNum / minNum drop: a vacuous constraint, never
// a cull.)
constexpr uint32_t kWideCodeBigLeaf = 0xFEu;
constexpr uint32_t kWideCodeEmpty = 0xFFu;

struct WideRows {
    uint32_t near;                       // near-plane rows of x, y, z in bytes 0, 1, 2 (far = 3 + a - (near - a))
    float cnx, cny, cnz, cfx, cfy, cfz;  // fma offsets of near / far planes
};
__device__ __forceinline__ WideRows wideRows(const RayK& r) {
    WideRows w;
    w.near = (r.ix >= 0.0f ? 0u : 3u) | (r.iy >= 0.0f ? 1u : 4u) << 8 | (r.iz >= 0.0f ? 2u : 5u) << 16;
    const float ax = -r.ox * r.ix, ay = -r.oy * r.iy, az = -r.oz * r.iz;
    const float ex = fabsf(ax) * (1.0f / 4194304.0f), ey = fabsf(ay) * (1.0f / 4194304.0f),
                ez = fabsf(az) * (1.0f / 4194304.0f);
    w.cnx = ax - ex;
    w.cny = ay - ey;
    w.cnz = az - ez;
    w.cfx = ax + ex;
    w.cfy = ay + ey;
    w.cfz = az + ez;
    return w;
}

struct WideStep {
    uint32_t key[4];    // entry distance bits of a hit child (t >= 0, -0 as +0), 0xFFFFFFFF for a miss
    uint32_t entry[4];  // stack entry codes (x) of the children, by slot
    uint32_t hitMask;
};

__device__ __forceinline__ float f4At(const float4& v, int k) {
    return k == 0 ? v.x : (k == 1 ? v.y : (k == 2 ? v.z : v.w));
}
__device__ __forceinline__ uint32_t u4At(const float4& v, int k) { return __float_as_uint(f4At(v, k)); }

// Box tests of the four children of wide node w.  FMA: plane distances as
// fused multiply-adds with the offsets of `rows` (one op per plane); else as
// slab()'s (plane - o) * inv, which keeps six fewer values live (the path
// tracer's trace kernel runs at a register budget).
template <bool FMA = true>
__device__ __forceinline__ WideStep wideTest(const float4* __restrict__ wide, uint32_t w, const RayK& r,
                                             const WideRows& rows, float limit) {
#pragma clang fp contract(off)
    const float4* p = wide + size_t(w) * 8;
    const uint32_t rx = rows.near & 0xFFu, ry = (rows.near >> 8) & 0xFFu, rz = rows.near >> 16;
    const float4 nx = ldGlobal4(p + rx), ny = ldGlobal4(p + ry), nz = ldGlobal4(p + rz);
    const float4 fx = ldGlobal4(p + (3u - rx)), fy = ldGlobal4(p + (5u - ry)), fz = ldGlobal4(p + (7u - rz));
    const float4 ch = ldGlobal4(p + 6), cd = ldGlobal4(p + 7);
    WideStep s;
    s.hitMask = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t code = u4At(cd, k);
        float tnx, tfx, tny, tfy, tnz, tfz;
        if constexpr (FMA) {
            tnx = fmaf(f4At(nx, k), r.ix, rows.cnx);
            tfx = fmaf(f4At(fx, k), r.ix, rows.cfx);
            tny = fmaf(f4At(ny, k), r.iy, rows.cny);
            tfy = fmaf(f4At(fy, k), r.iy, rows.cfy);
            tnz = fmaf(f4At(nz, k), r.iz, rows.cnz);
            tfz = fmaf(f4At(fz, k), r.iz, rows.cfz);
        } else {
            tnx = (f4At(nx, k) - r.ox) * r.ix;
            tfx = (f4At(fx, k) - r.ox) * r.ix;
            tny = (f4At(ny, k) - r.oy) * r.iy;
            tfy = (f4At(fy, k) - r.oy) * r.iy;
            tnz = (f4At(nz, k) - r.oz) * r.iz;
            tfz = (f4At(fz, k) - r.oz) * r.iz;
        }
        const float enter = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, 0.0f));  // >= 0 or -0, never NaN
        const float far = fminf(fminf(tfx, tfy), tfz) * kFarWiden;
        // bitwise & on the conditions, and selects, keep the step free of
        // branches (the compiler turns && and nested ?: into exec-mask
        // branches per child)
        const bool hit = (code != kWideCodeEmpty) & (enter <= fminf(far, limit));
        s.key[k] = hit ? (__float_as_uint(enter) & 0x7FFFFFFFu) : 0xFFFFFFFFu;
        s.hitMask |= uint32_t(hit) << k;
        const uint32_t child = u4At(ch, k);
        s.entry[k] = code == kWideCodeBigLeaf ? (kEntryRefetch << kEntryIndexBits) | child
                                              : (code << kEntryIndexBits) | child;
    }
    return s;
}

// Cursor of a stack entry of the wide walk (big leaves from the binary tree).
__device__ __forceinline__ Cursor wideEntryCursor(const Node* __restrict__ nodes, uint2 e) {
    const uint32_t code = entryCode(e), idx = e.x & kEntryIndexMask;
    if (code == kEntryRefetch) return cursorOf(loadNode(nodes, idx));
    return Cursor{ idx, code };
}

// Sorts the children nearest first, misses last (a sorting network over
// (key, entry) pairs; ties keep slot order), so that key[0..hits-1] and
// entry[0..hits-1] are the hit children in visiting order.  Returns the
// number of hits.
__device__ __forceinline__ int wideSort(WideStep& s) {
    auto cx = [&](int a, int b) {
        const bool sw = s.key[b] < s.key[a];
        const uint32_t ka = s.key[a], kb = s.key[b], ea = s.entry[a], eb = s.entry[b];
        s.key[a] = sw ? kb : ka;
        s.key[b] = sw ? ka : kb;
        s.entry[a] = sw ? eb : ea;
        s.entry[b] = sw ? ea : eb;
    };
    cx(0, 1);
    cx(2, 3);
    cx(0, 2);
    cx(1, 3);
    cx(1, 2);
    return __builtin_popcount(s.hitMask);
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
