// Oracle implementation: CPU restatement of the reference hot path.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h).  Compiled with g++ -O3
// -ffp-contract=off: each float expression below is evaluated as separate
// IEEE-754 single-precision operations in source order.
//
// Pinning status (what checks each function beyond this restatement; the
// tests named are under tests/):
//   or_pcg4d / or_rng1 / or_rng2   known-answer vectors computed by hand with
//                                  exact integer arithmetic in Python
//                                  (test_oracle_rng.py).
//   or_isect_mt                    hand-built rays with closed-form answers
//                                  (test_oracle_geometry.py).
//   or_isect_wt                    same closed-form rays, edge/vertex rays that
//                                  must hit exactly one of two triangles
//                                  sharing the edge (watertightness), and
//                                  agreement with or_isect_mt away from edges.
//   or_bvh_build                   topology invariants of tests/bvh_tests.cpp
//                                  :91-119 and degenerate inputs :165-188; the
//                                  SAH split choice against an exhaustive
//                                  restatement in numpy on small inputs.
//   or_brute_closest / or_bvh_closest / or_ref_traverse / or_glsl_traverse
//                                  traversal == brute force (bvh_tests.cpp
//                                  :121-163), brute force == an independent
//                                  float64 numpy intersector on small cases.
//   Exact BVH topology of the reference binary (built with libc++) is parity
//   unpinned: std::partition / std::nth_element orders are library-defined.

#include "oracle.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

// ---------------------------------------------------------------- vectors --
struct V3 {
    float x, y, z;
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
inline V3 v3(const float* p) { return { p[0], p[1], p[2] }; }
inline V3 sub(V3 a, V3 b) { return { a.x - b.x, a.y - b.y, a.z - b.z }; }
inline V3 mul(V3 a, V3 b) { return { a.x * b.x, a.y * b.y, a.z * b.z }; }
// glm::cross component order.
inline V3 cross(V3 a, V3 b) {
    return { a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y };
}
// glm::dot: (x*x' + y*y') + z*z'.
inline float dot(V3 a, V3 b) {
    const float px = a.x * b.x, py = a.y * b.y, pz = a.z * b.z;
    return (px + py) + pz;
}
// glm::min / glm::max semantics: (y < x) ? y : x.
inline float gmin(float x, float y) { return (y < x) ? y : x; }
inline float gmax(float x, float y) { return (x < y) ? y : x; }

// ------------------------------------------------------------------- AABB --
// Reference core/include/mrt/MathTypes.hpp:20-48.
struct Box {
    V3 lo{ FLT_MAX, FLT_MAX, FLT_MAX };
    V3 hi{ -FLT_MAX, -FLT_MAX, -FLT_MAX };
    void expand(V3 p) {
        lo = { gmin(lo.x, p.x), gmin(lo.y, p.y), gmin(lo.z, p.z) };
        hi = { gmax(hi.x, p.x), gmax(hi.y, p.y), gmax(hi.z, p.z) };
    }
    void expand(const Box& b) {
        lo = { gmin(lo.x, b.lo.x), gmin(lo.y, b.lo.y), gmin(lo.z, b.lo.z) };
        hi = { gmax(hi.x, b.hi.x), gmax(hi.y, b.hi.y), gmax(hi.z, b.hi.z) };
    }
    V3 extent() const { return sub(hi, lo); }
    V3 center() const {
        return { 0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z) };
    }
    bool valid() const { return lo.x <= hi.x; }
    float area() const {
        if (!valid()) return 0.0f;
        const V3 e = extent();
        const float s = (e.x * e.y + e.y * e.z) + e.z * e.x;
        return 2.0f * s;
    }
};

// ------------------------------------------------------------- BVH build --
// Restates core/src/accel/Bvh.cpp:13-156.
constexpr uint32_t kBins = 32;
constexpr float kTrav = 1.0f;
constexpr float kIsect = 1.5f;

struct Prim {
    Box box;
    V3 c;
    uint32_t index;
};

struct Builder {
    std::vector<Prim>& prims;
    std::vector<OrNode>& nodes;
    uint32_t maxLeaf;

    Box rangeBox(uint32_t b, uint32_t e) const {
        Box r;
        for (uint32_t i = b; i < e; ++i) r.expand(prims[i].box);
        return r;
    }

    static void setBox(OrNode& n, const Box& b) {
        n.lo[0] = b.lo.x; n.lo[1] = b.lo.y; n.lo[2] = b.lo.z;
        n.hi[0] = b.hi.x; n.hi[1] = b.hi.y; n.hi[2] = b.hi.z;
    }
    static void leaf(OrNode& n, uint32_t begin, uint32_t count, uint32_t axis) {
        n.leftOrPrim = begin;
        n.packed = (count << 2) | axis;
    }

    // Bvh.cpp:41-133
    void build(uint32_t ni, uint32_t begin, uint32_t end) {
        const uint32_t count = end - begin;
        const Box bounds = rangeBox(begin, end);
        setBox(nodes[ni], bounds);

        Box cb;
        for (uint32_t i = begin; i < end; ++i) cb.expand(prims[i].c);
        const V3 ce = cb.extent();
        uint32_t axis = 0;
        if (ce.y > ce.x) axis = 1;
        if (ce.z > ce[axis]) axis = 2;

        if (count <= maxLeaf || ce[axis] < 1e-8f) {
            leaf(nodes[ni], begin, count, axis);
            return;
        }

        Box binBox[kBins];
        uint32_t binCount[kBins] = {};
        const float k = float(kBins) * (1.0f - 1e-6f) / ce[axis];
        const float origin = cb.lo[axis];
        for (uint32_t i = begin; i < end; ++i) {
            const uint32_t b = static_cast<uint32_t>(k * (prims[i].c[axis] - origin));
            binBox[b].expand(prims[i].box);
            binCount[b]++;
        }

        float rightArea[kBins];
        uint32_t rightCount[kBins];
        {
            Box acc;
            uint32_t cnt = 0;
            for (int i = int(kBins) - 1; i > 0; --i) {
                acc.expand(binBox[i]);
                cnt += binCount[i];
                rightArea[i] = acc.area();
                rightCount[i] = cnt;
            }
        }
        float bestCost = kIsect * float(count);
        int bestSplit = -1;
        {
            Box acc;
            uint32_t cnt = 0;
            const float invRoot = 1.0f / std::max(bounds.area(), 1e-12f);
            for (uint32_t i = 0; i < kBins - 1; ++i) {
                acc.expand(binBox[i]);
                cnt += binCount[i];
                if (cnt == 0 || rightCount[i + 1] == 0) continue;
                const float l = acc.area() * float(cnt);
                const float r = rightArea[i + 1] * float(rightCount[i + 1]);
                const float cost = kTrav + (kIsect * invRoot) * (l + r);
                if (cost < bestCost) { bestCost = cost; bestSplit = int(i); }
            }
        }

        uint32_t mid;
        if (bestSplit < 0) {
            if (count <= maxLeaf * 2) {
                leaf(nodes[ni], begin, count, axis);
                return;
            }
            mid = begin + count / 2;
            std::nth_element(prims.begin() + begin, prims.begin() + mid, prims.begin() + end,
                             [axis](const Prim& a, const Prim& b) { return a.c[axis] < b.c[axis]; });
        } else {
            auto it = std::partition(prims.begin() + begin, prims.begin() + end,
                                     [&](const Prim& p) {
                                         const uint32_t b = static_cast<uint32_t>(k * (p.c[axis] - origin));
                                         return b <= static_cast<uint32_t>(bestSplit);
                                     });
            mid = static_cast<uint32_t>(it - prims.begin());
            if (mid == begin || mid == end) mid = begin + count / 2;
        }

        const uint32_t left = static_cast<uint32_t>(nodes.size());
        nodes.emplace_back();
        nodes.emplace_back();
        nodes[ni].leftOrPrim = left;
        nodes[ni].packed = axis;
        build(left, begin, mid);
        build(left + 1, mid, end);
    }
};

// ------------------------------------------------------------ intersection --
// tests/bvh_tests.cpp:19-34 (== core/shaders/intersect.glsl:21-45).
float isectMT(V3 a, V3 b, V3 c, V3 o, V3 d, float tMin, float tMax, float* bu, float* bv) {
    const V3 e1 = sub(b, a);
    const V3 e2 = sub(c, a);
    const V3 p = cross(d, e2);
    const float det = dot(e1, p);
    if (std::fabs(det) < 1e-12f) return -1.0f;
    const float invDet = 1.0f / det;
    const V3 s = sub(o, a);
    const float u = dot(s, p) * invDet;
    if (u < 0.0f || u > 1.0f) return -1.0f;
    const V3 q = cross(s, e1);
    const float v = dot(d, q) * invDet;
    if (v < 0.0f || u + v > 1.0f) return -1.0f;
    const float t = dot(e2, q) * invDet;
    if (!(t > tMin && t < tMax)) return -1.0f;
    if (bu) { *bu = u; *bv = v; }
    return t;
}

// Watertight test, Woop/Benthin/Wald 2013 (JCGT 2(1):65-82), section 3.
struct WtRay {
    V3 o;
    int kx, ky, kz;
    float sx, sy, sz;
};
WtRay wtSetup(V3 o, V3 d) {
    WtRay r;
    r.o = o;
    const float ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z);
    int kz = 0;
    if (ay > ax) kz = 1;
    if (az > (kz == 0 ? ax : ay)) kz = 2;
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    if (d[kz] < 0.0f) std::swap(kx, ky);
    r.kx = kx; r.ky = ky; r.kz = kz;
    r.sx = d[kx] / d[kz];
    r.sy = d[ky] / d[kz];
    r.sz = 1.0f / d[kz];
    return r;
}
bool wtIsect(const WtRay& r, V3 v0, V3 v1, V3 v2, float tMin, float tLimit,
             float* tOut, float* b1, float* b2) {
    const V3 A = sub(v0, r.o), B = sub(v1, r.o), C = sub(v2, r.o);
    const float Akz = A[r.kz], Bkz = B[r.kz], Ckz = C[r.kz];
    const float Ax = A[r.kx] - r.sx * Akz, Ay = A[r.ky] - r.sy * Akz;
    const float Bx = B[r.kx] - r.sx * Bkz, By = B[r.ky] - r.sy * Bkz;
    const float Cx = C[r.kx] - r.sx * Ckz, Cy = C[r.ky] - r.sy * Ckz;
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
    if (!(t > tMin && t <= tLimit)) return false;
    *tOut = t;
    *b1 = V * rcp;
    *b2 = W * rcp;
    return true;
}

// --------------------------------------------------------------- slab tests --
// tests/bvh_tests.cpp:36-46.
bool slabRef(const OrNode& n, V3 o, V3 inv, float tMax) {
    const V3 lo = v3(n.lo), hi = v3(n.hi);
    const V3 t0 = mul(sub(lo, o), inv), t1 = mul(sub(hi, o), inv);
    const V3 mn = { gmin(t0.x, t1.x), gmin(t0.y, t1.y), gmin(t0.z, t1.z) };
    const V3 mx = { gmax(t0.x, t1.x), gmax(t0.y, t1.y), gmax(t0.z, t1.z) };
    // std::max({a,b,c,d}) / std::min({...}) fold left to right.
    float enter = mn.x;
    enter = std::max(enter, mn.y);
    enter = std::max(enter, mn.z);
    enter = std::max(enter, 0.0f);
    float exit = mx.x;
    exit = std::min(exit, mx.y);
    exit = std::min(exit, mx.z);
    exit = std::min(exit, tMax);
    return enter <= exit;
}
// core/shaders/intersect.glsl:10-18: entry distance or tMax on a miss.
float slabGlsl(const OrNode& n, V3 o, V3 inv, float tMax) {
    const V3 lo = v3(n.lo), hi = v3(n.hi);
    const V3 t0 = mul(sub(lo, o), inv), t1 = mul(sub(hi, o), inv);
    const float mnx = std::fmin(t0.x, t1.x), mny = std::fmin(t0.y, t1.y), mnz = std::fmin(t0.z, t1.z);
    const float mxx = std::fmax(t0.x, t1.x), mxy = std::fmax(t0.y, t1.y), mxz = std::fmax(t0.z, t1.z);
    const float enter = std::fmax(std::fmax(mnx, mny), std::fmax(mnz, 0.0f));
    const float exit = std::fmin(std::fmin(mxx, mxy), std::fmin(mxz, tMax));
    return enter <= exit ? enter : tMax;
}
// Conservative slab test used with the (t, id) tie-break.  Near/far planes
// are picked from the sign of the reciprocal direction, so the only NaN
// (0 * inf: a ray parallel to and lying on a box face) enters the max over
// near distances or the min over far distances, where IEEE maxNum/minNum drop
// it (a vacuous constraint).  Far distances are widened by 2^-20 and the
// limit by 2^-16 so rounding never culls a box holding a hit at t <= limit.
constexpr float kFarWiden = 1.0f + 1.0f / 1048576.0f;
constexpr float kLimitWiden = 1.0f + 1.0f / 65536.0f;
bool slabSafe(const OrNode& n, V3 o, V3 inv, float limit, float* tEnter) {
    const bool px = inv.x >= 0.0f, py = inv.y >= 0.0f, pz = inv.z >= 0.0f;
    const float nx = ((px ? n.lo[0] : n.hi[0]) - o.x) * inv.x, fx = ((px ? n.hi[0] : n.lo[0]) - o.x) * inv.x;
    const float ny = ((py ? n.lo[1] : n.hi[1]) - o.y) * inv.y, fy = ((py ? n.hi[1] : n.lo[1]) - o.y) * inv.y;
    const float nz = ((pz ? n.lo[2] : n.hi[2]) - o.z) * inv.z, fz = ((pz ? n.hi[2] : n.lo[2]) - o.z) * inv.z;
    const float enter = std::fmax(std::fmax(nx, ny), std::fmax(nz, 0.0f));
    const float far = std::fmin(std::fmin(fx, fy), fz) * kFarWiden;
    *tEnter = enter;
    return enter <= std::fmin(far, limit);
}

struct Ray {
    V3 o, d;
    float tmin, tmax;
};
inline Ray loadRay(const float* r) {
    return { { r[0], r[1], r[2] }, { r[4], r[5], r[6] }, r[3], r[7] };
}
inline V3 invDir(V3 d) { return { 1.0f / d.x, 1.0f / d.y, 1.0f / d.z }; }
inline V3 vert(const float* verts9, uint32_t tri, int k) { return v3(verts9 + size_t(tri) * 9 + k * 3); }

template <typename F>
void parallelFor(uint64_t n, int threads, F f) {
    if (threads <= 0) threads = or_hardware_threads();
    if (threads <= 1 || n < 256) {
        f(uint64_t(0), n);
        return;
    }
    std::vector<std::thread> pool;
    const uint64_t chunk = (n + uint64_t(threads) - 1) / uint64_t(threads);
    for (int t = 0; t < threads; ++t) {
        const uint64_t b = uint64_t(t) * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        pool.emplace_back([=] { f(b, e); });
    }
    for (auto& th : pool) th.join();
}

inline uint32_t fbits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float bitsf(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

} // namespace

extern "C" {

int or_hardware_threads(void) {
    const unsigned h = std::thread::hardware_concurrency();
    return h ? int(h) : 1;
}

// random.glsl:6-13
void or_pcg4d(uint32_t v[4]) {
    for (int i = 0; i < 4; ++i) v[i] = v[i] * 1664525u + 1013904223u;
    v[0] += v[1] * v[3]; v[1] += v[2] * v[0]; v[2] += v[0] * v[1]; v[3] += v[1] * v[2];
    for (int i = 0; i < 4; ++i) v[i] ^= v[i] >> 16u;
    v[0] += v[1] * v[3]; v[1] += v[2] * v[0]; v[2] += v[0] * v[1]; v[3] += v[1] * v[2];
}
// random.glsl:22-31
float or_rng1(uint32_t s[4]) {
    s[3]++;
    uint32_t v[4] = { s[0], s[1], s[2], s[3] };
    or_pcg4d(v);
    return float(v[0]) * (1.0f / 4294967296.0f);
}
void or_rng2(uint32_t s[4], float out[2]) {
    s[3]++;
    uint32_t v[4] = { s[0], s[1], s[2], s[3] };
    or_pcg4d(v);
    out[0] = float(v[0]) * (1.0f / 4294967296.0f);
    out[1] = float(v[1]) * (1.0f / 4294967296.0f);
}

// Bvh.cpp:137-156
int64_t or_bvh_build(const float* bounds, uint32_t count, uint32_t maxLeaf,
                     OrNode* nodesOut, uint64_t nodeCap, uint32_t* primOrder) {
    if (count == 0) return 0;
    std::vector<Prim> prims(count);
    for (uint32_t i = 0; i < count; ++i) {
        Box b;
        b.lo = v3(bounds + size_t(i) * 6);
        b.hi = v3(bounds + size_t(i) * 6 + 3);
        prims[i] = { b, b.center(), i };
    }
    std::vector<OrNode> nodes;
    nodes.reserve(size_t(2) * count);
    nodes.emplace_back();
    Builder bld{ prims, nodes, std::max(1u, maxLeaf) };
    bld.build(0, 0, count);
    if (nodes.size() > nodeCap) return -1;
    std::memcpy(nodesOut, nodes.data(), nodes.size() * sizeof(OrNode));
    for (uint32_t i = 0; i < count; ++i) primOrder[i] = prims[i].index;
    return int64_t(nodes.size());
}

// Renderer.cpp:348-356
void or_tri_bounds(const float* verts9, uint32_t triCount, float* out) {
    for (uint32_t t = 0; t < triCount; ++t) {
        Box b;
        for (int k = 0; k < 3; ++k) b.expand(vert(verts9, t, k));
        float* o = out + size_t(t) * 6;
        o[0] = b.lo.x; o[1] = b.lo.y; o[2] = b.lo.z;
        o[3] = b.hi.x; o[4] = b.hi.y; o[5] = b.hi.z;
    }
}

float or_isect_mt(const float* v0, const float* v1, const float* v2,
                  const float* o, const float* d, float tMax) {
    return isectMT(v3(v0), v3(v1), v3(v2), v3(o), v3(d), 1e-4f, tMax, nullptr, nullptr);
}

int or_isect_wt(const float* v0, const float* v1, const float* v2,
                const float* o, const float* d, float tMin, float tLimit,
                float* tOut, float* b1Out, float* b2Out) {
    const WtRay r = wtSetup(v3(o), v3(d));
    return wtIsect(r, v3(v0), v3(v1), v3(v2), tMin, tLimit, tOut, b1Out, b2Out) ? 1 : 0;
}

void or_brute_closest(const float* verts9, uint32_t triCount, const float* rays,
                      uint64_t n, uint32_t* hits, int threads) {
    parallelFor(n, threads, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; ++i) {
            const Ray ray = loadRay(rays + i * 8);
            const WtRay wr = wtSetup(ray.o, ray.d);
            float best = ray.tmax;
            uint32_t id = 0xFFFFFFFFu;
            for (uint32_t t = 0; t < triCount; ++t) {
                float th, b1, b2;
                if (wtIsect(wr, vert(verts9, t, 0), vert(verts9, t, 1), vert(verts9, t, 2),
                            ray.tmin, best, &th, &b1, &b2) && th < best) {
                    best = th;
                    id = t;
                }
            }
            hits[i * 2 + 0] = id;
            hits[i * 2 + 1] = fbits(id == 0xFFFFFFFFu ? ray.tmax : best);
        }
    });
}

void or_bvh_closest(const OrNode* nodes, const uint32_t* primOrder, const float* verts9,
                    const float* rays, uint64_t n, uint32_t* hits, int threads) {
    parallelFor(n, threads, [&](uint64_t b, uint64_t e) {
        std::vector<uint32_t> stack;
        stack.reserve(256);
        for (uint64_t i = b; i < e; ++i) {
            const Ray ray = loadRay(rays + i * 8);
            const WtRay wr = wtSetup(ray.o, ray.d);
            const V3 inv = invDir(ray.d);
            float best = ray.tmax;
            uint32_t id = 0xFFFFFFFFu;
            bool hit = false;
            stack.clear();
            stack.push_back(0);
            while (!stack.empty()) {
                const OrNode& nd = nodes[stack.back()];
                stack.pop_back();
                const float limit = hit ? best * kLimitWiden : ray.tmax * kLimitWiden;
                float te;
                if (!slabSafe(nd, ray.o, inv, limit, &te)) continue;
                const uint32_t cnt = nd.packed >> 2;
                if (cnt == 0) {
                    stack.push_back(nd.leftOrPrim);
                    stack.push_back(nd.leftOrPrim + 1);
                    continue;
                }
                for (uint32_t k = 0; k < cnt; ++k) {
                    const uint32_t orig = primOrder[nd.leftOrPrim + k];
                    float th, b1, b2;
                    if (!wtIsect(wr, vert(verts9, orig, 0), vert(verts9, orig, 1),
                                 vert(verts9, orig, 2), ray.tmin, best, &th, &b1, &b2))
                        continue;
                    if (th < best || (hit && th == best && orig < id)) {
                        best = th;
                        id = orig;
                        hit = true;
                    }
                }
            }
            hits[i * 2 + 0] = id;
            hits[i * 2 + 1] = fbits(hit ? best : ray.tmax);
        }
    });
}

void or_ref_traverse(const OrNode* nodes, const uint32_t* primOrder, const float* verts9,
                     const float* rays, uint64_t n, uint32_t* hits, int threads) {
    parallelFor(n, threads, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; ++i) {
            const Ray ray = loadRay(rays + i * 8);
            const V3 inv = invDir(ray.d);
            float tMax = ray.tmax;
            int64_t hitTri = -1;
            uint32_t stack[64];
            int sp = 0;
            stack[sp++] = 0;
            while (sp > 0) {
                const OrNode& nd = nodes[stack[--sp]];
                if (!slabRef(nd, ray.o, inv, tMax)) continue;
                const uint32_t cnt = nd.packed >> 2;
                if (cnt == 0) {
                    if (sp + 2 > 64) { hitTri = -2; break; }
                    stack[sp++] = nd.leftOrPrim;
                    stack[sp++] = nd.leftOrPrim + 1;
                } else {
                    for (uint32_t k = 0; k < cnt; ++k) {
                        const uint32_t orig = primOrder[nd.leftOrPrim + k];
                        const float t = isectMT(vert(verts9, orig, 0), vert(verts9, orig, 1),
                                                vert(verts9, orig, 2), ray.o, ray.d,
                                                ray.tmin, tMax, nullptr, nullptr);
                        if (t > 0.0f) { tMax = t; hitTri = int64_t(nd.leftOrPrim + k); }
                    }
                }
            }
            hits[i * 2 + 0] = hitTri >= 0 ? primOrder[hitTri] : 0xFFFFFFFFu;
            hits[i * 2 + 1] = fbits(tMax);
        }
    });
}

void or_glsl_traverse(const OrNode* nodes, const uint32_t* primOrder, const float* verts9,
                      const float* rays, uint64_t n, uint32_t* hits, uint32_t* counts,
                      uint64_t sums[2], int threads) {
    if (threads <= 0) threads = or_hardware_threads();
    std::vector<uint64_t> partial(size_t(threads) * 2 + 2, 0);
    int slot = 0;
    std::vector<std::pair<uint64_t, uint64_t>> ranges;
    const uint64_t chunk = (n + uint64_t(threads) - 1) / uint64_t(threads);
    for (int t = 0; t < threads; ++t) {
        const uint64_t b = uint64_t(t) * chunk, e = std::min(n, b + chunk);
        if (b < e) ranges.emplace_back(b, e);
    }
    std::vector<std::thread> pool;
    for (auto [b, e] : ranges) {
        uint64_t* acc = &partial[size_t(slot++) * 2];
        pool.emplace_back([=] {
            uint64_t nodeSum = 0, triSum = 0;
            std::vector<uint32_t> stack;
            stack.reserve(128);
            for (uint64_t i = b; i < e; ++i) {
                const Ray ray = loadRay(rays + i * 8);
                const V3 inv = invDir(ray.d);
                float tMax = ray.tmax;
                int64_t hitTri = -1;
                uint32_t nNode = 0, nTri = 0;
                stack.clear();
                stack.push_back(0);
                while (!stack.empty()) {
                    const OrNode& nd = nodes[stack.back()];
                    stack.pop_back();
                    ++nNode;
                    if (slabGlsl(nd, ray.o, inv, tMax) >= tMax) continue;
                    const uint32_t cnt = nd.packed >> 2;
                    if (cnt == 0) {
                        const uint32_t axis = nd.packed & 3u;
                        uint32_t nearC = nd.leftOrPrim, farC = nd.leftOrPrim + 1;
                        if (ray.d[int(axis)] < 0.0f) std::swap(nearC, farC);
                        stack.push_back(farC);
                        stack.push_back(nearC);
                    } else {
                        for (uint32_t k = 0; k < cnt; ++k) {
                            ++nTri;
                            const uint32_t orig = primOrder[nd.leftOrPrim + k];
                            const float t = isectMT(vert(verts9, orig, 0), vert(verts9, orig, 1),
                                                    vert(verts9, orig, 2), ray.o, ray.d,
                                                    ray.tmin, tMax, nullptr, nullptr);
                            if (t > 0.0f) { tMax = t; hitTri = int64_t(nd.leftOrPrim + k); }
                        }
                    }
                }
                if (hits) {
                    hits[i * 2 + 0] = hitTri >= 0 ? primOrder[hitTri] : 0xFFFFFFFFu;
                    hits[i * 2 + 1] = fbits(tMax);
                }
                if (counts) { counts[i * 2] = nNode; counts[i * 2 + 1] = nTri; }
                nodeSum += nNode;
                triSum += nTri;
            }
            acc[0] = nodeSum;
            acc[1] = triSum;
        });
    }
    for (auto& th : pool) th.join();
    uint64_t s0 = 0, s1 = 0;
    for (int t = 0; t < slot; ++t) { s0 += partial[size_t(t) * 2]; s1 += partial[size_t(t) * 2 + 1]; }
    if (sums) { sums[0] = s0; sums[1] = s1; }
    (void)bitsf;
}

void or_brute_any(const float* verts9, uint32_t triCount, const float* rays,
                  uint64_t n, uint8_t* occluded, int threads) {
    parallelFor(n, threads, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; ++i) {
            const Ray ray = loadRay(rays + i * 8);
            const WtRay wr = wtSetup(ray.o, ray.d);
            uint8_t occ = 0;
            for (uint32_t t = 0; t < triCount && !occ; ++t) {
                float th, b1, b2;
                if (wtIsect(wr, vert(verts9, t, 0), vert(verts9, t, 1), vert(verts9, t, 2),
                            ray.tmin, ray.tmax, &th, &b1, &b2) && th < ray.tmax)
                    occ = 1;
            }
            occluded[i] = occ;
        }
    });
}

} // extern "C"
