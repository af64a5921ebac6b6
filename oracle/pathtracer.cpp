// Oracle path tracer (FP64); see pathtracer.h.  TEST INFRASTRUCTURE ONLY.
#include "pathtracer.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoPi = 6.28318530717958647692;
constexpr double kInvPi = 0.31830988618379067154;
constexpr double kEps = 1e-4;  // common.glsl:9 (EPS_RAY)

struct D3 {
    double x, y, z;
    double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
inline D3 d3(double x, double y, double z) { return { x, y, z }; }
inline D3 operator+(D3 a, D3 b) { return { a.x + b.x, a.y + b.y, a.z + b.z }; }
inline D3 operator-(D3 a, D3 b) { return { a.x - b.x, a.y - b.y, a.z - b.z }; }
inline D3 operator-(D3 a) { return { -a.x, -a.y, -a.z }; }
inline D3 operator*(D3 a, D3 b) { return { a.x * b.x, a.y * b.y, a.z * b.z }; }
inline D3 operator*(D3 a, double s) { return { a.x * s, a.y * s, a.z * s }; }
inline D3 operator/(D3 a, double s) { return { a.x / s, a.y / s, a.z / s }; }
inline double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline D3 cross(D3 a, D3 b) { return { a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x }; }
inline double len(D3 a) { return std::sqrt(dot(a, a)); }
inline D3 norm(D3 a) { return a / len(a); }
inline double maxc(D3 a) { return std::max(a.x, std::max(a.y, a.z)); }
inline double lum(D3 c) { return 0.2126 * c.x + 0.7152 * c.y + 0.0722 * c.z; }
inline D3 f3(const float* p) { return { p[0], p[1], p[2] }; }
inline uint32_t fbits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float bitsf(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
inline double clampd(double x, double a, double b) { return std::min(std::max(x, a), b); }

// ---------------------------------------------------------------------- RNG --
struct Rng {
    uint32_t s[4];
    double r1() { return double(or_rng1(s)); }
    void r2(double& a, double& b) {
        float o[2];
        or_rng2(s, o);
        a = o[0];
        b = o[1];
    }
};

// ---------------------------------------------------------------- env / CDF --
struct Env {
    const float* px = nullptr;
    uint32_t w = 0, h = 0;
    double rot = 0, intensity = 1, integral = 1;
    std::vector<float> cdf;  // [conditional h*w][marginal h]
    const uint32_t* alias = nullptr;

    bool valid() const { return px != nullptr; }
    D3 texel(int x, int y) const {
        const float* p = px + (size_t(y) * w + x) * 4;
        return { p[0], p[1], p[2] };
    }
    // Bilinear, REPEAT on both axes.
    D3 lookup(double u, double v) const {
        const double fx = u * w - 0.5, fy = v * h - 0.5;
        const double x0f = std::floor(fx), y0f = std::floor(fy);
        auto wrap = [](long i, long n) { long m = i % n; return int(m < 0 ? m + n : m); };
        const int x0 = wrap(long(x0f), w), y0 = wrap(long(y0f), h);
        const int x1 = (x0 + 1) % int(w), y1 = (y0 + 1) % int(h);
        const double ax = fx - x0f, ay = fy - y0f;
        const D3 a = texel(x0, y0) * (1 - ax) + texel(x1, y0) * ax;
        const D3 b = texel(x0, y1) * (1 - ax) + texel(x1, y1) * ax;
        return a * (1 - ay) + b * ay;
    }
    void toUv(D3 d, double& u, double& v) const {
        const double phi = std::atan2(d.z, d.x) - rot;
        const double theta = std::acos(clampd(d.y, -1.0, 1.0));
        const double x = phi / kTwoPi + 1.0;
        u = x - std::floor(x);
        v = theta / kPi;
    }
    D3 fromUv(double u, double v) const {
        const double phi = u * kTwoPi + rot, theta = v * kPi;
        const double st = std::sin(theta);
        return { st * std::cos(phi), std::cos(theta), st * std::sin(phi) };
    }
    D3 radiance(D3 d) const {
        if (!valid()) return { 0, 0, 0 };
        double u, v;
        toUv(d, u, v);
        return lookup(u, v) * intensity;
    }
    double pdf(D3 d) const {
        if (!valid()) return 0.0;
        double u, v;
        toUv(d, u, v);
        return lum(lookup(u, v)) / integral;
    }
    uint32_t search(size_t base, uint32_t count, float u) const {
        uint32_t lo = 0, hi = count - 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (cdf[base + mid] < u) lo = mid + 1;
            else hi = mid;
        }
        return lo;
    }
};

// Renderer.cpp:620-655 in float/double exactly as the reference packs it.
double buildCdf(const float* px, uint32_t w, uint32_t h, std::vector<float>& cdf) {
    cdf.assign(size_t(w) * h + h, 0.0f);
    float* cond = cdf.data();
    float* marg = cdf.data() + size_t(w) * h;
    const float dTheta = float(kPi) / float(h);
    const float dPhi = 2.0f * float(kPi) / float(w);
    double integral = 0.0;
    std::vector<double> rows(h, 0.0);
    for (uint32_t y = 0; y < h; ++y) {
        const float sinT = std::sin((float(y) + 0.5f) * dTheta);
        double acc = 0.0;
        for (uint32_t x = 0; x < w; ++x) {
            const float* p = px + (size_t(y) * w + x) * 4;
            const float l = (0.2126f * p[0] + 0.7152f * p[1]) + 0.0722f * p[2];
            acc += double(l) * sinT;
            cond[size_t(y) * w + x] = float(acc);
        }
        rows[y] = acc;
        integral += acc * dTheta * dPhi;
        const float inv = acc > 0.0 ? float(1.0 / acc) : 0.0f;
        for (uint32_t x = 0; x < w; ++x) cond[size_t(y) * w + x] *= inv;
    }
    double total = 0.0, run = 0.0;
    for (uint32_t y = 0; y < h; ++y) total += rows[y];
    for (uint32_t y = 0; y < h; ++y) {
        run += rows[y];
        marg[y] = total > 0.0 ? float(run / total) : 0.0f;
    }
    return integral;
}

// ----------------------------------------------------------------- textures --
// Level-0 bilinear lookup with REPEAT addressing (the reference's sampler,
// GpuContext.cpp:190-198; a compute shader's implicit LOD is the base level),
// texels decoded to linear first (sRGB formats decode before filtering).
using D4 = std::array<double, 4>;
double srgbDecode(double c) { return c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4); }
D4 texelOf(const OrTexture& t, uint32_t x, uint32_t y) {
    const size_t i = size_t(y) * t.width + x;
    if (t.format == 2) {
        const float* p = static_cast<const float*>(t.pixels) + i * 4;
        return { p[0], p[1], p[2], p[3] };
    }
    const uint8_t* p = static_cast<const uint8_t*>(t.pixels) + i * 4;
    D4 c{ p[0] / 255.0, p[1] / 255.0, p[2] / 255.0, p[3] / 255.0 };
    if (t.format == 0)
        for (int k = 0; k < 3; ++k) c[k] = srgbDecode(c[k]);
    return c;
}
D4 texSample(const OrTexture* t, double u, double v) {
    if (!t || t->width == 0 || t->height == 0 || !t->pixels) return { 1, 1, 1, 1 };  // dummy: white
    // non-finite or huge coordinates read texel (0, 0)'s neighbourhood
    if (!(u == u) || !(v == v) || std::fabs(u) > 1e7 || std::fabs(v) > 1e7) u = v = 0;
    const double fx = u * t->width - 0.5, fy = v * t->height - 0.5;
    const double x0f = std::floor(fx), y0f = std::floor(fy);
    auto wrap = [](double i, uint32_t n) {
        long m = long(i) % long(n);
        return uint32_t(m < 0 ? m + long(n) : m);
    };
    const uint32_t x0 = wrap(x0f, t->width), y0 = wrap(y0f, t->height);
    const uint32_t x1 = (x0 + 1) % t->width, y1 = (y0 + 1) % t->height;
    const double ax = fx - x0f, ay = fy - y0f;
    const D4 c00 = texelOf(*t, x0, y0), c10 = texelOf(*t, x1, y0), c01 = texelOf(*t, x0, y1), c11 = texelOf(*t, x1, y1);
    D4 r;
    for (int k = 0; k < 4; ++k) {
        const double a = c00[k] * (1 - ax) + c10[k] * ax, b = c01[k] * (1 - ax) + c11[k] * ax;
        r[k] = a * (1 - ay) + b * ay;
    }
    return r;
}

// -------------------------------------------------------------------- scene --
struct Inst {
    double o2w[16], w2o[16];
    uint32_t blasRoot, material;
};

struct Hit {
    double t, b1, b2;
    uint32_t tri, inst;
};

struct Scene {
    const OrPtScene* s;
    std::vector<Inst> inst;
    Env env;
};

// Conservative slab test in double against float boxes.
bool slabD(const OrNode& n, D3 o, D3 inv, double limit) {
    double enter = 0.0, far = limit * (1 + 1e-9);
    for (int a = 0; a < 3; ++a) {
        const double lo = n.lo[a], hi = n.hi[a];
        const double oi = o[a], ii = inv[a];
        if (std::isinf(ii)) {
            if (oi < lo || oi > hi) return false;
            continue;
        }
        double t0 = (lo - oi) * ii, t1 = (hi - oi) * ii;
        if (t0 > t1) std::swap(t0, t1);
        enter = std::max(enter, t0);
        far = std::min(far, t1 * (1 + 1e-9));
    }
    return enter <= far;
}

// Watertight test in double.
bool isectD(D3 o, D3 d, int kx, int ky, int kz, double sx, double sy, double sz, D3 v0, D3 v1, D3 v2,
            double tmax, double& t, double& b1, double& b2) {
    const D3 A = v0 - o, B = v1 - o, C = v2 - o;
    const double Ax = A[kx] - sx * A[kz], Ay = A[ky] - sy * A[kz];
    const double Bx = B[kx] - sx * B[kz], By = B[ky] - sy * B[kz];
    const double Cx = C[kx] - sx * C[kz], Cy = C[ky] - sy * C[kz];
    const double U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
    if ((U < 0 || V < 0 || W < 0) && (U > 0 || V > 0 || W > 0)) return false;
    const double det = U + V + W;
    if (det == 0.0) return false;
    const double T = U * (sz * A[kz]) + V * (sz * B[kz]) + W * (sz * C[kz]);
    const double tt = T / det;
    if (!(tt > kEps && tt < tmax)) return false;
    t = tt;
    b1 = V / det;
    b2 = W / det;
    return true;
}

D3 xfPoint(const double* m, D3 p) {
    return { m[0] * p.x + m[4] * p.y + m[8] * p.z + m[12], m[1] * p.x + m[5] * p.y + m[9] * p.z + m[13],
             m[2] * p.x + m[6] * p.y + m[10] * p.z + m[14] };
}
D3 xfDir(const double* m, D3 p) {
    return { m[0] * p.x + m[4] * p.y + m[8] * p.z, m[1] * p.x + m[5] * p.y + m[9] * p.z,
             m[2] * p.x + m[6] * p.y + m[10] * p.z };
}
// transpose(mat3(m)) * n
D3 xfNormal(const double* m, D3 n) {
    return { m[0] * n.x + m[1] * n.y + m[2] * n.z, m[4] * n.x + m[5] * n.y + m[6] * n.z,
             m[8] * n.x + m[9] * n.y + m[10] * n.z };
}

// Two-level traversal; any = true stops at the first hit with t < tmax.
bool trace(const Scene& sc, D3 o, D3 d, double tmax, bool any, Hit& hit) {
    const OrPtScene& s = *sc.s;
    if (s.instanceCount == 0) return false;
    const D3 inv = { 1.0 / d.x, 1.0 / d.y, 1.0 / d.z };
    double best = tmax;
    bool found = false;
    thread_local std::vector<uint32_t> tstack, bstack;
    tstack.clear();
    tstack.push_back(s.tlasRoot);
    while (!tstack.empty()) {
        const OrNode& n = s.nodes[tstack.back()];
        tstack.pop_back();
        if (!slabD(n, o, inv, best)) continue;
        const uint32_t cnt = n.packed >> 2;
        if (cnt == 0) {
            tstack.push_back(n.leftOrPrim);
            tstack.push_back(n.leftOrPrim + 1);
            continue;
        }
        for (uint32_t j = 0; j < cnt; ++j) {
            const uint32_t ii = n.leftOrPrim + j;
            const Inst& in = sc.inst[ii];
            const D3 oo = xfPoint(in.w2o, o), dd = xfDir(in.w2o, d);
            const D3 oinv = { 1.0 / dd.x, 1.0 / dd.y, 1.0 / dd.z };
            const double ax = std::fabs(dd.x), ay = std::fabs(dd.y), az = std::fabs(dd.z);
            int kz = 0;
            if (ay > ax) kz = 1;
            if (az > (kz == 0 ? ax : ay)) kz = 2;
            int kx = (kz + 1) % 3, ky = (kx + 1) % 3;
            if (dd[kz] < 0) std::swap(kx, ky);
            const double sx = dd[kx] / dd[kz], sy = dd[ky] / dd[kz], sz = 1.0 / dd[kz];
            bstack.clear();
            bstack.push_back(in.blasRoot);
            while (!bstack.empty()) {
                const OrNode& b = s.nodes[bstack.back()];
                bstack.pop_back();
                if (!slabD(b, oo, oinv, best)) continue;
                const uint32_t bc = b.packed >> 2;
                if (bc == 0) {
                    bstack.push_back(b.leftOrPrim);
                    bstack.push_back(b.leftOrPrim + 1);
                    continue;
                }
                for (uint32_t k = b.leftOrPrim; k < b.leftOrPrim + bc; ++k) {
                    const uint32_t* tr = s.triangles + size_t(k) * 4;
                    double t, b1, b2;
                    if (!isectD(oo, dd, kx, ky, kz, sx, sy, sz, f3(s.positions + size_t(tr[0]) * 4),
                                f3(s.positions + size_t(tr[1]) * 4), f3(s.positions + size_t(tr[2]) * 4), best, t,
                                b1, b2))
                        continue;
                    if (any) return true;
                    best = t;
                    found = true;
                    hit = { t, b1, b2, k, ii };
                }
            }
        }
    }
    return found;
}

// ---------------------------------------------- reference-order fetch counts --
// core/shaders/intersect.glsl:10-86 (closest) and :128-178 (shadow), in float
// with Moeller-Trumbore, counting node pops and triangle tests.
struct F3 { float x, y, z; float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); } };
inline F3 ff(const float* p) { return { p[0], p[1], p[2] }; }
inline F3 fsub(F3 a, F3 b) { return { a.x - b.x, a.y - b.y, a.z - b.z }; }
inline F3 fcross(F3 a, F3 b) { return { a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y }; }
inline float fdot(F3 a, F3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }

float slabGlsl(const OrNode& n, F3 o, F3 inv, float tMax) {
    float tmn[3], tmx[3];
    for (int a = 0; a < 3; ++a) {
        const float t0 = (n.lo[a] - o[a]) * inv[a], t1 = (n.hi[a] - o[a]) * inv[a];
        tmn[a] = std::fmin(t0, t1);
        tmx[a] = std::fmax(t0, t1);
    }
    const float enter = std::fmax(std::fmax(tmn[0], tmn[1]), std::fmax(tmn[2], 0.0f));
    const float exit = std::fmin(std::fmin(tmx[0], tmx[1]), std::fmin(tmx[2], tMax));
    return enter <= exit ? enter : tMax;
}
float mtGlsl(const OrPtScene& s, uint32_t tri, F3 o, F3 d, float tMax) {
    const uint32_t* t = s.triangles + size_t(tri) * 4;
    const F3 v0 = ff(s.positions + size_t(t[0]) * 4);
    const F3 e1 = fsub(ff(s.positions + size_t(t[1]) * 4), v0), e2 = fsub(ff(s.positions + size_t(t[2]) * 4), v0);
    const F3 p = fcross(d, e2);
    const float det = fdot(e1, p);
    if (std::fabs(det) < 1e-12f) return -1.0f;
    const float inv = 1.0f / det;
    const F3 sv = fsub(o, v0);
    const float u = fdot(sv, p) * inv;
    if (u < 0.0f || u > 1.0f) return -1.0f;
    const F3 q = fcross(sv, e1);
    const float v = fdot(d, q) * inv;
    if (v < 0.0f || u + v > 1.0f) return -1.0f;
    const float t2 = fdot(e2, q) * inv;
    if (t2 <= 1e-4f || t2 >= tMax) return -1.0f;
    return t2;
}
F3 objRay(const float* instF, F3 v, float w) {
    // worldToObject (columns at offset 16) * (v, w), GLSL mat4 * vec4
    const float* m = instF + 16;
    return { m[0] * v.x + m[4] * v.y + m[8] * v.z + m[12] * w, m[1] * v.x + m[5] * v.y + m[9] * v.z + m[13] * w,
             m[2] * v.x + m[6] * v.y + m[10] * v.z + m[14] * w };
}

void countClosest(const OrPtScene& s, F3 o, F3 d, uint64_t& nodes, uint64_t& tris) {
    if (s.instanceCount == 0) return;
    float tMax = 1e30f;
    const F3 inv = { 1.0f / d.x, 1.0f / d.y, 1.0f / d.z };
    std::vector<uint32_t> st{ s.tlasRoot };
    std::vector<uint32_t> bs;
    while (!st.empty()) {
        const OrNode& n = s.nodes[st.back()];
        st.pop_back();
        ++nodes;
        if (slabGlsl(n, o, inv, tMax) >= tMax) continue;
        const uint32_t cnt = n.packed >> 2;
        if (cnt == 0) {
            uint32_t nearC = n.leftOrPrim, farC = n.leftOrPrim + 1;
            if (d[int(n.packed & 3u)] < 0.0f) std::swap(nearC, farC);
            st.push_back(farC);
            st.push_back(nearC);
            continue;
        }
        const float* inst = s.instances + size_t(n.leftOrPrim) * 36;
        uint32_t root;
        std::memcpy(&root, inst + 32, 4);
        const F3 oo = objRay(inst, o, 1.0f), dd = objRay(inst, d, 0.0f);
        const F3 oinv = { 1.0f / dd.x, 1.0f / dd.y, 1.0f / dd.z };
        bs.assign(1, root);
        while (!bs.empty()) {
            const OrNode& b = s.nodes[bs.back()];
            bs.pop_back();
            ++nodes;
            if (slabGlsl(b, oo, oinv, tMax) >= tMax) continue;
            const uint32_t bc = b.packed >> 2;
            if (bc == 0) {
                uint32_t nearC = b.leftOrPrim, farC = b.leftOrPrim + 1;
                if (dd[int(b.packed & 3u)] < 0.0f) std::swap(nearC, farC);
                bs.push_back(farC);
                bs.push_back(nearC);
                continue;
            }
            for (uint32_t k = 0; k < bc; ++k) {
                ++tris;
                const float t = mtGlsl(s, b.leftOrPrim + k, oo, dd, tMax);
                if (t > 0.0f) tMax = t;
            }
        }
    }
}

void countShadow(const OrPtScene& s, F3 o, F3 d, float tMax, uint64_t& nodes, uint64_t& tris) {
    if (s.instanceCount == 0) return;
    const F3 inv = { 1.0f / d.x, 1.0f / d.y, 1.0f / d.z };
    std::vector<uint32_t> st{ s.tlasRoot };
    std::vector<uint32_t> bs;
    while (!st.empty()) {
        const OrNode& n = s.nodes[st.back()];
        st.pop_back();
        ++nodes;
        if (slabGlsl(n, o, inv, tMax) >= tMax) continue;
        const uint32_t cnt = n.packed >> 2;
        if (cnt == 0) {
            st.push_back(n.leftOrPrim);
            st.push_back(n.leftOrPrim + 1);
            continue;
        }
        const float* inst = s.instances + size_t(n.leftOrPrim) * 36;
        uint32_t root;
        std::memcpy(&root, inst + 32, 4);
        const F3 oo = objRay(inst, o, 1.0f), dd = objRay(inst, d, 0.0f);
        const F3 oinv = { 1.0f / dd.x, 1.0f / dd.y, 1.0f / dd.z };
        bs.assign(1, root);
        while (!bs.empty()) {
            const OrNode& b = s.nodes[bs.back()];
            bs.pop_back();
            ++nodes;
            if (slabGlsl(b, oo, oinv, tMax) >= tMax) continue;
            const uint32_t bc = b.packed >> 2;
            if (bc == 0) {
                bs.push_back(b.leftOrPrim);
                bs.push_back(b.leftOrPrim + 1);
                continue;
            }
            for (uint32_t k = 0; k < bc; ++k) {
                ++tris;
                if (mtGlsl(s, b.leftOrPrim + k, oo, dd, tMax) > 0.0f) return;
            }
        }
    }
}

// --------------------------------------------------------------------- BSDF --
struct Mat {
    D3 base, emission;
    double opacity, rough, metal, trans, ior;
};
struct Bsdf {
    D3 base;
    double rough, metal, trans, ior;
    bool entering;
};
double ggxD(D3 h, double a) {
    const double a2 = a * a, c2 = h.z * h.z, dd = c2 * (a2 - 1) + 1;
    return a2 / std::max(kPi * dd * dd, 1e-12);
}
double ggxL(D3 w, double a) {
    const double c2 = w.z * w.z, t2 = std::max(1 - c2, 0.0) / std::max(c2, 1e-12);
    return 0.5 * (std::sqrt(1 + a * a * t2) - 1);
}
double G1(D3 w, double a) { return 1 / (1 + ggxL(w, a)); }
double G2(D3 wo, D3 wi, double a) { return 1 / (1 + ggxL(wo, a) + ggxL(wi, a)); }
D3 vndf(D3 wo, double a, double u0, double u1) {
    const D3 vh = norm(d3(a * wo.x, a * wo.y, wo.z));
    const double ls = vh.x * vh.x + vh.y * vh.y;
    const D3 T1 = ls > 0 ? d3(-vh.y, vh.x, 0) / std::sqrt(ls) : d3(1, 0, 0);
    const D3 T2 = cross(vh, T1);
    const double r = std::sqrt(u0), phi = kTwoPi * u1;
    const double t1 = r * std::cos(phi);
    double t2 = r * std::sin(phi);
    const double sm = 0.5 * (1 + vh.z);
    t2 = (1 - sm) * std::sqrt(std::max(1 - t1 * t1, 0.0)) + sm * t2;
    const D3 nh = T1 * t1 + T2 * t2 + vh * std::sqrt(std::max(1 - t1 * t1 - t2 * t2, 0.0));
    return norm(d3(a * nh.x, a * nh.y, std::max(nh.z, 1e-6)));
}
D3 schlick(D3 f0, double c) {
    const double m = clampd(1 - c, 0, 1), m2 = m * m, w = m2 * m2 * m;
    return f0 + (d3(1, 1, 1) - f0) * w;
}
double fresnelD(double ci, double eta) {
    ci = clampd(ci, 0, 1);
    const double s2 = (1 - ci * ci) / (eta * eta);
    if (s2 >= 1) return 1;
    const double ct = std::sqrt(1 - s2);
    const double rs = (ci - eta * ct) / (ci + eta * ct), rp = (eta * ci - ct) / (eta * ci + ct);
    return 0.5 * (rs * rs + rp * rp);
}
D3 f0Of(const Bsdf& m) { return d3(0.04, 0.04, 0.04) * (1 - m.metal) + m.base * m.metal; }
void lobes(const Bsdf& m, D3 wo, double& wd, double& ws, double& wt) {
    ws = lum(schlick(f0Of(m), std::fabs(wo.z)));
    wd = (1 - m.metal) * (1 - m.trans) * lum(m.base);
    wt = (1 - m.metal) * m.trans;
    double tot = wd + ws + wt;
    if (tot < 1e-7) { wd = 1; ws = 0; wt = 0; tot = 1; }
    wd /= tot; ws /= tot; wt /= tot;
}
void bsdfEval(const Bsdf& m, D3 wo, D3 wi, D3& f, double& pdf) {
    f = d3(0, 0, 0);
    pdf = 0;
    const double a = std::max(m.rough * m.rough, 1e-3);
    double wd, ws, wt;
    lobes(m, wo, wd, ws, wt);
    const double eta = m.entering ? m.ior : 1 / m.ior;
    if (wi.z > 0) {
        if (wd > 0) {
            f = f + m.base * ((1 - m.metal) * (1 - m.trans) * kInvPi);
            pdf += wd * wi.z * kInvPi;
        }
        const D3 h = norm(wo + wi);
        const double dd = ggxD(h, a), g2 = G2(wo, wi, a);
        f = f + schlick(f0Of(m), dot(wo, h)) * (dd * g2 / std::max(4 * wo.z * wi.z, 1e-7));
        pdf += ws * G1(wo, a) * dd / std::max(4 * wo.z, 1e-7);
    } else if (wt > 0) {
        D3 h = norm(wo + wi * eta);
        if (h.z < 0) h = -h;
        const double woH = dot(wo, h), wiH = dot(wi, h);
        if (woH * wiH < 0) {
            const double F = fresnelD(std::fabs(woH), eta), dd = ggxD(h, a);
            const double g2 = G2(wo, d3(wi.x, wi.y, -wi.z), a);
            double den = woH + eta * wiH;
            den = std::max(den * den, 1e-7);
            const double jac = eta * eta * std::fabs(wiH) / den;
            f = f + m.base * ((1 - F) * dd * g2 * std::fabs(woH) * jac / std::max(std::fabs(wo.z * wi.z), 1e-7) *
                              (1 - m.metal) * m.trans);
            pdf += wt * (1 - F) * G1(wo, a) * dd * std::fabs(woH) / std::max(wo.z, 1e-7) * jac;
        }
    }
}
D3 reflectD(D3 i, D3 n) { return i - n * (2 * dot(n, i)); }
D3 refractD(D3 i, D3 n, double eta) {
    const double ni = dot(n, i), k = 1 - eta * eta * (1 - ni * ni);
    if (k < 0) return d3(0, 0, 0);
    return i * eta - n * (eta * ni + std::sqrt(k));
}
// bsdf.glsl:147-192 with the draws given: pick, (u0, u1), and fres() for the
// transmission lobe's Fresnel choice (drawn only on that lobe, after u0/u1).
template <class Fres>
bool bsdfSampleU(const Bsdf& m, D3 wo, double pick, double u0, double u1, Fres fres, D3& wi, D3& f, double& pdf,
                 int& branch) {
    const double a = std::max(m.rough * m.rough, 1e-3);
    double wd, ws, wt;
    lobes(m, wo, wd, ws, wt);
    const double eta = m.entering ? m.ior : 1 / m.ior;
    if (pick < wd) {
        branch = 0;
        const double r = std::sqrt(u0), phi = kTwoPi * u1;
        wi = d3(r * std::cos(phi), r * std::sin(phi), std::sqrt(std::max(1 - u0, 0.0)));
    } else if (pick < wd + ws) {
        branch = 1;
        const D3 h = vndf(wo, a, u0, u1);
        wi = reflectD(-wo, h);
        if (wi.z <= 0) return false;
    } else {
        const D3 h = vndf(wo, a, u0, u1);
        const double F = fresnelD(std::fabs(dot(wo, h)), eta);
        if (fres() < F) {
            branch = 2;
            wi = reflectD(-wo, h);
            if (wi.z <= 0) return false;
        } else {
            branch = 3;
            wi = refractD(-wo, h, 1 / eta);
            if (wi.z >= 0 || len(wi) < 1e-6) return false;
            wi = norm(wi);
        }
    }
    bsdfEval(m, wo, wi, f, pdf);
    return pdf >= 1e-9;
}
bool bsdfSample(const Bsdf& m, D3 wo, Rng& rng, D3& wi, D3& f, double& pdf) {
    const double pick = rng.r1();
    double u0, u1;
    rng.r2(u0, u1);
    int branch;
    return bsdfSampleU(m, wo, pick, u0, u1, [&] { return rng.r1(); }, wi, f, pdf, branch);
}

// ------------------------------------------------------------------- lights --
struct LS {
    D3 wi, rad;
    double pdf = 0, dist = 0;
    bool delta = false, valid = false;
};
void onbD(D3 n, D3& t, D3& b) {
    const double s = n.z >= 0 ? 1 : -1, a = -1 / (s + n.z), c = n.x * n.y * a;
    t = d3(1 + s * n.x * n.x * a, s * c, -s * n.x);
    b = d3(c, s + n.y * n.y * a, -n.y);
}
bool sampleEnvD(const Env& env, Rng& rng, D3& dir, D3& rad, double& pdf) {
    const double u1 = rng.r1(), u2 = rng.r1();
    if (env.w == 0 || env.h == 0) return false;
    uint32_t row, col;
    if (env.alias) {
        const uint32_t n = env.w * env.h;
        uint32_t i = uint32_t(float(u1) * float(n));
        if (i >= n) i = n - 1;
        const float prob = bitsf(env.alias[size_t(i) * 2]);
        const uint32_t texel = float(u2) < prob ? i : env.alias[size_t(i) * 2 + 1];
        row = texel / env.w;
        col = texel % env.w;
    } else {
        row = env.search(size_t(env.w) * env.h, env.h, float(u1));
        col = env.search(size_t(row) * env.w, env.w, float(u2));
    }
    const double u = (col + 0.5) / env.w, v = (row + 0.5) / env.h;
    dir = env.fromUv(u, v);
    rad = env.lookup(u, v) * env.intensity;
    pdf = env.pdf(dir);
    return pdf > 1e-9;
}
// Analytic light record (sun, point, rect) at p; u0, u1 = the light's rng2
// (unused for a point light).
LS sampleAnalytic(const float* l, double pick, D3 p, double u0, double u1) {
    LS ls;
    int32_t type;
    std::memcpy(&type, l + 16, 4);
    if (type == 1) {
        const double cosMax = l[3];
        const double cosT = 1 + (cosMax - 1) * u0, sinT = std::sqrt(std::max(1 - cosT * cosT, 0.0));
        const double phi = kTwoPi * u1;
        D3 t, b;
        const D3 ax = f3(l);
        onbD(ax, t, b);
        ls.wi = norm(t * (sinT * std::cos(phi)) + b * (sinT * std::sin(phi)) + ax * cosT);
        ls.rad = f3(l + 4);
        ls.pdf = pick / (kTwoPi * std::max(1 - cosMax, 1e-7));
        ls.dist = 1e30; ls.delta = true; ls.valid = true;
    } else if (type == 2) {
        const D3 dv = f3(l) - p;
        const double d2 = std::max(dot(dv, dv), 1e-8);
        ls.dist = std::sqrt(d2);
        ls.wi = dv / ls.dist;
        ls.rad = f3(l + 4) / d2;
        ls.pdf = pick; ls.delta = true; ls.valid = true;
    } else if (type == 3) {
        const D3 q = f3(l) + f3(l + 4) * u0 + f3(l + 8) * u1;
        const D3 dv = q - p;
        const double d2 = std::max(dot(dv, dv), 1e-8);
        ls.dist = std::sqrt(d2);
        ls.wi = dv / ls.dist;
        const D3 n = norm(cross(f3(l + 4), f3(l + 8)));
        double cosL = dot(n, -ls.wi);
        if (l[15] > 0.5f) cosL = std::fabs(cosL);
        if (cosL < 1e-6) return ls;
        ls.rad = f3(l + 12);
        ls.pdf = d2 / (cosL * std::max(double(l[3]), 1e-8)) * pick;
        ls.delta = true; ls.valid = true;
    }
    return ls;
}

LS sampleLight(const Scene& sc, D3 p, Rng& rng) {
    LS ls;
    const OrPtScene& s = *sc.s;
    if (s.lightCount == 0) return ls;
    uint32_t idx = uint32_t(float(rng.r1()) * float(s.lightCount));
    idx = std::min(idx, s.lightCount - 1);
    const double pick = 1.0 / s.lightCount;
    const float* l = s.lights + size_t(idx) * 20;
    int32_t type;
    std::memcpy(&type, l + 16, 4);
    if (type == 0) {
        D3 dir, rad;
        double pdf;
        if (!sampleEnvD(sc.env, rng, dir, rad, pdf)) return ls;
        ls.wi = dir; ls.rad = rad; ls.pdf = pdf * pick; ls.dist = 1e30; ls.valid = true;
        return ls;
    }
    double u0 = 0, u1 = 0;
    if (type == 1 || type == 3) rng.r2(u0, u1);
    return sampleAnalytic(l, pick, p, u0, u1);
}

double balance(double a, double b) { return a / std::max(a + b, 1e-12); }

// One path sample (pathtrace.comp:100-225); returns the clamped radiance.
D3 samplePath(const Scene& sc, const OrPtCamera& cam, const OrPtParams& P, uint32_t px, uint32_t py,
              uint32_t frame, uint64_t* cnt) {
    const OrPtScene& s = *sc.s;
    Rng rng{ { px, py, frame ^ P.seedMix, 0u } };
    double jx, jy;
    rng.r2(jx, jy);
    const double ux = (px + jx) / P.width, uy = (py + jy) / P.height;
    const double sx = cam.frustum[0] * (1 - ux) + cam.frustum[1] * ux;
    const double sy = cam.frustum[3] * (1 - uy) + cam.frustum[2] * uy;
    D3 o, d;
    if (cam.ortho) {
        o = f3(cam.pos) + f3(cam.right) * sx + f3(cam.up) * sy;
        d = f3(cam.forward);
    } else {
        o = f3(cam.pos);
        d = norm(f3(cam.forward) + f3(cam.right) * sx + f3(cam.up) * sy);
    }
    D3 L = d3(0, 0, 0), beta = d3(1, 1, 1);
    double prevPdf = 0;
    const uint32_t maxB = std::min(P.maxBounces, 16u);
    const double envPick = (sc.env.valid() && s.lightCount > 0) ? 1.0 / s.lightCount : 0.0;
    for (uint32_t bounce = 0; bounce <= maxB; ++bounce) {
        if (cnt) {
            ++cnt[0];
            countClosest(s, F3{ float(o.x), float(o.y), float(o.z) }, F3{ float(d.x), float(d.y), float(d.z) },
                         cnt[1], cnt[2]);
        }
        Hit h;
        if (!trace(sc, o, d, 1e30, false, h)) {
            const D3 env = sc.env.radiance(d);
            if (maxc(env) > 0) {
                double w = 1;
                if (bounce > 0) w = balance(prevPdf, sc.env.pdf(d) * envPick);
                L = L + beta * env * w;
            }
            break;
        }
        // surface (pathtrace.comp:39-80)
        const Inst& in = sc.inst[h.inst];
        const uint32_t* tr = s.triangles + size_t(h.tri) * 4;
        const double b0 = 1 - h.b1 - h.b2;
        const D3 p0 = f3(s.positions + size_t(tr[0]) * 4), p1 = f3(s.positions + size_t(tr[1]) * 4),
                 p2 = f3(s.positions + size_t(tr[2]) * 4);
        const D3 pos = xfPoint(in.o2w, p0 * b0 + p1 * h.b1 + p2 * h.b2);
        D3 ng = norm(xfNormal(in.w2o, cross(p1 - p0, p2 - p0)));
        const D3 no = f3(s.normals + size_t(tr[0]) * 4) * b0 + f3(s.normals + size_t(tr[1]) * 4) * h.b1 +
                      f3(s.normals + size_t(tr[2]) * 4) * h.b2;
        D3 ns = norm(xfNormal(in.w2o, no));
        const bool front = dot(d, ng) < 0;
        if (!front) { ng = -ng; ns = -ns; }
        const float* mf = s.materials + size_t(in.material) * 20;
        Mat m{ f3(mf), f3(mf + 4), mf[3], mf[7], mf[8], mf[9], mf[10] };
        int32_t slot[5];  // base colour, roughness, metallic, normal, emission
        std::memcpy(slot, mf + 12, sizeof(slot));
        auto tex = [&](int k) -> const OrTexture* {
            if (slot[k] < 0) return nullptr;
            static const OrTexture kWhite{ 0, 0, 0, nullptr };
            return uint32_t(slot[k]) < s.textureCount ? &s.textures[slot[k]] : &kWhite;
        };
        double tu = 0, tv = 0;
        if (s.uvs) {
            const float* q0 = s.uvs + size_t(tr[0]) * 2, *q1 = s.uvs + size_t(tr[1]) * 2, *q2 = s.uvs + size_t(tr[2]) * 2;
            tu = q0[0] * b0 + q1[0] * h.b1 + q2[0] * h.b2;
            tv = q0[1] * b0 + q1[1] * h.b1 + q2[1] * h.b2;
        }
        // normal mapping, pathtrace.comp:65-71 (tangent space, +z up)
        if (slot[3] >= 0) {
            double th[4] = { 1, 0, 0, 1 };
            if (s.tangents) {
                const float* t0 = s.tangents + size_t(tr[0]) * 4, *t1 = s.tangents + size_t(tr[1]) * 4,
                            *t2 = s.tangents + size_t(tr[2]) * 4;
                for (int k = 0; k < 4; ++k) th[k] = t0[k] * b0 + t1[k] * h.b1 + t2[k] * h.b2;
            } else {
                th[0] = b0 + h.b1 + h.b2;
                th[3] = th[0];
            }
            const D3 tangent = norm(xfDir(in.o2w, d3(th[0], th[1], th[2])));
            const D4 c = texSample(tex(3), tu, tv);
            const D3 tn = d3(c[0] * 2 - 1, c[1] * 2 - 1, c[2] * 2 - 1);
            const D3 bit = cross(ns, tangent) * th[3];
            ns = norm(tangent * tn.x + bit * tn.y + ns * tn.z);
            if (dot(ns, ng) < 0) ns = ng;
        }
        // textured parameters, pathtrace.comp:82-94 and :158-159
        if (slot[0] >= 0) {
            const D4 c = texSample(tex(0), tu, tv);
            m.base = m.base * d3(c[0], c[1], c[2]);
        }
        if (slot[1] >= 0) m.rough *= texSample(tex(1), tu, tv)[0];
        if (slot[2] >= 0) m.metal *= texSample(tex(2), tu, tv)[0];
        if (m.opacity < 1 && rng.r1() >= m.opacity) {
            o = pos + d * kEps;
            continue;
        }
        D3 emission = m.emission;
        if (slot[4] >= 0) {
            const D4 c = texSample(tex(4), tu, tv);
            emission = emission * d3(c[0], c[1], c[2]);
        }
        if (maxc(emission) > 0 && front) L = L + beta * emission;
        const Bsdf bs{ m.base, m.rough, m.metal, m.trans, m.ior, front };
        D3 T, B;
        onbD(ns, T, B);
        D3 wo = d3(dot(-d, T), dot(-d, B), dot(-d, ns));
        if (wo.z < 1e-6) wo.z = 1e-6;
        wo = norm(wo);
        const LS ls = sampleLight(sc, pos, rng);
        if (ls.valid && ls.pdf > 1e-9) {
            const D3 wi = d3(dot(ls.wi, T), dot(ls.wi, B), dot(ls.wi, ns));
            D3 f;
            double bp;
            bsdfEval(bs, wo, wi, f, bp);
            if (maxc(f) > 0) {
                const D3 so = pos + ng * (wi.z > 0 ? kEps : -kEps);
                const double smax = ls.dist - 2 * kEps;
                if (cnt) {
                    ++cnt[3];
                    countShadow(s, F3{ float(so.x), float(so.y), float(so.z) },
                                F3{ float(ls.wi.x), float(ls.wi.y), float(ls.wi.z) }, float(smax), cnt[4], cnt[5]);
                }
                Hit sh;
                if (!trace(sc, so, ls.wi, smax, true, sh)) {
                    const double w = ls.delta ? 1 : balance(ls.pdf, bp);
                    L = L + beta * f * ls.rad * (std::fabs(wi.z) * w / ls.pdf);
                }
            }
        }
        if (bounce == maxB) break;  // the last segment's BSDF sample is unused
        D3 wi, f;
        double pdf;
        if (!bsdfSample(bs, wo, rng, wi, f, pdf)) break;
        const D3 wiW = norm(T * wi.x + B * wi.y + ns * wi.z);
        beta = beta * f * (std::fabs(wi.z) / pdf);
        prevPdf = pdf;
        if (bounce >= 3) {
            const double q = clampd(maxc(beta), 0.05, 0.95);
            if (rng.r1() >= q) break;
            beta = beta / q;
        }
        o = pos + ng * (wi.z > 0 ? kEps : -kEps);
        d = wiW;
    }
    if (P.fireflyClamp > 0) {
        const double mm = maxc(L);
        if (mm > P.fireflyClamp) L = L * (P.fireflyClamp / mm);
    }
    if (!std::isfinite(L.x) || !std::isfinite(L.y) || !std::isfinite(L.z)) L = d3(0, 0, 0);
    return L;
}

} // namespace

extern "C" {

void or_camera_lookat(const float pos[3], const float target[3], const float up[3], float fovYDeg,
                      uint32_t width, uint32_t height, OrPtCamera* out) {
    float f[3] = { target[0] - pos[0], target[1] - pos[1], target[2] - pos[2] };
    float inv = 1.0f / std::sqrt((f[0] * f[0] + f[1] * f[1]) + f[2] * f[2]);
    for (float& x : f) x *= inv;
    const float aspect = float(width) / float(height);
    const float tanHalf = std::tan(fovYDeg * 0.0174532925199432957692f * 0.5f);
    float r[3] = { f[1] * up[2] - up[1] * f[2], f[2] * up[0] - up[2] * f[0], f[0] * up[1] - up[0] * f[1] };
    const float rl = std::sqrt((r[0] * r[0] + r[1] * r[1]) + r[2] * r[2]);
    if (rl < 1e-6f) { r[0] = 1; r[1] = 0; r[2] = 0; }
    else { const float ri = 1.0f / rl; for (float& x : r) x *= ri; }
    const float u[3] = { r[1] * f[2] - f[1] * r[2], r[2] * f[0] - f[2] * r[0], r[0] * f[1] - f[0] * r[1] };
    for (int i = 0; i < 3; ++i) {
        out->pos[i] = pos[i];
        out->forward[i] = f[i];
        out->right[i] = r[i];
        out->up[i] = u[i];
    }
    out->frustum[0] = -tanHalf * aspect;
    out->frustum[1] = tanHalf * aspect;
    out->frustum[2] = -tanHalf;
    out->frustum[3] = tanHalf;
    out->ortho = 0;
}

void or_bsdf_eval(const OrBsdfParams* p, const double wo[3], uint64_t n, const double* wi, double* f, double* pdf) {
    const Bsdf m{ d3(p->base[0], p->base[1], p->base[2]), p->roughness, p->metallic, p->transmission, p->ior,
                  p->entering != 0 };
    const D3 o = d3(wo[0], wo[1], wo[2]);
    for (uint64_t i = 0; i < n; ++i) {
        D3 ff;
        bsdfEval(m, o, d3(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]), ff, pdf[i]);
        f[3 * i] = ff.x;
        f[3 * i + 1] = ff.y;
        f[3 * i + 2] = ff.z;
    }
}

void or_bsdf_sample(const OrBsdfParams* p, const double wo[3], uint64_t n, const double* u, double* wi, double* f,
                    double* pdf, int32_t* branch, uint8_t* valid) {
    const Bsdf m{ d3(p->base[0], p->base[1], p->base[2]), p->roughness, p->metallic, p->transmission, p->ior,
                  p->entering != 0 };
    const D3 o = d3(wo[0], wo[1], wo[2]);
    for (uint64_t i = 0; i < n; ++i) {
        const double* ui = u + 4 * i;
        D3 w{ 0, 0, 0 }, ff{ 0, 0, 0 };
        double pp = 0;
        int br = -1;
        const bool ok = bsdfSampleU(m, o, ui[0], ui[1], ui[2], [&] { return ui[3]; }, w, ff, pp, br);
        wi[3 * i] = w.x; wi[3 * i + 1] = w.y; wi[3 * i + 2] = w.z;
        f[3 * i] = ff.x; f[3 * i + 1] = ff.y; f[3 * i + 2] = ff.z;
        pdf[i] = pp;
        branch[i] = br;
        valid[i] = ok ? 1 : 0;
    }
}

void or_light_sample(const float* l, uint32_t pickCount, const double p[3], const double u[3], double out[10]) {
    const LS ls = sampleAnalytic(l, 1.0 / std::max(pickCount, 1u), d3(p[0], p[1], p[2]), u[1], u[2]);
    const double v[10] = { ls.wi.x, ls.wi.y, ls.wi.z, ls.rad.x, ls.rad.y, ls.rad.z, ls.pdf, ls.dist,
                           ls.delta ? 1.0 : 0.0, ls.valid ? 1.0 : 0.0 };
    std::memcpy(out, v, sizeof(v));
}

void or_texture_sample(const OrTexture* t, double u, double v, double out[4]) {
    const D4 c = texSample(t, u, v);
    for (int k = 0; k < 4; ++k) out[k] = c[k];
}

double or_env_cdf_probs(const float* env, uint32_t w, uint32_t h, double* prob) {
    std::vector<float> cdf;
    const double integral = buildCdf(env, w, h, cdf);
    const float* cond = cdf.data();
    const float* marg = cdf.data() + size_t(w) * h;
    // searchCdf returns the first index whose CDF value is >= u; the last
    // index also absorbs u above the final CDF value.
    for (uint32_t y = 0; y < h; ++y) {
        double pr = double(marg[y]) - (y ? double(marg[y - 1]) : 0.0);
        if (y + 1 == h) pr += 1.0 - double(marg[y]);
        for (uint32_t x = 0; x < w; ++x) {
            const float* c = cond + size_t(y) * w;
            double pc = double(c[x]) - (x ? double(c[x - 1]) : 0.0);
            if (x + 1 == w) pc += 1.0 - double(c[x]);
            prob[size_t(y) * w + x] = pr * pc;
        }
    }
    return integral;
}

void or_pt_render(const OrPtScene* scene, const OrPtCamera* cam, const OrPtParams* P, double* out,
                  uint64_t* counts) {
    Scene sc;
    sc.s = scene;
    sc.inst.resize(scene->instanceCount);
    for (uint32_t i = 0; i < scene->instanceCount; ++i) {
        const float* f = scene->instances + size_t(i) * 36;
        for (int k = 0; k < 16; ++k) {
            sc.inst[i].o2w[k] = f[k];
            sc.inst[i].w2o[k] = f[16 + k];
        }
        std::memcpy(&sc.inst[i].blasRoot, f + 32, 4);
        std::memcpy(&sc.inst[i].material, f + 33, 4);
    }
    if (scene->env) {
        sc.env.px = scene->env;
        sc.env.w = scene->envW;
        sc.env.h = scene->envH;
        sc.env.rot = scene->envRotation;
        sc.env.intensity = scene->envIntensity;
        const double integral = buildCdf(scene->env, scene->envW, scene->envH, sc.env.cdf);
        sc.env.integral = std::max(float(integral), 1e-8f);
        sc.env.alias = scene->envAlias;
    }
    int threads = P->threads > 0 ? P->threads : or_hardware_threads();
    threads = std::max(1, std::min<int>(threads, int(P->ch)));
    std::vector<std::array<uint64_t, 6>> part(threads);
    std::atomic<uint32_t> nextRow{ 0 };
    auto work = [&](int t) {
        std::array<uint64_t, 6> c{};
        for (uint32_t r; (r = nextRow.fetch_add(1)) < P->ch;) {
            for (uint32_t x = 0; x < P->cw; ++x) {
                D3 acc = d3(0, 0, 0);
                for (uint32_t f = 0; f < P->spp; ++f)
                    acc = acc + samplePath(sc, *cam, *P, P->x0 + x, P->y0 + r, P->frame0 + f,
                                           P->countFetches ? c.data() : nullptr);
                double* o = out + (size_t(r) * P->cw + x) * 3;
                o[0] = acc.x;
                o[1] = acc.y;
                o[2] = acc.z;
            }
        }
        part[t] = c;
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t) pool.emplace_back(work, t);
    work(0);
    for (auto& th : pool) th.join();
    if (counts) {
        for (int k = 0; k < 6; ++k) counts[k] = 0;
        for (auto& c : part)
            for (int k = 0; k < 6; ++k) counts[k] += c[k];
    }
}

} // extern "C"
