// Device math for the path tracer: vectors, RNG, BSDF, lights, environment,
// textures.  Each block names the reference function whose semantics it
// follows (core/shaders/*.glsl); the code is written for HIP/gfx950.
#pragma once

#include "common/gpu_types.h"
#include "pt_params.h"

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mrt {
namespace dev {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kTwoPi = 6.28318530717958647692f;
constexpr float kInvPi = 0.31830988618379067154f;
constexpr float kEpsRay = 1e-4f;  // common.glsl:9

// ------------------------------------------------------------------ vectors --
struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return { x, y, z }; }
__device__ __forceinline__ V3 v3(float4 a) { return { a.x, a.y, a.z }; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return { a.x + b.x, a.y + b.y, a.z + b.z }; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return { a.x - b.x, a.y - b.y, a.z - b.z }; }
__device__ __forceinline__ V3 operator-(V3 a) { return { -a.x, -a.y, -a.z }; }
__device__ __forceinline__ V3 operator*(V3 a, V3 b) { return { a.x * b.x, a.y * b.y, a.z * b.z }; }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return { a.x * s, a.y * s, a.z * s }; }
__device__ __forceinline__ V3 operator*(float s, V3 a) { return { a.x * s, a.y * s, a.z * s }; }
__device__ __forceinline__ V3 operator/(V3 a, float s) { return { a.x / s, a.y / s, a.z / s }; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return { a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x };
}
__device__ __forceinline__ float length(V3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ V3 normalize(V3 a) { return a * (1.0f / sqrtf(dot(a, a))); }
__device__ __forceinline__ float maxComp(V3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }
__device__ __forceinline__ float luminance(V3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }
__device__ __forceinline__ V3 lerp3(V3 a, V3 b, float t) { return a * (1.0f - t) + b * t; }
__device__ __forceinline__ float4 f4(V3 a, float w) { return make_float4(a.x, a.y, a.z, w); }
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }

// GLSL reflect / refract.
__device__ __forceinline__ V3 reflect(V3 i, V3 n) { return i - n * (2.0f * dot(n, i)); }
__device__ __forceinline__ V3 refract(V3 i, V3 n, float eta) {
    const float ni = dot(n, i);
    const float k = 1.0f - eta * eta * (1.0f - ni * ni);
    if (k < 0.0f) return v3(0.0f, 0.0f, 0.0f);
    return i * eta - n * (eta * ni + sqrtf(k));
}

// Orthonormal basis around a unit normal (Duff et al. 2017), common.glsl:94-100.
__device__ __forceinline__ void onb(V3 n, V3& t, V3& b) {
    const float s = n.z >= 0.0f ? 1.0f : -1.0f;
    const float a = -1.0f / (s + n.z);
    const float c = n.x * n.y * a;
    t = v3(1.0f + s * n.x * n.x * a, s * c, -s * n.x);
    b = v3(c, s + n.y * n.y * a, -n.y);
}

// ---------------------------------------------------------------------- RNG --
// pcg4d (Jarzynski & Olano 2020), random.glsl:6-31.  State: (px, py, frame,
// counter); each draw pre-increments the counter and hashes the state.
__device__ __forceinline__ uint4 pcg4d(uint4 v) {
    v.x = v.x * 1664525u + 1013904223u;
    v.y = v.y * 1664525u + 1013904223u;
    v.z = v.z * 1664525u + 1013904223u;
    v.w = v.w * 1664525u + 1013904223u;
    v.x += v.y * v.w; v.y += v.z * v.x; v.z += v.x * v.y; v.w += v.y * v.z;
    v.x ^= v.x >> 16u; v.y ^= v.y >> 16u; v.z ^= v.z >> 16u; v.w ^= v.w >> 16u;
    v.x += v.y * v.w; v.y += v.z * v.x; v.z += v.x * v.y; v.w += v.y * v.z;
    return v;
}

struct Rng {
    uint32_t px, py, frame, counter;
    __device__ __forceinline__ float next1() {
        ++counter;
        const uint4 h = pcg4d(make_uint4(px, py, frame, counter));
        return float(h.x) * (1.0f / 4294967296.0f);
    }
    __device__ __forceinline__ void next2(float& a, float& b) {
        ++counter;
        const uint4 h = pcg4d(make_uint4(px, py, frame, counter));
        a = float(h.x) * (1.0f / 4294967296.0f);
        b = float(h.y) * (1.0f / 4294967296.0f);
    }
};

// ---------------------------------------------------------------- textures --
// Bilinear filtering at level 0 with REPEAT addressing on both axes (the
// reference samples through a linear/REPEAT sampler, GpuContext.cpp:190-198,
// from compute shaders where the implicit level is 0).  sRGB texels are
// decoded before filtering, as the hardware does for *_SRGB formats.
__device__ __forceinline__ float srgbToLinear(float c) {
    return c <= 0.04045f ? c * (1.0f / 12.92f) : powf((c + 0.055f) * (1.0f / 1.055f), 2.4f);
}

__device__ __forceinline__ float4 texel8(const uint32_t* pool, uint32_t idx, bool srgb) {
    const uint32_t p = pool[idx];
    float r = float(p & 0xFFu) * (1.0f / 255.0f);
    float g = float((p >> 8) & 0xFFu) * (1.0f / 255.0f);
    float b = float((p >> 16) & 0xFFu) * (1.0f / 255.0f);
    const float a = float(p >> 24) * (1.0f / 255.0f);
    if (srgb) {
        r = srgbToLinear(r);
        g = srgbToLinear(g);
        b = srgbToLinear(b);
    }
    return make_float4(r, g, b, a);
}

__device__ __forceinline__ int wrapi(int i, int n) {
    int m = i % n;
    return m < 0 ? m + n : m;
}

struct Bilerp {
    uint32_t i00, i10, i01, i11;
    float fx, fy;
};
__device__ __forceinline__ Bilerp bilerpSetup(float u, float v, uint32_t w, uint32_t h) {
    const float x = u * float(w) - 0.5f, y = v * float(h) - 0.5f;
    const float x0f = floorf(x), y0f = floorf(y);
    const int x0 = wrapi(int(x0f), int(w)), y0 = wrapi(int(y0f), int(h));
    const int x1 = x0 + 1 == int(w) ? 0 : x0 + 1;
    const int y1 = y0 + 1 == int(h) ? 0 : y0 + 1;
    Bilerp b;
    b.i00 = uint32_t(y0) * w + uint32_t(x0);
    b.i10 = uint32_t(y0) * w + uint32_t(x1);
    b.i01 = uint32_t(y1) * w + uint32_t(x0);
    b.i11 = uint32_t(y1) * w + uint32_t(x1);
    b.fx = x - x0f;
    b.fy = y - y0f;
    return b;
}
__device__ __forceinline__ float4 mix4(float4 a, float4 b, float t) {
    return make_float4(a.x + (b.x - a.x) * t, a.y + (b.y - a.y) * t, a.z + (b.z - a.z) * t,
                       a.w + (b.w - a.w) * t);
}

__device__ __forceinline__ float4 sampleTexture(const SceneView& sc, int slot, float u, float v) {
    const GpuTexture t = sc.texTable[slot];
    if (t.width == 0) return make_float4(1.0f, 1.0f, 1.0f, 1.0f);  // removed slot: white dummy
    if (!(u == u) || !(v == v) || fabsf(u) > 1e7f || fabsf(v) > 1e7f) { u = 0.0f; v = 0.0f; }
    const Bilerp b = bilerpSetup(u, v, t.width, t.height);
    float4 c00, c10, c01, c11;
    if (t.format == kTexRgba32f) {
        c00 = sc.tex32[t.offset + b.i00]; c10 = sc.tex32[t.offset + b.i10];
        c01 = sc.tex32[t.offset + b.i01]; c11 = sc.tex32[t.offset + b.i11];
    } else {
        const bool srgb = t.format == kTexRgba8Srgb;
        c00 = texel8(sc.tex8, t.offset + b.i00, srgb); c10 = texel8(sc.tex8, t.offset + b.i10, srgb);
        c01 = texel8(sc.tex8, t.offset + b.i01, srgb); c11 = texel8(sc.tex8, t.offset + b.i11, srgb);
    }
    return mix4(mix4(c00, c10, b.fx), mix4(c01, c11, b.fx), b.fy);
}

// ------------------------------------------------------------- environment --
// lights.glsl:11-58 (equirect mapping, bilinear radiance and pdf).  Sampling
// uses a Vose alias table over the same discrete texel distribution as the
// reference's marginal/conditional CDFs (Renderer.cpp:620-652).
__device__ __forceinline__ void dirToEquirect(V3 d, float rot, float& u, float& v) {
    const float phi = atan2f(d.z, d.x) - rot;
    const float theta = acosf(clampf(d.y, -1.0f, 1.0f));
    const float x = phi * (1.0f / kTwoPi) + 1.0f;
    u = x - floorf(x);
    v = theta * kInvPi;
}
__device__ __forceinline__ V3 equirectToDir(float u, float v, float rot) {
    const float phi = u * kTwoPi + rot;
    const float theta = v * kPi;
    float st, ct, sp, cp;
    sincosf(theta, &st, &ct);
    sincosf(phi, &sp, &cp);
    return v3(st * cp, ct, st * sp);
}
__device__ __forceinline__ float4 envLookup(const SceneView& sc, const FrameConsts& g, float u, float v) {
    const Bilerp b = bilerpSetup(u, v, g.envWidth, g.envHeight);
    const float4 c00 = sc.envMap[b.i00], c10 = sc.envMap[b.i10];
    const float4 c01 = sc.envMap[b.i01], c11 = sc.envMap[b.i11];
    return mix4(mix4(c00, c10, b.fx), mix4(c01, c11, b.fx), b.fy);
}
__device__ __forceinline__ V3 envRadiance(const SceneView& sc, const FrameConsts& g, V3 d) {
    if (!(g.flags & kFlagHasEnv)) return v3(0.0f, 0.0f, 0.0f);
    float u, v;
    dirToEquirect(d, g.envRotation, u, v);
    return v3(envLookup(sc, g, u, v)) * g.envIntensity;
}
// Solid-angle pdf of environment sampling: lum(bilinear(dir)) / integral.
__device__ __forceinline__ float envPdf(const SceneView& sc, const FrameConsts& g, V3 d) {
    if (!(g.flags & kFlagHasEnv)) return 0.0f;
    float u, v;
    dirToEquirect(d, g.envRotation, u, v);
    return luminance(v3(envLookup(sc, g, u, v))) / g.envIntegral;
}
__device__ __forceinline__ float envPickPdf(const FrameConsts& g) {
    return ((g.flags & kFlagHasEnv) && g.lightCount > 0) ? 1.0f / float(g.lightCount) : 0.0f;
}

// Two draws, as the reference's row and column CDF searches consume: the
// first picks a texel slot uniformly, the second keeps it or takes its alias.
__device__ __forceinline__ bool sampleEnv(const SceneView& sc, const FrameConsts& g, Rng& rng, V3& dir,
                                          V3& rad, float& pdf) {
    const uint32_t w = g.envWidth, h = g.envHeight;
    const float u1 = rng.next1();
    const float u2 = rng.next1();
    if (w == 0 || h == 0) return false;
    const uint32_t n = w * h;
    uint32_t i = uint32_t(u1 * float(n));
    if (i >= n) i = n - 1;
    const AliasEntry e = sc.envAlias[i];
    const uint32_t texel = u2 < e.prob ? i : e.alias;
    const uint32_t row = texel / w, col = texel - row * w;
    const float u = (float(col) + 0.5f) / float(w), v = (float(row) + 0.5f) / float(h);
    dir = equirectToDir(u, v, g.envRotation);
    rad = v3(envLookup(sc, g, u, v)) * g.envIntensity;
    pdf = envPdf(sc, g, dir);
    return pdf > 1e-9f;
}

// ----------------------------------------------- environment strata (guided) --
// Guided shading may draw environment NEE from two strata (FrameConsts::
// envSplitA): the bright texels (a sun) with probability s and the rest with
// 1 - s, each through its own alias table over the same texel weights.  A
// texel is then drawn with sampleEnv's probability times r = s / share or
// (1 - s) / (1 - share); s comes from per-cell statistics of the shading
// point (neeSelect).  The sample's importance weight divides by that
// probability, but its MIS weight (and that of a continuation reaching the
// environment) uses sampleEnv's pdf: NEE returns texel-centre directions
// while a continuation's pdf is the bilinear lookup, and a ratio that jumps
// at the stratum border would weigh the two differently across the sun's
// edge texels (guided rendered the room 0.14 % brighter than plain).  With
// plain MIS weights, each technique's expectation is the plain tracer's for
// any s in (0, 1).
// sampleEnv over the strata: the same two draws (the first chooses the
// stratum and, rescaled, its entry; the second keeps the entry or takes its
// alias).  pdf: of the draw; plainPdf: sampleEnv's.  stratum: 1 bright, 0
// the rest.
__device__ __forceinline__ bool sampleEnvStrata(const SceneView& sc, const FrameConsts& g, Rng& rng, float s,
                                                V3& dir, V3& rad, float& pdf, float& plainPdf, int& stratum) {
    const uint32_t w = g.envWidth, h = g.envHeight;
    const float u1 = rng.next1();
    const float u2 = rng.next1();
    if (w == 0 || h == 0) return false;
    const bool bright = u1 < s;
    const float uu = bright ? u1 / s : (u1 - s) / (1.0f - s);
    const uint32_t n = bright ? g.envCountA : g.envCountB;
    const EnvStratumEntry* t = bright ? sc.envA : sc.envB;
    uint32_t i = uint32_t(uu * float(n));
    if (i >= n) i = n - 1;
    const EnvStratumEntry e = t[i];
    const uint32_t texel = u2 < e.prob ? e.texel : t[e.alias].texel;
    const uint32_t row = texel / w, col = texel - row * w;
    const float u = (float(col) + 0.5f) / float(w), v = (float(row) + 0.5f) / float(h);
    dir = equirectToDir(u, v, g.envRotation);
    rad = v3(envLookup(sc, g, u, v)) * g.envIntensity;
    plainPdf = envPdf(sc, g, dir);
    pdf = plainPdf * (bright ? s / g.envSplitA : (1.0f - s) / (1.0f - g.envSplitA));
    stratum = bright ? 1 : 0;
    return plainPdf > 1e-9f;
}
// Cell of a shading point for the stratum statistics.
__device__ __forceinline__ uint32_t neeCell(const FrameConsts& g, V3 pos, V3 n) {
    const V3 q = (pos - v3(g.guideBoxMin)) * v3(g.guideBoxScale);
    const int G = int(kNeeGrid);
    const int ix = min(max(int(q.x * float(G)), 0), G - 1);
    const int iy = min(max(int(q.y * float(G)), 0), G - 1);
    const int iz = min(max(int(q.z * float(G)), 0), G - 1);
    const float ax = fabsf(n.x), ay = fabsf(n.y), az = fabsf(n.z);
    const int axis = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
    const float c = axis == 0 ? n.x : (axis == 1 ? n.y : n.z);
    return uint32_t(((ix * G + iy) * G + iz) * 6 + axis * 2 + (c < 0.0f ? 1 : 0));
}
// Bright-stratum probability of a cell: the variance-minimising choice of a
// one-sample estimator over two strata, s = sqrt(M_A) / (sqrt(M_A) +
// sqrt(M_B)) with M the mean squared stratum-conditional contribution, kept
// in [0.05, 0.95]; the plain share until both strata have 16 samples.
__device__ __forceinline__ float neeSelect(const float4* __restrict__ stats, uint32_t cell, float share) {
    const float4 c = stats[cell];
    if (c.y < 16.0f || c.w < 16.0f) return share;
    const float a = sqrtf(c.x / c.y), b = sqrtf(c.z / c.w);
    if (!(a + b > 0.0f)) return share;
    return clampf(a / (a + b), 0.05f, 0.95f);
}

// ------------------------------------------------------------------- lights --
struct LightSample {
    V3 wi;
    V3 radiance;
    float pdf;     // of the draw (the contribution divides by it)
    float misPdf;  // of the plain light distribution (MIS weights and validity use it)
    float dist;
    bool delta;
    bool valid;
    int stratum;  // environment samples of sampleEnvStrata: 1 bright, 0 the rest; -1 otherwise
};

// Uniform pick over the light list (environment at slot 0 when present),
// lights.glsl:70-134.  STRATA: the environment is drawn over its strata with
// bright-stratum probability s (guided shading).
template <bool STRATA = false>
__device__ __forceinline__ LightSample sampleOneLight(const SceneView& sc, const FrameConsts& g, V3 p,
                                                      Rng& rng, float s = 0.0f) {
    LightSample ls;
    ls.valid = false;
    ls.delta = false;
    ls.pdf = 0.0f;
    ls.misPdf = 0.0f;
    ls.dist = 0.0f;
    ls.stratum = -1;
    if (g.lightCount == 0) return ls;
    uint32_t idx = uint32_t(rng.next1() * float(g.lightCount));
    if (idx > g.lightCount - 1) idx = g.lightCount - 1;
    const float pick = 1.0f / float(g.lightCount);
    const GpuLight l = sc.lights[idx];
    if (l.type == kLightEnv) {
        V3 dir, rad;
        float pdf, plainPdf;
        if (STRATA) {
            if (!sampleEnvStrata(sc, g, rng, s, dir, rad, pdf, plainPdf, ls.stratum)) return ls;
        } else if (!sampleEnv(sc, g, rng, dir, rad, pdf)) {
            return ls;
        } else {
            plainPdf = pdf;
        }
        ls.wi = dir;
        ls.radiance = rad;
        ls.pdf = pdf * pick;
        ls.misPdf = plainPdf * pick;
        ls.dist = 1e30f;
        ls.valid = true;
    } else if (l.type == kLightSun) {
        const float cosMax = l.p0.w;
        float u0, u1;
        rng.next2(u0, u1);
        const float cosT = 1.0f + (cosMax - 1.0f) * u0;
        const float sinT = sqrtf(fmaxf(1.0f - cosT * cosT, 0.0f));
        float sp, cp;
        sincosf(kTwoPi * u1, &sp, &cp);
        V3 t, b;
        const V3 axis = v3(l.p0);
        onb(axis, t, b);
        ls.wi = normalize(t * (sinT * cp) + b * (sinT * sp) + axis * cosT);
        ls.radiance = v3(l.p1);
        ls.pdf = pick / (kTwoPi * fmaxf(1.0f - cosMax, 1e-7f));
        ls.dist = 1e30f;
        ls.delta = true;
        ls.valid = true;
    } else if (l.type == kLightPoint) {
        const V3 d = v3(l.p0) - p;
        const float d2 = fmaxf(dot(d, d), 1e-8f);
        ls.dist = sqrtf(d2);
        ls.wi = d / ls.dist;
        ls.radiance = v3(l.p1) / d2;
        ls.pdf = pick;
        ls.delta = true;
        ls.valid = true;
    } else {
        float u0, u1;
        rng.next2(u0, u1);
        const V3 q = v3(l.p0) + v3(l.p1) * u0 + v3(l.p2) * u1;
        const V3 d = q - p;
        const float d2 = fmaxf(dot(d, d), 1e-8f);
        ls.dist = sqrtf(d2);
        ls.wi = d / ls.dist;
        const V3 n = normalize(cross(v3(l.p1), v3(l.p2)));
        float cosL = dot(n, -ls.wi);
        if (l.p3.w > 0.5f) cosL = fabsf(cosL);
        if (cosL < 1e-6f) return ls;
        ls.radiance = v3(l.p3);
        ls.pdf = (d2 / (cosL * fmaxf(l.p0.w, 1e-8f))) * pick;
        ls.delta = true;
        ls.valid = true;
    }
    if (l.type != kLightEnv) ls.misPdf = ls.pdf;
    return ls;
}

// --------------------------------------------------------------------- BSDF --
// "Disney-lite" Lambert + GGX reflection + GGX transmission, bsdf.glsl:29-192.
struct Bsdf {
    V3 base;
    float roughness, metallic, transmission, ior;
    bool entering;
};

__device__ __forceinline__ float ggxD(V3 h, float alpha) {
    const float a2 = alpha * alpha;
    const float c2 = h.z * h.z;
    const float d = c2 * (a2 - 1.0f) + 1.0f;
    return a2 / fmaxf(kPi * d * d, 1e-12f);
}
__device__ __forceinline__ float ggxLambda(V3 w, float alpha) {
    const float c2 = w.z * w.z;
    const float t2 = fmaxf(1.0f - c2, 0.0f) / fmaxf(c2, 1e-12f);
    return 0.5f * (sqrtf(1.0f + alpha * alpha * t2) - 1.0f);
}
__device__ __forceinline__ float ggxG1(V3 w, float alpha) { return 1.0f / (1.0f + ggxLambda(w, alpha)); }
__device__ __forceinline__ float ggxG2(V3 wo, V3 wi, float alpha) {
    return 1.0f / (1.0f + ggxLambda(wo, alpha) + ggxLambda(wi, alpha));
}
// Visible-normal sampling (Heitz 2018).
__device__ __forceinline__ V3 sampleVndf(V3 wo, float alpha, float u0, float u1) {
    const V3 vh = normalize(v3(alpha * wo.x, alpha * wo.y, wo.z));
    const float lensq = vh.x * vh.x + vh.y * vh.y;
    const V3 T1 = lensq > 0.0f ? v3(-vh.y, vh.x, 0.0f) * (1.0f / sqrtf(lensq)) : v3(1.0f, 0.0f, 0.0f);
    const V3 T2 = cross(vh, T1);
    const float r = sqrtf(u0);
    float sp, cp;
    sincosf(kTwoPi * u1, &sp, &cp);
    const float t1 = r * cp;
    float t2 = r * sp;
    const float s = 0.5f * (1.0f + vh.z);
    t2 = (1.0f - s) * sqrtf(fmaxf(1.0f - t1 * t1, 0.0f)) + s * t2;
    const V3 nh = T1 * t1 + T2 * t2 + vh * sqrtf(fmaxf(1.0f - t1 * t1 - t2 * t2, 0.0f));
    return normalize(v3(alpha * nh.x, alpha * nh.y, fmaxf(nh.z, 1e-6f)));
}
__device__ __forceinline__ V3 schlick(V3 f0, float cosT) {
    const float m = clampf(1.0f - cosT, 0.0f, 1.0f);
    const float m2 = m * m;
    const float w = m2 * m2 * m;
    return f0 + (v3(1.0f, 1.0f, 1.0f) - f0) * w;
}
__device__ __forceinline__ float fresnelDielectric(float cosI, float eta) {
    cosI = clampf(cosI, 0.0f, 1.0f);
    const float sin2T = (1.0f - cosI * cosI) / (eta * eta);
    if (sin2T >= 1.0f) return 1.0f;
    const float cosT = sqrtf(1.0f - sin2T);
    const float rs = (cosI - eta * cosT) / (cosI + eta * cosT);
    const float rp = (eta * cosI - cosT) / (eta * cosI + cosT);
    return 0.5f * (rs * rs + rp * rp);
}

struct LobeW {
    float diff, spec, trans;
};
__device__ __forceinline__ V3 f0Of(const Bsdf& m) {
    return lerp3(v3(0.04f, 0.04f, 0.04f), m.base, m.metallic);
}
__device__ __forceinline__ LobeW lobeWeights(const Bsdf& m, V3 wo) {
    LobeW w;
    w.spec = luminance(schlick(f0Of(m), fabsf(wo.z)));
    w.diff = (1.0f - m.metallic) * (1.0f - m.transmission) * luminance(m.base);
    w.trans = (1.0f - m.metallic) * m.transmission;
    float total = w.diff + w.spec + w.trans;
    if (total < 1e-7f) {
        w.diff = 1.0f; w.spec = 0.0f; w.trans = 0.0f; total = 1.0f;
    }
    w.diff /= total; w.spec /= total; w.trans /= total;
    return w;
}

// f and the combined sampling pdf for (wo, wi) in the local frame (wo.z > 0).
__device__ __forceinline__ void bsdfEval(const Bsdf& m, V3 wo, V3 wi, V3& f, float& pdf) {
    f = v3(0.0f, 0.0f, 0.0f);
    pdf = 0.0f;
    const float alpha = fmaxf(m.roughness * m.roughness, 1e-3f);
    const LobeW w = lobeWeights(m, wo);
    const float eta = m.entering ? m.ior : 1.0f / m.ior;
    if (wi.z > 0.0f) {
        if (w.diff > 0.0f) {
            f = f + m.base * ((1.0f - m.metallic) * (1.0f - m.transmission) * kInvPi);
            pdf += w.diff * wi.z * kInvPi;
        }
        const V3 h = normalize(wo + wi);
        const float d = ggxD(h, alpha);
        const float g2 = ggxG2(wo, wi, alpha);
        const V3 F = schlick(f0Of(m), dot(wo, h));
        f = f + F * (d * g2 / fmaxf(4.0f * wo.z * wi.z, 1e-7f));
        pdf += w.spec * ggxG1(wo, alpha) * d / fmaxf(4.0f * wo.z, 1e-7f);
    } else if (w.trans > 0.0f) {
        V3 h = normalize(wo + wi * eta);
        if (h.z < 0.0f) h = -h;
        const float woH = dot(wo, h), wiH = dot(wi, h);
        if (woH * wiH < 0.0f) {
            const float F = fresnelDielectric(fabsf(woH), eta);
            const float d = ggxD(h, alpha);
            const float g2 = ggxG2(wo, v3(wi.x, wi.y, -wi.z), alpha);
            float denom = woH + eta * wiH;
            denom = fmaxf(denom * denom, 1e-7f);
            const float jac = eta * eta * fabsf(wiH) / denom;
            const float s = (1.0f - F) * d * g2 * fabsf(woH) * jac / fmaxf(fabsf(wo.z * wi.z), 1e-7f) *
                            (1.0f - m.metallic) * m.transmission;
            f = f + m.base * s;
            pdf += w.trans * (1.0f - F) * ggxG1(wo, alpha) * d * fabsf(woH) / fmaxf(wo.z, 1e-7f) * jac;
        }
    }
}

struct BsdfSample {
    V3 wi, f;
    float pdf;
    bool valid;
};

// Lobe pick (rng1), direction (rng2), Fresnel choice for transmission (rng1);
// bsdf.glsl:147-192.
__device__ __forceinline__ BsdfSample bsdfSample(const Bsdf& m, V3 wo, Rng& rng) {
    BsdfSample s;
    s.valid = false;
    const float alpha = fmaxf(m.roughness * m.roughness, 1e-3f);
    const LobeW w = lobeWeights(m, wo);
    const float eta = m.entering ? m.ior : 1.0f / m.ior;
    const float pick = rng.next1();
    V3 wi;
    float u0, u1;
    rng.next2(u0, u1);
    if (pick < w.diff) {
        const float r = sqrtf(u0);
        float sp, cp;
        sincosf(kTwoPi * u1, &sp, &cp);
        wi = v3(r * cp, r * sp, sqrtf(fmaxf(1.0f - u0, 0.0f)));
    } else if (pick < w.diff + w.spec) {
        const V3 h = sampleVndf(wo, alpha, u0, u1);
        wi = reflect(-wo, h);
        if (wi.z <= 0.0f) return s;
    } else {
        const V3 h = sampleVndf(wo, alpha, u0, u1);
        const float woH = dot(wo, h);
        const float F = fresnelDielectric(fabsf(woH), eta);
        if (rng.next1() < F) {
            wi = reflect(-wo, h);
            if (wi.z <= 0.0f) return s;
        } else {
            wi = refract(-wo, h, 1.0f / eta);
            if (wi.z >= 0.0f || length(wi) < 1e-6f) return s;
            wi = normalize(wi);
        }
    }
    V3 f;
    float pdf;
    bsdfEval(m, wo, wi, f, pdf);
    if (pdf < 1e-9f) return s;
    s.wi = wi;
    s.f = f;
    s.pdf = pdf;
    s.valid = true;
    return s;
}

__device__ __forceinline__ float balanceHeuristic(float a, float b) { return a / fmaxf(a + b, 1e-12f); }

// -------------------------------------------------------------- wave helpers --
__device__ __forceinline__ uint32_t laneId() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}
// The XCD this wave runs on, modulo 8 (hardware register HW_REG_XCC_ID = 20,
// bits [3:0]); wave-uniform.  Work placement only: any value is correct.
__device__ __forceinline__ uint32_t xccId() {
    return __builtin_amdgcn_s_getreg(20 | (0 << 6) | (3 << 11)) & 7u;
}
// Wave-aggregated append: one atomic per wave; returns this lane's slot
// (only meaningful where pred is true).  All lanes of the wave that are
// active must call it.
__device__ __forceinline__ uint32_t waveAppend(uint32_t* counter, bool pred) {
    const uint64_t mask = __ballot(pred);
    if (mask == 0) return 0;
    const uint32_t rank = __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32),
                                                    __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
    const uint32_t leader = uint32_t(__builtin_ctzll(mask));
    uint32_t base = 0;
    if (laneId() == leader) base = atomicAdd(counter, uint32_t(__builtin_popcountll(mask)));
    base = __builtin_amdgcn_readlane(base, leader);
    return base + rank;
}
// Persistent-kernel work fetch: 64 consecutive items per wave per call.
__device__ __forceinline__ uint32_t waveFetch(uint32_t* cursor) {
    uint32_t base = 0;
    if (laneId() == 0) base = atomicAdd(cursor, 64u);
    return __builtin_amdgcn_readfirstlane(base);
}

} // namespace dev
} // namespace mrt
