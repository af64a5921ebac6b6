// Path-guiding distribution on the device: feature encoding of a shading
// point, decoding of the network's 40 outputs into a mixture of 8 von
// Mises-Fisher lobes, and the mixture's pdf and sampling routine.
//
// No counterpart in the reference (one of the three additions on top of the
// reference path tracer).  Numerics follow the FP64 oracle (oracle/guide.py):
//   vMF(w; mu, k) = k / (2 pi (1 - e^{-2k})) * exp(k (mu.w - 1))
//   log k clamped to [-5, 7]; lobe weights = softmax of the logits.
#pragma once

#include "pt_device.h"

namespace mrt {
namespace dev {

constexpr int kGuideIn = 78;       // encoded features
constexpr int kGuideInPad = 80;    // padded to a multiple of the MFMA k step
constexpr int kGuideHidden = 64;
constexpr int kGuideOut = 40;      // 8 lobes x (mu 3, log kappa, logit)
constexpr int kGuideLobes = 8;
constexpr int kGuideBands = 12;
constexpr int kGuideParams = 78 * 64 + 64 + 3 * (64 * 64 + 64) + 64 * 40 + 40;  // 20,136
constexpr float kLogKappaLo = -5.0f, kLogKappaHi = 7.0f;

// Offsets of each layer's weights and biases in the flat parameter vector
// (weights output-major: row n holds the weights of output n).
__host__ __device__ constexpr int guideLayerIn(int l) { return l == 0 ? kGuideIn : kGuideHidden; }
__host__ __device__ constexpr int guideLayerOut(int l) { return l == 4 ? kGuideOut : kGuideHidden; }
__host__ __device__ constexpr int guideWeightOffset(int l) {
    return l == 0 ? 0 : guideWeightOffset(l - 1) + guideLayerIn(l - 1) * guideLayerOut(l - 1) + guideLayerOut(l - 1);
}
__host__ __device__ constexpr int guideBiasOffset(int l) {
    return guideWeightOffset(l) + guideLayerIn(l) * guideLayerOut(l);
}
static_assert(guideBiasOffset(4) + kGuideOut == kGuideParams, "guide parameter layout");

// Features of a shading point (pos01 = position normalised to the scene
// box): for axis a (x, y, z) and band b < 12 the pair sin(pi 2^b p_a),
// cos(pi 2^b p_a) at 24 a + 2 b; then the normal (72..74) and wo (75..77);
// 78 and 79 are zero pads.  guide.hip's encodeRows forms them.

struct GuideMix {
    V3 mu[kGuideLobes];
    float kappa[kGuideLobes];
    float w[kGuideLobes];
};

// Mean direction and concentration of lobe k from its raw outputs.
__device__ __forceinline__ void decodeLobe(const float* raw, int k, V3& mu, float& kappa) {
    const V3 v = v3(raw[5 * k], raw[5 * k + 1], raw[5 * k + 2]);
    const float len = sqrtf(dot(v, v));
    mu = len > 1e-12f ? v * (1.0f / len) : v3(0.0f, 0.0f, 1.0f);
    kappa = expf(clampf(raw[5 * k + 3], kLogKappaLo, kLogKappaHi));
}

// The softmax sums lobes 0..3 and 4..7 separately and then adds the two, the
// order storeMixHalf uses, so both decode the same weights.
__device__ __forceinline__ void decodeMix(const float* raw, GuideMix& m) {
    float mx = -1e30f;
    for (int k = 0; k < kGuideLobes; ++k) mx = fmaxf(mx, raw[5 * k + 4]);
    float s[2] = { 0.0f, 0.0f };
    for (int k = 0; k < kGuideLobes; ++k) {
        decodeLobe(raw, k, m.mu[k], m.kappa[k]);
        m.w[k] = expf(raw[5 * k + 4] - mx);
        s[k / 4] += m.w[k];
    }
    const float inv = 1.0f / (s[0] + s[1]);
    for (int k = 0; k < kGuideLobes; ++k) m.w[k] *= inv;
}

__device__ __forceinline__ float vmfPdf(V3 mu, float kappa, V3 w) {
    // kappa / (2 pi (1 - e^{-2 kappa})) without cancellation for small kappa
    const float norm = kappa / (kTwoPi * -expm1f(-2.0f * kappa));
    return norm * expf(kappa * (dot(mu, w) - 1.0f));
}

__device__ __forceinline__ float mixPdf(const GuideMix& m, V3 w) {
    float p = 0.0f;
    for (int k = 0; k < kGuideLobes; ++k) p += m.w[k] * vmfPdf(m.mu[k], m.kappa[k], w);
    return p;
}

// Direction from vMF(mu, kappa) by (u1, u2) (Wenzel Jakob, "Numerically
// stable sampling of the von Mises-Fisher distribution on S^2", 2012).
__device__ __forceinline__ V3 vmfSample(V3 mu, float kap, float u1, float u2) {
    float wz = 1.0f + log1pf((1.0f - u1) * expm1f(-2.0f * kap)) / kap;
    wz = clampf(wz, -1.0f, 1.0f);
    const float sz = sqrtf(fmaxf(1.0f - wz * wz, 0.0f));
    float sp, cp;
    sincosf(kTwoPi * u2, &sp, &cp);
    V3 t, b;
    onb(mu, t, b);
    return normalize(t * (sz * cp) + b * (sz * sp) + mu * wz);
}

// Lobe by u0 over the cumulative weights, then a vMF direction from (u1, u2).
__device__ __forceinline__ V3 mixSample(const GuideMix& m, float u0, float u1, float u2) {
    int k = 0;
    float c = m.w[0];
    while (k < kGuideLobes - 1 && u0 >= c) {
        ++k;
        c += m.w[k];
    }
    return vmfSample(m.mu[k], m.kappa[k], u1, u2);
}

// Decoded mixture as stored for the renderer, kGuideMixF4 float4 per query:
//   [0..7]   (mu.xyz, kappa) of lobe k
//   [8..9]   selection weights w_0..w_7
//   [10..11] pdf prefactors w_k kappa_k / (2 pi (1 - e^{-2 kappa_k}))
constexpr int kGuideMixF4 = 12;

__device__ __forceinline__ void storeMix(const GuideMix& m, float4* __restrict__ out) {
    float pre[kGuideLobes];
    for (int k = 0; k < kGuideLobes; ++k) {
        out[k] = f4(m.mu[k], m.kappa[k]);
        pre[k] = m.w[k] * m.kappa[k] / (kTwoPi * -expm1f(-2.0f * m.kappa[k]));
    }
    out[8] = make_float4(m.w[0], m.w[1], m.w[2], m.w[3]);
    out[9] = make_float4(m.w[4], m.w[5], m.w[6], m.w[7]);
    out[10] = make_float4(pre[0], pre[1], pre[2], pre[3]);
    out[11] = make_float4(pre[4], pre[5], pre[6], pre[7]);
}

// storeMix of lobes 4 h .. 4 h + 3 of one row by one lane; the lanes of the
// other half (lane ^ 32) hold the other four lobes of the same row.  All
// lanes of the wave must call it (the softmax is shared through shuffles).
__device__ __forceinline__ void storeMixHalf(const float* raw, int h, bool store, float4* __restrict__ out) {
    float mx = -1e30f;
    for (int k = 4 * h; k < 4 * h + 4; ++k) mx = fmaxf(mx, raw[5 * k + 4]);
    mx = fmaxf(mx, __shfl_xor(mx, 32, 64));
    V3 mu[4];
    float kappa[4], w[4], s = 0.0f;
    for (int i = 0; i < 4; ++i) {
        const int k = 4 * h + i;
        decodeLobe(raw, k, mu[i], kappa[i]);
        w[i] = expf(raw[5 * k + 4] - mx);
        s += w[i];
    }
    const float other = __shfl_xor(s, 32, 64);
    const float inv = 1.0f / (h == 0 ? s + other : other + s);
    float pre[4];
    for (int i = 0; i < 4; ++i) {
        w[i] *= inv;
        pre[i] = w[i] * kappa[i] / (kTwoPi * -expm1f(-2.0f * kappa[i]));
    }
    if (!store) return;
    for (int i = 0; i < 4; ++i) out[4 * h + i] = f4(mu[i], kappa[i]);
    out[8 + h] = make_float4(w[0], w[1], w[2], w[3]);
    out[10 + h] = make_float4(pre[0], pre[1], pre[2], pre[3]);
}

__device__ __forceinline__ float storedMixPdf(const float4* __restrict__ mix, V3 w) {
    const float4 a = mix[10], b = mix[11];
    const float pre[kGuideLobes] = { a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w };
    float p = 0.0f;
    for (int k = 0; k < kGuideLobes; ++k) {
        const float4 l = mix[k];
        p += pre[k] * expf(l.w * (l.x * w.x + l.y * w.y + l.z * w.z - 1.0f));
    }
    return p;
}

__device__ __forceinline__ V3 storedMixSample(const float4* __restrict__ mix, float u0, float u1, float u2) {
    const float4 a = mix[8], b = mix[9];
    const float w[kGuideLobes] = { a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w };
    // Unrolled scan with constant indices keeps w in registers.
    int k = kGuideLobes - 1;
    float c = 0.0f;
    bool found = false;
#pragma unroll
    for (int j = 0; j < kGuideLobes - 1; ++j) {
        c += w[j];
        if (!found && u0 < c) {
            k = j;
            found = true;
        }
    }
    const float4 l = mix[k];
    return vmfSample(v3(l), l.w, u1, u2);
}

} // namespace dev
} // namespace mrt
