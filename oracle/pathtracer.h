// Oracle path tracer: FP64 restatement of the reference progressive path
// tracer (core/shaders/pathtrace.comp:39-233, bsdf.glsl:29-192,
// lights.glsl:11-140, random.glsl:6-31).
//
// TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// It renders a crop of the image for a range of frames with the reference's
// random-number stream (pcg4d keyed by pixel, frame and a per-path counter,
// same draw order), accumulating radiance in double.  Geometry is read from a
// packed scene (the same node / triangle / instance / material / light
// layouts the GPU consumes), traversed in double precision with a watertight
// test.  Environment light sampling uses either the reference's
// marginal/conditional CDFs (built here, Renderer.cpp:605-662) or a supplied
// alias table.  Textures are sampled at level 0 with bilinear filtering and
// REPEAT addressing (a compute shader's implicit LOD is the base level),
// after per-texel sRGB decoding for RGBA8_SRGB; normal maps follow
// pathtrace.comp:65-71.
#pragma once

#include "oracle.h"

#include <cstdint>

extern "C" {

// Texture slot (the engine's 64-entry table): format 0 = RGBA8 sRGB,
// 1 = RGBA8 UNORM, 2 = RGBA32F (mini_raytrace.h texture formats).  width 0
// = empty slot (samples as white, the reference's dummy texture).
struct OrTexture {
    uint32_t width, height, format;
    const void* pixels;  // row-major, top row first; 4 bytes or 4 floats per texel
};

struct OrPtScene {
    const OrNode* nodes;
    uint32_t tlasRoot;
    uint32_t instanceCount;
    const uint32_t* triangles;  // 4 per triangle (i0, i1, i2, pad), leaf order
    const float* positions;     // 4 per vertex
    const float* normals;       // 4 per vertex
    const float* instances;     // 36 per instance: objectToWorld cols, worldToObject cols, blasRoot, material
    const float* materials;     // 20 per material (GpuMaterial layout)
    const float* lights;        // 20 per light (GpuLight layout)
    uint32_t lightCount;
    const float* env;           // envW x envH x 4 (NULL = no environment)
    uint32_t envW, envH;
    float envRotation, envIntensity;
    const uint32_t* envAlias;   // envW*envH x (prob bits, alias) or NULL = reference CDFs
    const float* uvs;           // 2 per vertex (NULL = all zero)
    const float* tangents;      // 4 per vertex, w = handedness (NULL = (1,0,0,1))
    const OrTexture* textures;  // slot table indexed by the materials' texture slots
    uint32_t textureCount;
};

struct OrPtCamera {
    float pos[3], right[3], up[3], forward[3];
    float frustum[4];  // left, right, bottom, top
    int32_t ortho;
};

struct OrPtParams {
    uint32_t width, height;     // full image
    uint32_t x0, y0, cw, ch;    // crop
    uint32_t frame0, spp;       // frames frame0 .. frame0 + spp - 1
    uint32_t maxBounces;        // clamped to 16
    float fireflyClamp;         // 0 = off
    uint32_t seedMix;           // xor-ed into the frame word of the RNG key
    int32_t threads;            // 0 = all hardware threads
    int32_t countFetches;       // 1: also count reference-order fetches (float GLSL traversal)
};

// counts (may be NULL): [closest queries, nodes popped, triangles tested,
//                        shadow queries, shadow nodes, shadow triangles]
// in the reference traversal order of core/shaders/intersect.glsl:48-178.
void or_pt_render(const OrPtScene* scene, const OrPtCamera* cam, const OrPtParams* params,
                  double* outRgb, uint64_t* counts);

// Camera basis and frustum of the legacy look-at camera (Renderer.cpp:724-745).
void or_camera_lookat(const float pos[3], const float target[3], const float up[3], float fovYDeg,
                      uint32_t width, uint32_t height, OrPtCamera* out);

// ---- pieces of the path tracer, exported for the statistical and
// closed-form tests of tests/test_oracle_pt.py ---------------------------------
// BSDF of bsdf.glsl:29-192 in the local frame (z = shading normal).
struct OrBsdfParams {
    double base[3];
    double roughness, metallic, transmission, ior;
    int32_t entering;  // front-face hit
};
// n evaluations for one wo: f (3 per wi) and the combined lobe pdf.
void or_bsdf_eval(const OrBsdfParams* m, const double wo[3], uint64_t n, const double* wi, double* f, double* pdf);
// n samples for one wo; u: 4 per sample = (lobe pick, direction u0,
// direction u1, Fresnel choice), the draws bsdfSample makes in that order
// (the Fresnel draw only on the transmission lobe).  valid[i] = 1 for a
// valid sample; branch[i] = 0 diffuse, 1 specular, 2 transmission-lobe
// reflection, 3 refraction (also set for a rejected sample).
void or_bsdf_sample(const OrBsdfParams* m, const double wo[3], uint64_t n, const double* u, double* wi, double* f,
                    double* pdf, int32_t* branch, uint8_t* valid);
// sampleOneLight (lights.glsl:70-134) for an analytic light record (GpuLight
// layout, 20 floats) at point p with uniforms u (u[0] unused: the light pick
// is the caller's; u[1], u[2] the light's own rng2).  pickCount = lights in
// the scene.  out: wi xyz, radiance rgb, pdf, dist, delta, valid.
void or_light_sample(const float* light20, uint32_t pickCount, const double p[3], const double u[3], double out[10]);
// Texture lookup at (u, v): bilinear, REPEAT, level 0, decoded to linear.
void or_texture_sample(const OrTexture* tex, double u, double v, double out[4]);

// Environment integral and per-texel sampling probabilities of the
// reference's CDF construction (Renderer.cpp:620-652), in double:
// prob[y*w + x] = probability the CDF sampler returns texel (x, y).
double or_env_cdf_probs(const float* env, uint32_t w, uint32_t h, double* prob);
}
