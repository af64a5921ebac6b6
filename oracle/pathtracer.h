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
// alias table.  Textures are not modelled (parity unpinned for textured
// materials; the GPU tests pin texture sampling with exact-emission scenes).
#pragma once

#include "oracle.h"

#include <cstdint>

extern "C" {

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

// Environment integral and per-texel sampling probabilities of the
// reference's CDF construction (Renderer.cpp:620-652), in double:
// prob[y*w + x] = probability the CDF sampler returns texel (x, y).
double or_env_cdf_probs(const float* env, uint32_t w, uint32_t h, double* prob);
}
