// Device buffer layouts shared by the host packer and the HIP kernels.
//
// GpuInstance, GpuMaterial, GpuLight and GpuTriangle keep the byte layout of
// the reference's std430 mirrors (core/src/render/GpuTypes.hpp:40-87) so a
// packed scene can be compared with the reference packing rules field by
// field.  The per-frame constants (FrameConsts) start with the reference's
// GpuGlobals fields in the same order and append what the wavefront pipeline
// needs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mrt {

enum FrameFlags : uint32_t {
    kFlagHasEnv = 1u << 0,
    kFlagOrthographic = 1u << 2,
    kFlagPacketCamera = 1u << 3,  // camera rays traced by wave packets
};

enum LightKind : int32_t { kLightEnv = 0, kLightSun = 1, kLightPoint = 2, kLightRect = 3 };

enum TexFormat : uint32_t { kTexRgba8Srgb = 0, kTexRgba8Unorm = 1, kTexRgba32f = 2 };

struct FrameConsts {
    float4 camPos, camRight, camUp, camForward;
    float4 camFrustum;           // left, right, bottom, top
    uint32_t width, height, frameIndex, maxBounces;
    uint32_t lightCount, envWidth, envHeight, flags;
    float envRotation, envIntensity, envIntegral, exposure;
    float fireflyClamp;
    uint32_t tonemap, tlasRoot, instanceCount;
    uint32_t seedMix;            // xor-ed into the frame word of the RNG key (0 = reference keys)
    uint32_t guiding;            // 1 when continuation directions may come from the learned mixture
    float guideSelectProb;       // probability of drawing from the mixture at a guided vertex
    uint32_t guideTrain;         // 1 when shade writes training records this frame
    float guideMinRoughness;     // vertices with smaller roughness are not guided
    uint32_t guideStamp;         // nonzero, different for every rendered frame of an engine
    uint32_t guideTrainStride;   // training paths: p % guideTrainStride == guideTrainPhase
    uint32_t guideTrainPhase;
    float4 guideBoxMin;          // scene box: pos01 = (p - min) * scale
    float4 guideBoxScale;
    uint32_t tlasWideRoot;       // the TLAS's four-wide copy (SceneView::wide)
    // Environment strata (SceneView::envA / envB): the bright texels (each at
    // least 64 x the mean weight: a sun) and the rest.  envSplitA = the bright
    // stratum's share of the environment's weight, 0 when either is empty.
    float envSplitA;
    uint32_t envCountA, envCountB;
};

struct GpuInstance {             // 144 bytes
    float4 objectToWorld[4];     // column-major
    float4 worldToObject[4];     // column-major
    uint32_t blasRoot;
    uint32_t materialIndex;
    uint32_t pad0, pad1;
};
static_assert(sizeof(GpuInstance) == 144, "GpuInstance is 144 bytes");

// Compact traversal copy of an instance: rows of the 3x4 worldToObject and
// the BLAS root (64 bytes, one cache-friendly read per TLAS leaf).
struct TraceInstance {
    float4 row0, row1, row2;     // worldToObject rows [R | t]
    uint32_t blasRoot, materialIndex;
    uint32_t wideRoot;           // the BLAS's four-wide copy (SceneView::wide)
    uint32_t pad1;
};
static_assert(sizeof(TraceInstance) == 64, "TraceInstance is 64 bytes");

struct GpuMaterial {             // 80 bytes
    float4 baseColorOpacity;     // rgb + opacity
    float4 emissionRoughness;    // rgb emission + roughness
    float4 params;               // metallic, transmission, ior, pad
    int32_t baseColorTex, roughnessTex, metallicTex, normalTex;
    int32_t emissionTex, pad0, pad1, pad2;
};
static_assert(sizeof(GpuMaterial) == 80, "GpuMaterial is 80 bytes");

struct GpuLight {                // 80 bytes
    float4 p0;                   // sun: dir + cos(angular radius) | point: pos | rect: corner + area
    float4 p1;                   // sun: radiance | point: intensity | rect: edge0
    float4 p2;                   // rect: edge1
    float4 p3;                   // rect: radiance + twoSided
    int32_t type, pad0, pad1, pad2;
};
static_assert(sizeof(GpuLight) == 80, "GpuLight is 80 bytes");

struct GpuTriangle {             // global vertex indices + pad
    uint32_t i0, i1, i2, pad;
};

struct GpuTexture {              // one entry per texture slot
    uint32_t offset;             // texel offset into the pool of its format
    uint32_t width, height;
    uint32_t format;             // TexFormat; width == 0 -> empty slot
};

struct AliasEntry {              // environment alias table (Vose)
    float prob;                  // probability of keeping this texel
    uint32_t alias;              // texel taken otherwise
};

struct EnvStratumEntry {         // alias table over one stratum's texels
    float prob;                  // probability of keeping this entry
    uint32_t alias;              // entry taken otherwise
    uint32_t texel;              // the entry's texel
    uint32_t pad;
};

// Learned environment-stratum selection of guided shading: per cell (the
// shading point's position on a kNeeGrid^3 grid over the scene box, times the
// normal's dominant direction) the running sums of squared stratum-
// conditional NEE contributions and sample counts: (sumA, nA, sumB, nB).
constexpr uint32_t kNeeGrid = 32;
constexpr uint32_t kNeeCells = kNeeGrid * kNeeGrid * kNeeGrid * 6;

} // namespace mrt
