/*
 * mini_raytrace.h -- the MiniRaytrace C ABI, served by this project's
 * libmrt.so (HIP path tracer for gfx950).
 *
 * Every declaration below keeps the name, type, field order, enum value and
 * calling convention of the reference ABI (capi/include/mini_raytrace.h:26-237)
 * so that existing callers (the Rhino P/Invoke layer, the smoke test) bind to
 * this library unchanged.  Conventions the callers rely on:
 *   - opaque engine handle, mrtResult error codes; no C++ exception crosses
 *     this boundary;
 *   - every pointer argument is borrowed for the duration of the call only
 *     (the library deep-copies what it keeps);
 *   - calls on one engine come from one thread;
 *   - every descriptor struct starts with structSize; the library rejects a
 *     struct whose structSize is smaller than the one it was built with.
 *
 * Additive extensions (ray queries, batched rendering, path guiding) live in
 * mini_raytrace_ext.h; nothing here changes for them.
 */
#ifndef MINI_RAYTRACE_H
#define MINI_RAYTRACE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(_WIN32)
  #define MRT_API __declspec(dllexport)
#else
  #define MRT_API __attribute__((visibility("default")))
#endif

typedef struct mrtEngine mrtEngine;

/* MRT_ERROR_VULKAN_INIT keeps its value and name for binary compatibility;
 * this library returns it when the HIP device cannot be opened. */
typedef enum mrtResult {
    MRT_SUCCESS = 0,
    MRT_ERROR_UNKNOWN,
    MRT_ERROR_VULKAN_INIT,
    MRT_ERROR_OUT_OF_MEMORY,
    MRT_ERROR_INVALID_ARGUMENT,
    MRT_ERROR_INVALID_HANDLE,
    MRT_ERROR_SHADER_LOAD,
    MRT_ERROR_DEVICE_LOST,
} mrtResult;

/* Object ids come from one counter shared by every object kind, starting at 1. */
typedef uint32_t mrtMeshId;
typedef uint32_t mrtInstanceId;
typedef uint32_t mrtMaterialId;
typedef uint32_t mrtTextureId;
typedef uint32_t mrtLightId;
#define MRT_INVALID_ID 0u

/* Framebuffer readback formats.  Both are row-major with row 0 at the top of
 * the image; alpha is always 1.0. */
typedef enum mrtPixelFormat {
    MRT_PIXEL_RGBA8_SRGB = 0,   /* exposure + tonemap + sRGB OETF, 4 bytes/px */
    MRT_PIXEL_RGBA32F_LINEAR,   /* accumulated radiance / spp, 16 bytes/px   */
} mrtPixelFormat;

typedef enum mrtTextureFormat {
    MRT_TEX_RGBA8_SRGB = 0,
    MRT_TEX_RGBA8_UNORM,
    MRT_TEX_RGBA32F,
} mrtTextureFormat;

typedef enum mrtLightType {
    MRT_LIGHT_SUN = 0,
    MRT_LIGHT_POINT,
    MRT_LIGHT_RECT,
} mrtLightType;

typedef enum mrtTonemap {
    MRT_TONEMAP_LINEAR = 0,
    MRT_TONEMAP_ACES,
} mrtTonemap;

/* ---- descriptor structs ------------------------------------------------- */

typedef struct mrtEngineDesc {
    size_t   structSize;
    int32_t  enableValidation;  /* enables extra host-side argument checks */
    uint32_t width, height;     /* used only when both are > 0 */
    /* 0 = automatic device choice; N > 0 selects HIP device N-1.  An index
     * past the device count fails with MRT_ERROR_VULKAN_INIT and a message
     * containing "out of range". */
    uint32_t gpuIndex1;
} mrtEngineDesc;

typedef struct mrtMeshDesc {
    size_t          structSize;
    uint32_t        vertexCount;
    uint32_t        indexCount;     /* multiple of 3 */
    const float*    positions;      /* xyz per vertex, required */
    const float*    normals;        /* xyz per vertex, optional */
    const float*    uvs;            /* uv per vertex, optional */
    const uint32_t* indices;        /* required */
} mrtMeshDesc;

typedef struct mrtMaterialDesc {
    size_t  structSize;
    float   baseColor[3];
    float   opacity;
    float   roughness;
    float   metallic;
    float   transmission;
    float   ior;
    float   emission[3];
    mrtTextureId baseColorTex;      /* MRT_INVALID_ID = use the constant */
    mrtTextureId roughnessTex;
    mrtTextureId metallicTex;
    mrtTextureId normalTex;
    mrtTextureId emissionTex;
} mrtMaterialDesc;

typedef struct mrtTextureDesc {
    size_t           structSize;
    uint32_t         width, height;
    mrtTextureFormat format;
    const void*      pixels;        /* tightly packed RGBA rows */
    int32_t          generateMips;
} mrtTextureDesc;

typedef struct mrtLightDesc {
    size_t       structSize;
    mrtLightType type;
    float direction[3];             /* sun: unit vector towards the light */
    float position[3];              /* point light position */
    float radiance[3];              /* sun/rect radiance; point intensity (W/sr) */
    float angularRadius;            /* sun cone half-angle, radians */
    float corner[3], edge0[3], edge1[3]; /* rect light parallelogram */
    int32_t twoSided;               /* rect light emits from both faces */
} mrtLightDesc;

typedef struct mrtCameraDesc {
    size_t structSize;
    float position[3];
    float target[3];
    float up[3];
    float fovYDeg;
} mrtCameraDesc;

typedef enum mrtCameraProjection {
    MRT_CAMERA_PERSPECTIVE = 0,
    MRT_CAMERA_ORTHOGRAPHIC = 1,
} mrtCameraProjection;

/* Explicit frustum camera.  Perspective: left/right/bottom/top describe the
 * image rectangle on the plane at distance 1 along `forward`.  Orthographic:
 * they are absolute offsets of the ray origins from `position`. */
typedef struct mrtCameraDescEx {
    size_t  structSize;
    int32_t projection;      /* mrtCameraProjection */
    float   position[3];
    float   forward[3];      /* unit */
    float   up[3];           /* up hint; orthonormalised against forward */
    float   left, right, bottom, top;
} mrtCameraDescEx;

typedef struct mrtRenderSettings {
    size_t     structSize;
    uint32_t   width, height;
    uint32_t   maxBounces;          /* default 8, clamped to 16 */
    uint32_t   sppLimit;            /* default 1024 */
    float      exposureEV;          /* default 0 */
    float      fireflyClamp;        /* per-sample max-component clamp; 0 = off, default 50 */
    mrtTonemap tonemap;
    int32_t    denoise;             /* accepted; this build has no denoiser */
} mrtRenderSettings;

typedef struct mrtFrameInfo {
    size_t   structSize;
    uint32_t spp;
    int32_t  converged;
    int32_t  denoised;
    float    frameMs;
} mrtFrameInfo;

typedef struct mrtCommitStats {
    size_t   structSize;
    uint32_t blasRebuilt;   /* meshes whose BVH was rebuilt by this commit */
    int32_t  tlasRebuilt;   /* 1 when the instance BVH was rebuilt */
    float    commitMs;      /* host wall time of the commit */
} mrtCommitStats;

/* ---- lifecycle ----------------------------------------------------------- */
MRT_API mrtResult mrtCreateEngine(const mrtEngineDesc* desc, mrtEngine** outEngine);
MRT_API void      mrtDestroyEngine(mrtEngine* engine);

/* ---- scene: incremental edits -------------------------------------------- */
MRT_API mrtMeshId     mrtAddMesh(mrtEngine*, const mrtMeshDesc*);
MRT_API void          mrtUpdateMesh(mrtEngine*, mrtMeshId, const mrtMeshDesc*);
MRT_API void          mrtRemoveMesh(mrtEngine*, mrtMeshId);

/* xform3x4: row-major 3x4 matrix [R|t]; NULL = identity */
MRT_API mrtInstanceId mrtAddInstance(mrtEngine*, mrtMeshId, const float* xform3x4, mrtMaterialId);
MRT_API void          mrtSetInstanceTransform(mrtEngine*, mrtInstanceId, const float* xform3x4);
MRT_API void          mrtRemoveInstance(mrtEngine*, mrtInstanceId);

MRT_API mrtMaterialId mrtAddMaterial(mrtEngine*, const mrtMaterialDesc*);
MRT_API void          mrtUpdateMaterial(mrtEngine*, mrtMaterialId, const mrtMaterialDesc*);
/* Instances that reference a removed material render with the default one. */
MRT_API void          mrtRemoveMaterial(mrtEngine*, mrtMaterialId);
MRT_API mrtTextureId  mrtAddTexture(mrtEngine*, const mrtTextureDesc*);
/* Material slots that reference a removed texture fall back to their
 * constant value; the texture slot is reused by a later mrtAddTexture. */
MRT_API void          mrtRemoveTexture(mrtEngine*, mrtTextureId);

MRT_API mrtLightId    mrtAddLight(mrtEngine*, const mrtLightDesc*);
MRT_API void          mrtUpdateLight(mrtEngine*, mrtLightId, const mrtLightDesc*);
MRT_API void          mrtRemoveLight(mrtEngine*, mrtLightId);

/* hdri must be an MRT_TEX_RGBA32F equirectangular map (other formats are
 * ignored); NULL removes the environment. */
MRT_API void          mrtSetEnvironment(mrtEngine*, const mrtTextureDesc* hdri,
                                        float rotationRad, float intensity);
MRT_API void          mrtSetCamera(mrtEngine*, const mrtCameraDesc*);
MRT_API void          mrtSetCameraEx(mrtEngine*, const mrtCameraDescEx*);

/* Apply pending edits now (BVH builds, uploads, accumulation reset).  A call
 * with nothing pending is a no-op that reports zero rebuilds.  outStats may
 * be NULL. */
MRT_API mrtResult     mrtCommitScene(mrtEngine*, mrtCommitStats* outStats);

/* ---- render ---------------------------------------------------------------- */
MRT_API mrtResult mrtSetRenderSettings(mrtEngine*, const mrtRenderSettings*);
/* Blocking.  Adds one sample per pixel until sppLimit is reached. */
MRT_API mrtResult mrtRenderFrame(mrtEngine*, mrtFrameInfo* outInfo);
/* dst must hold width*height*4 (RGBA8) or width*height*16 (RGBA32F) bytes.
 * Internal staging memory is reused across calls. */
MRT_API mrtResult mrtReadFramebuffer(mrtEngine*, mrtPixelFormat, void* dst, size_t dstSize);
MRT_API void      mrtResetAccumulation(mrtEngine*);

/* ---- misc -------------------------------------------------------------------- */
typedef void (*mrtLogFn)(int level /*0=debug..3=error*/, const char* message, void* user);
MRT_API void        mrtSetLogCallback(mrtLogFn fn, void* user);
MRT_API const char* mrtResultToString(mrtResult r);
/* Detail text of the most recent failure on the calling thread; empty string
 * when nothing has failed.  Never NULL. */
MRT_API const char* mrtGetLastErrorMessage(void);

#ifdef __cplusplus
} /* extern "C" */
#endif
#endif /* MINI_RAYTRACE_H */
