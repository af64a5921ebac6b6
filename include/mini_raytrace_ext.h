/*
 * mini_raytrace_ext.h -- additive extensions of the MiniRaytrace C ABI served
 * by this project's libmrt.so.  Nothing in mini_raytrace.h changes; callers
 * that only know that header keep working.
 *
 * Extensions:
 *   1. Ray queries over a standalone triangle set (closest hit / any hit) on
 *      device-resident ray batches.  These expose, as a C entry point, what
 *      the reference runs internally in core/shaders/intersect.glsl:48-178
 *      and restates on the CPU in tests/bvh_tests.cpp:49-73.
 *   2. Per-engine path options: RNG seed, path guiding on/off and its
 *      parameters (the reference has no counterpart; see each declaration).
 *   3. Batched rendering of several engines in one persistent launch.
 *   4. Per-frame GPU timing and counters for profiling.
 *
 * All entry points follow the conventions of mini_raytrace.h: structSize
 * first, mrtResult codes, no exceptions across the boundary, borrowed
 * pointers, one thread per engine / query scene.
 */
#ifndef MINI_RAYTRACE_EXT_H
#define MINI_RAYTRACE_EXT_H

#include "mini_raytrace.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- 1. ray queries ------------------------------------------------------- */

/* Host-side BVH build (no device needed): the reference BvhBuilder::build
 * (core/src/accel/Bvh.cpp:137-156) as a C entry point.  bounds: 6 floats per
 * primitive (lo xyz, hi xyz).  Writes up to nodeCap 32-byte nodes and `count`
 * primitive indices; *outNodeCount receives the node count (0 for count 0).
 * Returns MRT_ERROR_INVALID_ARGUMENT when nodeCap is too small. */
MRT_API mrtResult mrtxBuildBvh(const float* bounds, uint32_t count, uint32_t maxLeaf,
                               void* nodes, uint64_t nodeCap, uint32_t* primOrder,
                               uint64_t* outNodeCount);

/* The same build on the GPU: identical nodes and primitive order (the
 * reference builder's, node for node), inputs and outputs in device memory of
 * the current HIP device.  dBounds: 6 floats per primitive; nodeCap must be
 * at least 2 * count - 1 (the largest possible tree).  Work is enqueued on
 * hipStream (NULL = the null stream); the call waits for it between tree
 * levels and returns once the outputs are written.  A node that needs the
 * reference's median split (SAH preferring a leaf of more than 2 * maxLeaf
 * primitives, core/src/accel/Bvh.cpp:104-111) is built on the CPU instead,
 * with the same result (an Info log message says so).  count must be below
 * 2^31.  The library keeps one builder workspace per device until exit;
 * concurrent calls on one device take turns. */
MRT_API mrtResult mrtxBuildBvhDevice(const float* dBounds, uint32_t count, uint32_t maxLeaf,
                                     void* dNodes, uint64_t nodeCap, uint32_t* dPrimOrder,
                                     uint64_t* outNodeCount, void* hipStream);

typedef struct mrtxQueryScene mrtxQueryScene;

/* Ray record on the device: 8 floats, 32 bytes:
 *   origin.x origin.y origin.z tmin  dir.x dir.y dir.z tmax
 * A hit must satisfy tmin < t < tmax. */
typedef struct mrtxRay {
    float org[3];
    float tmin;
    float dir[3];
    float tmax;
} mrtxRay;

/* Closest-hit record on the device: 8 bytes.  triId is the index of the
 * triangle in the order it was passed to mrtxCreateQueryScene, or
 * MRTX_NO_HIT; t is the hit distance (the ray's tmax on a miss).  Exact ties
 * in t resolve to the lowest triId. */
typedef struct mrtxHit {
    uint32_t triId;
    float    t;
} mrtxHit;
#define MRTX_NO_HIT 0xFFFFFFFFu

typedef struct mrtxQuerySceneInfo {
    size_t   structSize;
    uint32_t triCount;
    uint32_t nodeCount;     /* 32-byte BVH nodes (reference Bvh.hpp:15-26 layout) */
    uint32_t leafCount;
    uint32_t maxDepth;      /* root = depth 0 */
    float    buildMs;       /* host wall time of the BVH build */
    uint64_t deviceBytes;   /* device memory held by the scene */
} mrtxQuerySceneInfo;

/* Traversal launch shape.  0 in a field = library default (or the value the
 * autotuner picked for this scene). */
typedef struct mrtxTraversalConfig {
    size_t   structSize;
    uint32_t blockSize;     /* threads per workgroup: 64, 128, 256 */
    uint32_t stackDepth;    /* LDS stack entries per lane: 8, 16, 32 */
    uint32_t loopMode;      /* 0 = one node or leaf step per iteration, 1 = while-while;
                               2, 3 = the same over a four-wide tree collapsed from the
                               BVH (same hits) */
} mrtxTraversalConfig;

/* Builds a binned-SAH BVH (reference core/src/accel/Bvh.cpp:137-156
 * semantics, max leaf 4) over triCount triangles and uploads it to HIP device
 * gpuIndex1-1 (0 = current device).  positions: xyz per vertex; indices: 3 per
 * triangle.  Host pointers, borrowed. */
MRT_API mrtResult mrtxCreateQueryScene(const float* positions, uint32_t vertexCount,
                                       const uint32_t* indices, uint32_t triCount,
                                       uint32_t gpuIndex1, mrtxQueryScene** outScene);
MRT_API void      mrtxDestroyQueryScene(mrtxQueryScene* scene);
MRT_API mrtResult mrtxGetQuerySceneInfo(const mrtxQueryScene* scene, mrtxQuerySceneInfo* out);
/* Copies the BVH nodes (32 bytes each, nodeCount of them) and the primitive
 * order (triCount uint32) to host memory.  Either pointer may be NULL. */
MRT_API mrtResult mrtxExportQueryBvh(const mrtxQueryScene* scene, void* nodes, size_t nodesBytes,
                                     uint32_t* primOrder, size_t primCount);
MRT_API mrtResult mrtxSetTraversalConfig(mrtxQueryScene* scene, const mrtxTraversalConfig* cfg);
MRT_API mrtResult mrtxGetTraversalConfig(const mrtxQueryScene* scene, mrtxTraversalConfig* out);

/* Closest hit for n device-resident rays; results to device memory.  Work is
 * enqueued on `hipStream` (a hipStream_t; NULL = the null stream) and the
 * call returns without synchronising. */
MRT_API mrtResult mrtxTraceClosest(mrtxQueryScene* scene, const mrtxRay* dRays, uint32_t n,
                                   mrtxHit* dHits, void* hipStream);
/* Any hit: dOccluded[i] = 1 if some triangle has tmin < t < tmax, else 0. */
MRT_API mrtResult mrtxTraceAny(mrtxQueryScene* scene, const mrtxRay* dRays, uint32_t n,
                               uint32_t* dOccluded, void* hipStream);
/* Synchronises the device and returns sticky status bits of all traces so
 * far: bit 0 = a traversal stack overflowed (results of that call invalid). */
MRT_API mrtResult mrtxGetQueryStatus(mrtxQueryScene* scene, uint32_t* flags);
/* Times every supported launch shape (block size x LDS stack depth) on the
 * given rays and keeps the fastest for this scene.  Overwrites dHits. */
MRT_API mrtResult mrtxAutotuneTraversal(mrtxQueryScene* scene, const mrtxRay* dRays, uint32_t n,
                                        mrtxHit* dHits, mrtxTraversalConfig* chosen);

/* ---- 2. per-engine path options ------------------------------------------- */

/* No counterpart in the reference: the reference keys its RNG by (pixel,
 * frame) only and always runs one launch shape. */
typedef struct mrtxPathOptions {
    size_t   structSize;
    uint32_t seed;          /* 0 = the reference's RNG keys; else mixed into the frame key */
    uint32_t useGraph;      /* 1 (default) = replay the frame as a captured HIP graph */
    uint32_t extendBlock;   /* shape of the per-bounce trace launches: closest hits of a */
    uint32_t extendStack;   /* bounce with the previous bounce's shadow rays (0 = default / tuned) */
    uint32_t shadowBlock;   /* shape of the last launch: the deepest bounce's shadow rays */
    uint32_t shadowStack;
    uint32_t tailPaths;     /* once a bounce level holds at most this many rays (all batched
                             * requests together), one kernel finishes those paths instead of a
                             * launch per level; 0 = default (1280 per compute unit),
                             * MRTX_TAIL_DISABLED = never.  Images do not depend on it. */
    uint32_t cameraPerLane; /* 0 (default) = each wave traces its 8x8 tile of camera rays together
                             * (shared node, triangle and instance reads); 1 = one traversal per
                             * lane.  Hits, and so images, do not depend on it. */
} mrtxPathOptions;

#define MRTX_TAIL_DISABLED 0xFFFFFFFFu

MRT_API mrtResult mrtxSetPathOptions(mrtEngine* engine, const mrtxPathOptions* opts);
MRT_API mrtResult mrtxGetPathOptions(mrtEngine* engine, mrtxPathOptions* out);

/* ---- 4. profiling and inspection ------------------------------------------- */

#define MRTX_MAX_BOUNCE_LEVELS 18

/* Stage order of gpuMs: generate, extend, shade, shadow, accumulate (with the
 * RGBA8 resolve, one kernel), tail, guide (prep + network inference), train
 * (record collection + Adam steps).  Each bounce's closest-hit launch also
 * traces the previous bounce's shadow rays, so "extend" holds those and
 * "shadow" only the deepest bounce's shadow rays; "tail" is the kernel that
 * finishes the last few paths (their shading, closest hits and shadow rays;
 * see mrtxPathOptions::tailPaths).  rays / shadowRays count every ray traced
 * at a level, by the per-level launches or the tail kernel. */
typedef struct mrtxFrameStats {
    size_t   structSize;
    float    frameMs;                                /* host wall time of the last frame */
    float    gpuMs[8];                               /* per stage, profiling mode only */
    uint32_t rays[MRTX_MAX_BOUNCE_LEVELS];           /* rays traced per bounce level */
    uint32_t shadowRays[MRTX_MAX_BOUNCE_LEVELS];     /* shadow rays per bounce level */
    uint32_t status;                                 /* bit 0: traversal stack overflow */
    /* Appended fields, written when structSize covers them: */
    float    commitMs;          /* host wall time of the last commit that changed the scene */
    uint64_t commitUploadBytes; /* bytes that commit copied to the device */
    uint32_t geometryRepacks;   /* commits that repacked every mesh, since creation */
    uint32_t meshesWritten;     /* meshes rewritten alone (in place, or appended
                                   behind the others), since creation */
    uint32_t graphTailLevel;    /* level at which the last frame's graph handed every
                                   path left to the tail kernel and ended (0xFFFFFFFF:
                                   per-level launches to the last bounce, or no graph) */
} mrtxFrameStats;

/* Profiling mode: each frame runs stage by stage with HIP events between the
 * stages (no graph replay) and records ray counts. */
MRT_API mrtResult mrtxSetProfiling(mrtEngine* engine, int32_t enable);
MRT_API mrtResult mrtxGetFrameStats(mrtEngine* engine, mrtxFrameStats* out);

/* Copy of the packed device scene (commits pending edits first).  Counts are
 * always filled; each array pointer may be NULL; arrays must hold the counts
 * reported by a previous call. */
typedef struct mrtxSceneExport {
    size_t   structSize;
    uint32_t nodeCount, triCount, vertexCount, instanceCount;
    uint32_t materialCount, lightCount, tlasRoot, blasNodeCount;
    uint32_t envWidth, envHeight;
    float    envIntegral;
    void*    nodes;        /* nodeCount x 32 B: all BLAS nodes, then the TLAS */
    void*    triangles;    /* triCount x 16 B: global vertex indices, leaf order */
    float*   positions;    /* vertexCount x 4 floats */
    float*   normals;      /* vertexCount x 4 floats */
    float*   uvs;          /* vertexCount x 2 floats */
    void*    instances;    /* instanceCount x 144 B (TLAS leaf order) */
    void*    materials;    /* materialCount x 80 B (index 0 = default) */
    void*    lights;       /* lightCount x 80 B (environment first when present) */
    void*    envAlias;     /* envWidth*envHeight x (float prob, uint32 alias) */
    /* Appended fields (a structSize that ends at envAlias is still accepted
     * and leaves them untouched): */
    float*   tangents;     /* vertexCount x 4 floats, w = handedness */
    void*    textureTable; /* MRTX_TEXTURE_SLOTS x 16 B: texel offset into its pool, width, height,
                              format (width 0 = empty slot) */
    uint32_t tex8Count;    /* texels in the RGBA8 pool (one uint32 each) */
    uint32_t tex32Count;   /* texels in the RGBA32F pool (4 floats each) */
    void*    tex8;
    float*   tex32;
} mrtxSceneExport;
#define MRTX_TEXTURE_SLOTS 64

MRT_API mrtResult mrtxExportScene(mrtEngine* engine, mrtxSceneExport* inout);

/* ---- 5. batched rendering -----------------------------------------------------
 * No counterpart in the reference ABI: many engines (render requests of any
 * resolution and scene) progress together in one HIP graph launch.  Every
 * engine ends in exactly the state that `frames` calls of mrtRenderFrame
 * would leave (same images, bit for bit), clamped to its spp limit.  All
 * engines must be on one device and appear once.
 * Mode 0 (up to 64 requests, plain or guided): one set of persistent kernels
 * per batched frame works on the queues of all requests at once (each wave
 * takes 64-item chunks of any request's queue); a guided request adds its own
 * network inference and training launches.  Mode 1 (more than 64 requests):
 * each request's frames are captured as a chain and the chains are
 * independent nodes of the graph. */
typedef struct mrtxBatchRequest {
    mrtEngine* engine;
    uint32_t   frames;        /* samples per pixel to add */
    uint32_t   framesDone;    /* out: samples added (after the spp-limit clamp) */
} mrtxBatchRequest;

typedef struct mrtxBatchStats {
    size_t   structSize;
    float    captureMs;       /* host time to capture and instantiate the graph */
    float    launchMs;        /* graph launch to completion */
    uint32_t requests;        /* requests with work */
    uint32_t frames;          /* frames rendered in total */
    uint32_t mode;            /* 0 = shared persistent kernels, 1 = per-request graph chains */
    uint32_t extendBlock, extendStack, shadowBlock, shadowStack;  /* mode 0 launch shapes */
} mrtxBatchStats;

MRT_API mrtResult mrtxRenderBatch(mrtxBatchRequest* requests, uint32_t count, mrtxBatchStats* stats);

/* Path-tracer launch-shape autotuner: one profiled trial frame per candidate
 * (block 64/128/256 x LDS stack 8/16/32); the shape with the least trace time
 * over the frame is kept for both trace launch kinds and reported.  Results
 * are cached per (triangle count, pixel count, depth, compute units) class.
 * Images do not depend on the shape; the trial frames are discarded
 * (accumulation restarts) and do not train the engine's guiding network. */
MRT_API mrtResult mrtxAutotunePath(mrtEngine* engine, mrtxPathOptions* chosen);

/* ---- 6. path guiding -------------------------------------------------------
 * No counterpart in the reference: a learned directional distribution (an
 * MLP producing a mixture of 8 von Mises-Fisher lobes from the position,
 * shading normal and outgoing direction of a path vertex), trained online
 * on the engine's own paths and sampled with one-sample MIS against the
 * BSDF.  Network: 78 inputs (12-band sin/cos encoding of the position in the
 * scene box, normal, wo) -> 4 x 64 ReLU -> 40 outputs (per lobe: mean
 * direction xyz, log kappa clamped to [-5, 7], weight logit); BF16 matrix
 * products with FP32 accumulation, FP32 master weights, Adam. */

#define MRTX_GUIDE_PARAM_COUNT 20136
#define MRTX_GUIDE_OUTPUTS 40

typedef struct mrtxGuidingOptions {
    size_t   structSize;
    uint32_t enableSampling;  /* draw continuation directions from the mixture (default 0) */
    uint32_t enableTraining;  /* train on this engine's paths every frame (default 0) */
    float    selectProb;      /* probability of a mixture draw at a guided vertex (0.5) */
    float    learningRate;    /* Adam step size (1e-3) */
    uint32_t batchSize;       /* training records per Adam step (65,536) */
    uint32_t stepsPerFrame;   /* Adam steps per rendered frame (4) */
    float    minRoughness;    /* less rough or transmissive vertices use the BSDF only (0.2) */
    uint32_t seed;            /* network initialisation seed (used when the network is created) */
    float    positionScale;   /* network position input = (p - scene box min) / largest box extent
                                 x positionScale, encoded as sin/cos(2^b pi x), b < 12 (0 = 1) */
    /* Appended fields, read and written when structSize covers them (0 =
     * default).  The first warmupFrames trained frames of a network take
     * warmupStepsPerFrame Adam steps each instead of stepsPerFrame (the
     * pool then holds batchSize x warmupStepsPerFrame records, split over
     * the steps): an untrained network learns from its first, smaller
     * pools in more, smaller steps.  warmupStepsPerFrame = stepsPerFrame
     * turns the warm-up off. */
    uint32_t warmupFrames;        /* (32) */
    uint32_t warmupStepsPerFrame; /* (16) */
} mrtxGuidingOptions;

typedef struct mrtxGuidingStats {
    size_t   structSize;
    uint32_t steps;           /* Adam steps taken */
    uint32_t lastBatch;       /* records in the last step's batch */
    float    lastLoss;        /* weighted negative log-likelihood of that batch */
    uint32_t poolRecords;     /* training records collected in the last frame */
    uint32_t candidates;      /* positive-weight path records found in the last frame */
    uint32_t lastFrameRows;   /* network queries (inferred rows) of the last profiled frame */
    /* Appended fields, written when structSize covers them.  Guided
     * environment NEE draws from two strata (bright texels, the rest) with a
     * bright-stratum probability learned per cell of the scene: */
    float    envBrightShare;  /* the bright stratum's share of the environment (0: no strata) */
    uint32_t envCellsLearned; /* cells whose probability has left the share (both strata sampled) */
    float    envMeanSplit;    /* mean learned bright-stratum probability over those cells */
} mrtxGuidingStats;

/* Changing the options re-captures the frame graph; it does not reset the
 * accumulation (all sampling strategies are unbiased). */
MRT_API mrtResult mrtxSetGuiding(mrtEngine* engine, const mrtxGuidingOptions* opts);
MRT_API mrtResult mrtxGetGuiding(mrtEngine* engine, mrtxGuidingOptions* out);
MRT_API mrtResult mrtxGetGuidingStats(mrtEngine* engine, mrtxGuidingStats* out);
/* count must be MRTX_GUIDE_PARAM_COUNT; setting resets the Adam state. */
MRT_API mrtResult mrtxGetGuideParams(mrtEngine* engine, float* dst, size_t count);
MRT_API mrtResult mrtxSetGuideParams(mrtEngine* engine, const float* src, size_t count);

/* Standalone network (same kernels as the engine), host buffers in and out.
 * Record layout, 16 floats per record:
 *   pos01.xyz, weight, n.xyz, -, wo.xyz, -, wi.xyz, -
 * (inference reads pos01, n and wo; training also weight and wi). */
typedef struct mrtxGuideNetwork mrtxGuideNetwork;

MRT_API mrtResult mrtxCreateGuideNetwork(int32_t gpuIndex1, uint32_t maxBatch, uint32_t seed,
                                         mrtxGuideNetwork** out);
MRT_API void      mrtxDestroyGuideNetwork(mrtxGuideNetwork* net);
MRT_API mrtResult mrtxGuideNetworkSetParams(mrtxGuideNetwork* net, const float* src, size_t count);
MRT_API mrtResult mrtxGuideNetworkGetParams(mrtxGuideNetwork* net, float* dst, size_t count);
/* Adam state: first and second moments (count floats each, may be NULL) and
 * the step counter. */
MRT_API mrtResult mrtxGuideNetworkGetAdam(mrtxGuideNetwork* net, float* m, float* v, size_t count,
                                          uint32_t* step);
/* raw: count x MRTX_GUIDE_OUTPUTS network outputs. */
MRT_API mrtResult mrtxGuideNetworkInfer(mrtxGuideNetwork* net, const float* records, uint32_t count,
                                        float* raw);
/* Gradient of loss = sum w (-log q(wi)) / sum w over the records, without an
 * update; grad may be NULL. */
MRT_API mrtResult mrtxGuideNetworkGradient(mrtxGuideNetwork* net, const float* records, uint32_t count,
                                           float* grad, size_t gradCount, float* loss);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) on the records' gradient. */
MRT_API mrtResult mrtxGuideNetworkTrain(mrtxGuideNetwork* net, const float* records, uint32_t count,
                                        float learningRate, float* loss);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (work under HIP graph capture, tight serving or online-training
 * loops): the step of mrtxGuideNetworkTrain on count device-resident records
 * (same 16-float layout), enqueued on hipStream (NULL = the network's own
 * stream; another stream is ordered against the network's other calls by the
 * caller).  It calls no hipMalloc, hipFree or other allocating routine: the
 * training workspace lives in one fixed device array per GPU, sized for
 * MRTX_GUIDE_NOALLOC_MAX_RECORDS records (about 12 MB, held until the process
 * ends); more records return MRT_ERROR_INVALID_ARGUMENT.  Every network on a
 * GPU shares that GPU's array, so calls that overlap on one GPU (other
 * streams or threads) race.  loss may be NULL; when given, the call waits for
 * the stream and stores the step's loss (leave it NULL under graph capture).
 * Same parameters, Adam state and loss, bit for bit, as mrtxGuideNetworkTrain
 * on the same records. */
#define MRTX_GUIDE_NOALLOC_MAX_RECORDS 8192u
MRT_API mrtResult mrtxGuideNetworkTrainNoAlloc(mrtxGuideNetwork* net, const float* dRecords, uint32_t count,
                                               float learningRate, void* hipStream, float* loss);
/* Mixture sampling from raw outputs: u holds 3 uniforms per row; out holds
 * (direction.xyz, mixture pdf) per row. */
MRT_API mrtResult mrtxGuideNetworkSample(mrtxGuideNetwork* net, const float* raw, const float* u,
                                         uint32_t count, float* out);
/* Mixture pdf of unit directions (4 floats per row, xyz used). */
MRT_API mrtResult mrtxGuideNetworkPdf(mrtxGuideNetwork* net, const float* raw, const float* dirs,
                                      uint32_t count, float* pdf);

#ifdef __cplusplus
} /* extern "C" */
#endif
#endif /* MINI_RAYTRACE_EXT_H */
