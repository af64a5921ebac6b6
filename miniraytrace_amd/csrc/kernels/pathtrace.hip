// Wavefront progressive path tracer for gfx950.
//
// The reference runs one megakernel thread per pixel that loops over bounces
// (core/shaders/pathtrace.comp:100-234).  Here each bounce is split into
// stages that communicate through compacted queues in HBM:
//
//   generate   camera rays for every pixel -> ray queue 0
//   extend(b)  closest hit for ray queue b -> hit records (per queue slot)
//   shade(b)   surface, emission, NEE light sample, BSDF sample, Russian
//              roulette -> shadow queue b and ray queue b+1
//   shadow(b)  any hit for shadow queue b; unoccluded -> add contribution
//   accumulate per-pixel firefly clamp, NaN guard, accum += (L, 1)
//   resolve    accum -> RGBA8 (exposure, ACES, sRGB) / RGBA32F (accum/spp)
//
// Queues are appended with one atomic per wave (ballot + mbcnt prefix) and
// consumed by persistent waves that fetch 64 slots per atomic.  Per-path
// random numbers follow the reference's draw order exactly (pcg4d keyed by
// pixel, frame and a per-path counter), so a path consumes the same random
// numbers as the reference megakernel would.
//
// No floating-point contraction anywhere in this file (the pragma precedes
// the includes, so it covers the inlined device helpers too): a path's
// arithmetic is then the same whichever kernel shades it (per-level, tail,
// batched or alone), which keeps images independent of how work is split.
#pragma clang fp contract(off)
#include "guide_device.h"
#include "pt_device.h"
#include "pt_launch.h"
#include "trace_scene.h"

namespace mrt {
namespace dev {

struct Surface {
    V3 pos, nGeom, nShade;
    float u, v;
    bool front;
    GpuMaterial mat;
};

__device__ __forceinline__ V3 colDot3(const float4* cols, V3 n) {
    // transpose(mat3(M)) * n: component i = dot(column i, n)
    return v3(cols[0].x * n.x + cols[0].y * n.y + cols[0].z * n.z,
              cols[1].x * n.x + cols[1].y * n.y + cols[1].z * n.z,
              cols[2].x * n.x + cols[2].y * n.y + cols[2].z * n.z);
}
__device__ __forceinline__ V3 mulCols3(const float4* cols, V3 t) {
    // mat3(M) * t
    return v3(cols[0]) * t.x + v3(cols[1]) * t.y + v3(cols[2]) * t.z;
}

// pathtrace.comp:39-80
__device__ Surface makeSurface(const SceneView& sc, V3 dir, float b1, float b2, uint32_t triIdx,
                               uint32_t instIdx) {
    Surface s;
    const GpuInstance& inst = sc.instances[instIdx];
    const GpuTriangle tri = sc.tris[triIdx];
    const float b0 = 1.0f - b1 - b2;
    const V3 p0 = v3(sc.positions[tri.i0]), p1 = v3(sc.positions[tri.i1]), p2 = v3(sc.positions[tri.i2]);
    const V3 po = p0 * b0 + p1 * b1 + p2 * b2;
    s.pos = mulCols3(inst.objectToWorld, po) + v3(inst.objectToWorld[3]);
    V3 nGeom = normalize(colDot3(inst.worldToObject, cross(p1 - p0, p2 - p0)));
    const V3 no = v3(sc.normals[tri.i0]) * b0 + v3(sc.normals[tri.i1]) * b1 + v3(sc.normals[tri.i2]) * b2;
    V3 nShade = normalize(colDot3(inst.worldToObject, no));
    s.front = dot(dir, nGeom) < 0.0f;
    if (!s.front) {
        nGeom = -nGeom;
        nShade = -nShade;
    }
    s.mat = sc.materials[inst.materialIndex];
    // Texture coordinates feed only texture lookups: fetched when the
    // material has a texture (a texture slot >= 0 clears the AND's sign bit).
    s.u = 0.0f;
    s.v = 0.0f;
    if ((s.mat.baseColorTex & s.mat.roughnessTex & s.mat.metallicTex & s.mat.normalTex & s.mat.emissionTex) >= 0) {
        const float2 uv0 = sc.uvs[tri.i0], uv1 = sc.uvs[tri.i1], uv2 = sc.uvs[tri.i2];
        s.u = uv0.x * b0 + uv1.x * b1 + uv2.x * b2;
        s.v = uv0.y * b0 + uv1.y * b1 + uv2.y * b2;
    }
    if (s.mat.normalTex >= 0) {
        const float4 t0 = sc.tangents[tri.i0], t1 = sc.tangents[tri.i1], t2 = sc.tangents[tri.i2];
        const float4 tanH = make_float4(t0.x * b0 + t1.x * b1 + t2.x * b2, t0.y * b0 + t1.y * b1 + t2.y * b2,
                                        t0.z * b0 + t1.z * b1 + t2.z * b2, t0.w * b0 + t1.w * b1 + t2.w * b2);
        const V3 tangent = normalize(mulCols3(inst.objectToWorld, v3(tanH)));
        const float4 tn4 = sampleTexture(sc, s.mat.normalTex, s.u, s.v);
        const V3 tn = v3(tn4.x * 2.0f - 1.0f, tn4.y * 2.0f - 1.0f, tn4.z * 2.0f - 1.0f);
        const V3 bitan = cross(nShade, tangent) * tanH.w;
        nShade = normalize(tangent * tn.x + bitan * tn.y + nShade * tn.z);
        if (dot(nShade, nGeom) < 0.0f) nShade = nGeom;
    }
    s.nGeom = nGeom;
    s.nShade = nShade;
    return s;
}

// pathtrace.comp:82-94
__device__ __forceinline__ Bsdf makeBsdf(const SceneView& sc, const Surface& s) {
    Bsdf b;
    b.base = v3(s.mat.baseColorOpacity);
    if (s.mat.baseColorTex >= 0) b.base = b.base * v3(sampleTexture(sc, s.mat.baseColorTex, s.u, s.v));
    b.roughness = s.mat.emissionRoughness.w;
    if (s.mat.roughnessTex >= 0) b.roughness *= sampleTexture(sc, s.mat.roughnessTex, s.u, s.v).x;
    b.metallic = s.mat.params.x;
    if (s.mat.metallicTex >= 0) b.metallic *= sampleTexture(sc, s.mat.metallicTex, s.u, s.v).x;
    b.transmission = s.mat.params.y;
    b.ior = s.mat.params.z;
    b.entering = s.front;
    return b;
}

// ---------------------------------------------------------------- generate --
// Camera ray of queue slot k.  Slots walk the image in bands of 8 rows,
// column by column inside a band, so 64 consecutive slots (one wave) cover
// an 8 x 8 pixel tile: the wave's rays stay close through the BVH.  A
// bijection of [0, width * height) (the last band may be shorter).
__device__ __forceinline__ void generatePixel(const PathView& pv, const FrameConsts& g, uint32_t k) {
    const uint32_t band = k / (8u * g.width);
    const uint32_t rows = min(8u, g.height - 8u * band);
    const uint32_t j = k - band * 8u * g.width;
    const uint32_t px = j / rows, py = 8u * band + j % rows;
    const uint32_t p = py * g.width + px;
    Rng rng{ px, py, g.frameIndex ^ g.seedMix, 0u };
    float jx, jy;
    rng.next2(jx, jy);
    const float ux = (float(px) + jx) / float(g.width);
    const float uy = (float(py) + jy) / float(g.height);
    const float sx = g.camFrustum.x * (1.0f - ux) + g.camFrustum.y * ux;
    const float sy = g.camFrustum.w * (1.0f - uy) + g.camFrustum.z * uy;
    V3 o, d;
    if (g.flags & kFlagOrthographic) {
        o = v3(g.camPos) + v3(g.camRight) * sx + v3(g.camUp) * sy;
        d = v3(g.camForward);
    } else {
        o = v3(g.camPos);
        d = normalize(v3(g.camForward) + v3(g.camRight) * sx + v3(g.camUp) * sy);
    }
    pv.rayO[0][k] = f4(o, __uint_as_float(p));
    pv.rayD[0][k] = f4(d, 0.0f);
    pv.beta[p] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    pv.radiance[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    pv.rngCounter[p] = rng.counter;
    if (pv.rrScale) pv.rrScale[p] = 1.0f;
    if (pv.recMask && p % g.guideTrainStride == g.guideTrainPhase) pv.recMask[p / g.guideTrainStride] = 0u;
}

// ------------------------------------------------------------ batch helpers --
// Pixel slot -> request: segments start at multiples of 64, so a wave's 64
// slots belong to one request (the last one whose pixelBase <= slot).
__device__ __forceinline__ uint32_t batchRequestOfSlot(const BatchView& bv, uint32_t waveFirstSlot) {
    const uint32_t j = laneId();
    const bool le = j < bv.count && bv.entries[j].pixelBase <= waveFirstSlot;
    const uint64_t m = __ballot(le);
    return m ? 63u - uint32_t(__builtin_clzll(m)) : 0u;
}

// Work split of one queue segment of a stage over all requests: lane j holds
// request j's item count in the segment and the inclusive prefix of 64-item
// chunks.
struct BatchChunks {
    uint32_t n;      // lane j: items of request j in this split
    uint32_t incl;   // lane j: chunks of requests 0..j
    uint32_t total;
};

// Lane j: request j's counter array (null when request j is absent or does
// not render this frame).
__device__ __forceinline__ const uint32_t* laneCounters(const BatchView& bv) {
    const uint32_t j = laneId();
    return j < bv.count && bv.entries[j].active ? bv.entries[j].pv.counters : nullptr;
}

// Lane j: request j's items of (level, slot) in each segment.
__device__ __forceinline__ void laneSegmentCounts(const uint32_t* cnt, uint32_t level, uint32_t slot,
                                                  uint32_t (&n)[kQueueSegs]) {
#pragma unroll
    for (uint32_t s = 0; s < kQueueSegs; ++s) n[s] = cnt ? cnt[cntWord(level, slot, s)] : 0u;
}

// Lane s (< kQueueSegs): the 64-item chunks of segment s over the requests
// whose lane has `keep` set (lane j: request j's counts n).
__device__ __forceinline__ uint32_t segmentTotals(const uint32_t (&n)[kQueueSegs], bool keep) {
    uint32_t out = 0;
#pragma unroll
    for (uint32_t s = 0; s < kQueueSegs; ++s) {
        uint32_t c = keep ? (n[s] + 63u) / 64u : 0u;
        for (int off = 32; off >= 1; off >>= 1) c += __shfl_xor(c, off, 64);
        if (laneId() == s) out = c;
    }
    return out;
}

// Lane j: whether the tail kernel serves request j at a level where it holds
// n rays: all requests together hold at most tailPaths rays there and
// request j holds at most its own limit (BatchEntry::tailAt; guided requests
// hand over only their last few paths).  The per-level kernels serve the rest.
__device__ __forceinline__ bool batchTailLane(const BatchView& bv, const uint32_t (&n)[kQueueSegs],
                                              uint32_t tailPaths) {
    uint32_t mine = 0;
#pragma unroll
    for (uint32_t s = 0; s < kQueueSegs; ++s) mine += n[s];
    const uint32_t tailAt = mine > 0 ? bv.entries[laneId()].tailAt : 0u;
    uint32_t all = mine;
    for (int off = 32; off >= 1; off >>= 1) all += __shfl_xor(all, off, 64);
    return mine > 0 && all <= tailPaths && mine <= tailAt;
}

// Which part of a level's queues a kernel works on (tail: batchTailLane).
enum BatchPart : int { kPartAll = 0, kPartLevel = 1, kPartTail = 2 };

__device__ __forceinline__ BatchChunks batchChunks(const uint32_t* cnt, uint32_t bounce, uint32_t slot, uint32_t seg,
                                                   bool tail = false, int part = kPartAll) {
    const uint32_t j = laneId();
    BatchChunks bc;
    uint32_t n = cnt ? cnt[cntWord(bounce, slot, seg)] : 0u;
    if ((part == kPartLevel && tail) || (part == kPartTail && !tail)) n = 0;
    bc.n = n;
    uint32_t c = (n + 63u) / 64u;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(c, off, 64);
        if (j >= uint32_t(off)) c += t;
    }
    bc.incl = c;
    bc.total = __shfl(c, 63, 64);
    return bc;
}

// Request r (wave-uniform), its queue length and the first item of chunk
// `chunk` of a stage's work split.
__device__ __forceinline__ void batchLocate(const BatchChunks& bc, uint32_t chunk, uint32_t& r, uint32_t& count,
                                            uint32_t& base) {
    const uint64_t m = __ballot(bc.incl > chunk);
    r = __builtin_amdgcn_readfirstlane(uint32_t(__builtin_ctzll(m)));
    const uint32_t prev = __shfl(bc.incl, r == 0 ? 0 : r - 1, 64);
    base = (chunk - (r == 0 ? 0u : prev)) * 64u;
    count = __builtin_amdgcn_readfirstlane(__shfl(bc.n, r, 64));
}

// Next chunk number from a stage's work cursor (wave-uniform).
__device__ __forceinline__ uint32_t batchFetch(uint32_t* cursor) {
    uint32_t chunk = 0;
    if (laneId() == 0) chunk = atomicAdd(cursor, 1u);
    return __builtin_amdgcn_readfirstlane(chunk);
}

// Work loop of a stage over its queue segments (cursors: the stage's segment
// cursors, tot: lane s holds the chunks of segment s).  A wave takes chunks
// of its XCD's segment, then of the others in turn, one atomic per chunk;
// a segment without chunks costs nothing and a spent one a single atomic.
// prepare(seg) loads the work split of segment seg once the wave holds a
// chunk of it; work(seg, chunk) runs the chunk.
template <class Prepare, class Work>
__device__ __forceinline__ void segmentLoop(uint32_t* cursors, uint32_t tot, Prepare&& prepare, Work&& work) {
    const uint32_t home = xccId();
#pragma unroll 1
    for (uint32_t t = 0; t < kQueueSegs; ++t) {
        const uint32_t seg = (home + t) % kQueueSegs;
        const uint32_t total = __builtin_amdgcn_readlane(tot, seg);
        if (total == 0) continue;
        bool ready = false;
        while (true) {
            const uint32_t chunk = batchFetch(cursors + seg * kLineWords);
            if (chunk >= total) break;
            if (!ready) {
                prepare(seg);
                ready = true;
            }
            work(seg, chunk);
        }
    }
}

// Per frame: every request's counters and the batch's work cursors to zero
// (the first word of each counter and cursor line; the rest is padding).
__global__ __launch_bounds__(256) void kClearBatch(BatchView bv) {
    const uint32_t lines = bv.count * kCounterLines + kCursorLines;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < lines; i += gridDim.x * blockDim.x) {
        if (i < bv.count * kCounterLines)
            bv.entries[i / kCounterLines].pv.counters[(i % kCounterLines) * kLineWords] = 0u;
        else
            bv.cursors[(i - bv.count * kCounterLines) * kLineWords] = 0u;
    }
}

// Camera ray of pixel slot p goes to queue slot p: segment p / segCap.
__global__ __launch_bounds__(256) void kGenerateBatch(BatchView bv) {
    for (uint32_t w = (blockIdx.x * blockDim.x + threadIdx.x) & ~63u; w < bv.pixelSlots; w += gridDim.x * blockDim.x) {
        const uint32_t r = batchRequestOfSlot(bv, w);
        const BatchEntry& e = bv.entries[r];
        if (!e.active) continue;
        const uint32_t p = w - e.pixelBase + laneId();
        if (p < kQueueSegs) {
            const uint32_t lo = min(e.pixels, p * e.pv.segCap);
            e.pv.counters[cntWord(0, kCntRays, p)] = min(e.pv.segCap, e.pixels - lo);
        }
        if (p < e.pixels) generatePixel(e.pv, *e.g, p);
    }
}

// ------------------------------------------------------------------ extend --
// Closest hits for items [base, base + 64) of segment `seg` of one request's
// ray queue (count items in the segment).  PREP && prep (a guided request at
// a level the tail kernel does not take): the same wave then forms the
// guiding network's query record of every hit on a surface that may be
// guided (opaque, roughness >= guideMinRoughness) -- (pos01, n, wo) at a
// compacted row -- and the row number in guideIdx (~0u elsewhere), while the
// hit is still in registers.
template <int BLOCK, int S, bool PREP>
__device__ __forceinline__ void extendChunk(const SceneView& sc, const PathView& pv, const FrameConsts* __restrict__ gp,
                                            uint32_t bounce, uint32_t seg, uint32_t count, uint32_t base, uint2* lds,
                                            bool prep, uint32_t* waveStack) {
    const uint32_t i = base + laneId();
    const uint32_t k = seg * pv.segCap + i;  // queue slot
    bool overflow = false;
    bool hit = false;
    SceneHit h;
    V3 dir = v3(0.0f, 0.0f, 0.0f);
    if (i < count) {
        const float4 o = pv.rayO[bounce & 1][k], d = pv.rayD[bounce & 1][k];
        dir = v3(d);
        // Camera rays (level 0) come in 8 x 8 pixel tiles per wave: the wave
        // traverses together (same hits as the per-lane traversal).
        hit = bounce == 0 && (gp->flags & kFlagPacketCamera) ? traceScenePacket(sc, gp->tlasRoot, gp->instanceCount, make_float3(o.x, o.y, o.z),
                                             make_float3(d.x, d.y, d.z), kEpsRay, 1e30f, waveStack, h, overflow)
                          : traceSceneWide<false, S, BLOCK>(sc, gp->tlasWideRoot, gp->instanceCount,
                                                        make_float3(o.x, o.y, o.z), make_float3(d.x, d.y, d.z),
                                                        kEpsRay, 1e30f, lds, h, overflow);
        pv.hit[k] = hit ? make_float4(h.t, h.b1, h.b2, __uint_as_float(h.tri))
                        : make_float4(1e30f, 0.0f, 0.0f, __uint_as_float(kNoHit));
        pv.hitInst[k] = hit ? h.inst : kNoHit;
    }
    if (overflow) atomicOr(pv.counters + kStatusWord, 1u);
    if (PREP && prep) {  // wave-uniform
        bool guide = false;
        V3 pos01, n;
        if (hit) {
            const Surface s = makeSurface(sc, dir, h.b1, h.b2, h.tri, h.inst);
            const Bsdf b = makeBsdf(sc, s);
            guide = b.transmission <= 0.0f && b.roughness >= gp->guideMinRoughness;
            pos01 = (s.pos - v3(gp->guideBoxMin)) * v3(gp->guideBoxScale);
            n = s.nShade;
        }
        const uint32_t row = waveAppend(pv.counters + cntWord(bounce, kCntGuideRecords, 0), guide);
        if (i < count) pv.guideIdx[k] = guide ? row : kNoHit;
        if (guide) {
            float4* r = pv.guideRec + size_t(row) * 3;
            r[0] = f4(pos01, 0.0f);
            r[1] = f4(n, 0.0f);
            r[2] = f4(-dir, 0.0f);
        }
    }
}

// ------------------------------------------------------------------- shade --
// Outputs of shading one path vertex: the continuation ray and the shadow
// ray (NEE), each when emitted.
struct ShadeOut {
    bool emitRay, emitShadow;
    V3 newO, newD, shO, shD, shC;
    float shMax;
    uint2 shX;  // stratum-statistics record of the shadow ray (PathView::shadowX)
};

// Shades path p at the hit `h` (t, b1, b2, triangle bits; instance hitInst)
// of its bounce-`bounce` ray with direction `dir`: reads and updates the
// path's throughput, radiance and random-number counter.  guideRow: the
// vertex's row of decoded guiding mixtures (GUIDE), else kNoHit.
//
// GUIDE: the prep pass ran for this bounce.  At a guided vertex with
// FrameConsts::guiding set, the continuation direction comes from the learned
// mixture with probability guideSelectProb and from the BSDF otherwise
// (one-sample MIS: the path weight uses the combined pdf, which NEE's MIS
// weight uses too).  With guideTrain set, every guided vertex that continues
// writes a training record.  Without GUIDE (or with neither flag) this
// performs exactly the reference's operations.
template <bool GUIDE>
__device__ __forceinline__ ShadeOut shadePath(const SceneView& sc, const PathView& pv, const FrameConsts& g,
                                              uint32_t bounce, uint32_t p, V3 dir, float4 h, uint32_t hitInst,
                                              uint32_t guideRow) {
    const bool last = bounce >= g.maxBounces;
    ShadeOut o;
    o.emitRay = false;
    o.emitShadow = false;
    o.shMax = 0.0f;
    o.shX = make_uint2(~0u, 0u);
    const uint32_t tri = __float_as_uint(h.w);
    float4 beta4 = pv.beta[p];
    V3 beta = v3(beta4);
    float prevPdf = beta4.w;
    float4 L4 = pv.radiance[p];
    V3 L = v3(L4);
    Rng rng{ p % g.width, p / g.width, g.frameIndex ^ g.seedMix, pv.rngCounter[p] };
    if (tri == kNoHit) {
        // pathtrace.comp:135-147: environment, MIS against NEE after a BSDF bounce.
        const V3 env = envRadiance(sc, g, dir);
        if (maxComp(env) > 0.0f) {
            float w = 1.0f;
            if (bounce > 0) {
                // NEE's pdf of the plain light distribution (stratified
                // guided NEE weighs MIS with it too)
                const float pl = envPdf(sc, g, dir) * envPickPdf(g);
                w = balanceHeuristic(prevPdf, pl);
                if (pv.pathRec && g.guideTrain && p % g.guideTrainStride == g.guideTrainPhase) {
                    // Training target of the previous vertex: its own
                    // continuation's environment hit counts with the MIS
                    // weight of the BSDF pdf (L4.w), not of the guided
                    // mixture, so light that NEE already samples well
                    // (the sun) does not pull the mixture towards itself.
                    const float wTrain = balanceHeuristic(L4.w, pl);
                    const uint32_t ts = g.guideTrainStride;
                    float4* r = pv.pathRec + (size_t(bounce - 1) * ((pv.capacity + ts - 1) / ts) + p / ts) * 4;
                    if (wTrain != w && __float_as_uint(r[2].w) == g.guideStamp)
                        r[1].w -= luminance(beta * env) * (wTrain - w);
                }
            }
            L = L + beta * env * w;
        }
    } else {
        const Surface s = makeSurface(sc, dir, h.y, h.z, tri, hitInst);
        const float opacity = s.mat.baseColorOpacity.w;
        if (opacity < 1.0f && rng.next1() >= opacity) {
            // Stochastic alpha: pass through, consuming one bounce.
            o.newO = s.pos + dir * kEpsRay;
            o.newD = dir;
            o.emitRay = !last;
        } else {
            const V3 emission = s.mat.emissionTex >= 0
                                    ? v3(s.mat.emissionRoughness) * v3(sampleTexture(sc, s.mat.emissionTex, s.u, s.v))
                                    : v3(s.mat.emissionRoughness);
            if (maxComp(emission) > 0.0f && s.front) L = L + beta * emission;
            const Bsdf bsdf = makeBsdf(sc, s);
            V3 T, B;
            onb(s.nShade, T, B);
            V3 wo = v3(dot(-dir, T), dot(-dir, B), dot(-dir, s.nShade));
            if (wo.z < 1e-6f) wo.z = 1e-6f;
            wo = normalize(wo);
            const float4* mix = nullptr;  // learned mixture of this vertex
            float sp = 0.0f;
            bool record = false;
            if (GUIDE) {
                const uint32_t gi = guideRow;
                if (gi != kNoHit) {
                    record = g.guideTrain != 0 && p % g.guideTrainStride == g.guideTrainPhase;
                    if (g.guiding) {
                        mix = pv.guideMix + size_t(gi) * kGuideMixF4;
                        sp = g.guideSelectProb;
                    }
                }
            }
            // NEE (pathtrace.comp:181-197).  Guided shading with stratum
            // statistics draws the environment over its strata.
            const bool strata = GUIDE && pv.neeStats && g.envSplitA > 0.0f;
            uint32_t cell = 0;
            float split = g.envSplitA;
            if (strata) {
                cell = neeCell(g, s.pos, s.nShade);
                split = neeSelect(pv.neeStats, cell, g.envSplitA);
            }
            const LightSample ls = strata ? sampleOneLight<true>(sc, g, s.pos, rng, split)
                                          : sampleOneLight<false>(sc, g, s.pos, rng);
            if (ls.valid && ls.misPdf > 1e-9f) {
                const V3 wi = v3(dot(ls.wi, T), dot(ls.wi, B), dot(ls.wi, s.nShade));
                V3 f;
                float bsdfPdf;
                bsdfEval(bsdf, wo, wi, f, bsdfPdf);
                if (maxComp(f) > 0.0f) {
                    // MIS against the continuation's pdf.  At the last bounce
                    // no continuation is drawn and the reference weights NEE
                    // against the BSDF pdf alone; guided shading does the same,
                    // so both tracers converge to the same image.
                    float contPdf = bsdfPdf;
                    if (GUIDE && mix && !last) contPdf = sp * storedMixPdf(mix, ls.wi) + (1.0f - sp) * bsdfPdf;
                    const float w = ls.delta ? 1.0f : balanceHeuristic(ls.misPdf, contPdf);
                    o.shO = s.pos + s.nGeom * (wi.z > 0.0f ? kEpsRay : -kEpsRay);
                    o.shD = ls.wi;
                    o.shMax = ls.dist - 2.0f * kEpsRay;
                    o.shC = beta * f * ls.radiance * (fabsf(wi.z) * w / ls.pdf);
                    o.emitShadow = true;
                    if (strata && ls.stratum >= 0) {
                        // the contribution as its stratum's own estimator would give it
                        const float c = luminance(o.shC) * (ls.stratum ? split : 1.0f - split);
                        o.shX = make_uint2((cell << 1) | uint32_t(ls.stratum), __float_as_uint(c));
                    }
                }
            }
            // BSDF bounce and Russian roulette (pathtrace.comp:200-216)
            if (!last) {
                V3 wiL, wiW, f;
                float pdf = 0.0f, pBsdf = 0.0f;
                bool ok;
                if (GUIDE && mix && rng.next1() < sp) {
                    // Draw from the learned mixture.
                    const float u0 = rng.next1();
                    float u1, u2;
                    rng.next2(u1, u2);
                    wiW = storedMixSample(mix, u0, u1, u2);
                    wiL = v3(dot(wiW, T), dot(wiW, B), dot(wiW, s.nShade));
                    bsdfEval(bsdf, wo, wiL, f, pBsdf);
                    pdf = sp * storedMixPdf(mix, wiW) + (1.0f - sp) * pBsdf;
                    ok = maxComp(f) > 0.0f && pdf > 1e-9f;
                } else {
                    const BsdfSample bs = bsdfSample(bsdf, wo, rng);
                    ok = bs.valid;
                    wiL = bs.wi;
                    f = bs.f;
                    pdf = bs.pdf;
                    pBsdf = bs.pdf;
                    if (ok) wiW = normalize(T * bs.wi.x + B * bs.wi.y + s.nShade * bs.wi.z);
                    if (GUIDE && mix && ok) {
                        pdf = sp * storedMixPdf(mix, wiW) + (1.0f - sp) * bs.pdf;
                        ok = pdf > 1e-9f;
                    }
                }
                if (ok) {
                    const float lumBeta = luminance(beta);
                    beta = beta * f * (fabsf(wiL.z) / pdf);
                    prevPdf = pdf;
                    if (GUIDE) L4.w = pBsdf;  // BSDF pdf of this continuation
                    // Guided engines: roulette on the throughput of BSDF
                    // sampling alone.  The guided throughput is small exactly
                    // on paths drawn towards bright regions (high mixture
                    // pdf); rouletting on it would end those paths most often.
                    // Any survival probability in (0, 1] keeps the estimator
                    // unbiased.  Plain engines have no rrScale.
                    float rrS = 1.0f;
                    if (pv.rrScale) rrS = clampf(pv.rrScale[p] * (pdf / fmaxf(pBsdf, 1e-20f)), 1e-6f, 1e6f);
                    bool alive = true;
                    if (bounce >= 3) {
                        const float q = pv.rrScale ? clampf(maxComp(beta) * rrS, 0.05f, 0.95f)
                                                   : clampf(maxComp(beta), 0.05f, 0.95f);
                        if (rng.next1() >= q) alive = false;
                        else beta = beta * (1.0f / q);
                    }
                    if (pv.rrScale) pv.rrScale[p] = rrS;
                    if (alive) {
                        o.newO = s.pos + s.nGeom * (wiL.z > 0.0f ? kEpsRay : -kEpsRay);
                        o.newD = wiW;
                        o.emitRay = true;
                        if (GUIDE && record) {
                            const uint32_t trainPaths = (pv.capacity + g.guideTrainStride - 1) / g.guideTrainStride;
                            float4* r = pv.pathRec + (size_t(bounce) * trainPaths + p / g.guideTrainStride) * 4;
                            r[0] = f4((s.pos - v3(g.guideBoxMin)) * v3(g.guideBoxScale), lumBeta);
                            r[1] = f4(s.nShade, luminance(L));
                            r[2] = f4(-dir, __uint_as_float(g.guideStamp));
                            r[3] = f4(wiW, 0.0f);
                            pv.recMask[p / g.guideTrainStride] |= 1u << bounce;
                        }
                    }
                }
            }
        }
    }
    pv.beta[p] = f4(beta, prevPdf);
    pv.radiance[p] = f4(L, L4.w);
    pv.rngCounter[p] = rng.counter;
    return o;
}

// Shading of items [base, base + 64) of segment `seg` of ray queue
// `bounce`: new rays to the same segment of queue bounce + 1, shadow rays to
// the same segment of the shadow queue (each item adds at most one of each,
// so a segment never outgrows its segCap slots).
template <bool GUIDE>
__device__ __forceinline__ void shadeChunk(const SceneView& sc, const PathView& pv, const FrameConsts& g,
                                           uint32_t bounce, uint32_t seg, uint32_t count, uint32_t base) {
    const uint32_t segBase = seg * pv.segCap;
    const uint32_t i = base + laneId();
    const uint32_t k = segBase + i;  // queue slot
    ShadeOut o;
    o.emitRay = false;
    o.emitShadow = false;
    uint32_t p = 0;
    if (i < count) {
        const float4 ro = pv.rayO[bounce & 1][k], rd = pv.rayD[bounce & 1][k];
        p = __float_as_uint(ro.w);
        o = shadePath<GUIDE>(sc, pv, g, bounce, p, v3(rd), pv.hit[k], pv.hitInst[k],
                             GUIDE && pv.guideIdx ? pv.guideIdx[k] : kNoHit);
    }
    const uint32_t rs = segBase + waveAppend(pv.counters + cntWord(bounce + 1, kCntRays, seg), o.emitRay);
    if (o.emitRay) {
        pv.rayO[(bounce + 1) & 1][rs] = f4(o.newO, __uint_as_float(p));
        pv.rayD[(bounce + 1) & 1][rs] = f4(o.newD, 0.0f);
    }
    const uint32_t ss = segBase + waveAppend(pv.counters + cntWord(bounce, kCntShadows, seg), o.emitShadow);
    if (o.emitShadow) {
        pv.shadowO[ss] = f4(o.shO, o.shMax);
        pv.shadowD[ss] = f4(o.shD, __uint_as_float(p));
        pv.shadowC[ss] = f4(o.shC, 0.0f);
        if (pv.shadowX) pv.shadowX[ss] = o.shX;
    }
}

// Shading of every active request of a batched frame at level `bounce`,
// except the requests the tail kernel serves there.  GUIDE: requests with
// guiding buffers shade with their learned mixtures (the query records and
// mixtures of this level are ready); plain requests in the same batch shade
// exactly as the plain kernel does (no mixture: the same operations and
// random numbers).  Vertices the tail kernel shades are unguided (one-sample
// MIS with selection probability 0, still unbiased).
template <bool GUIDE>
__global__ __launch_bounds__(256) void kShadeBatch(BatchView bv, uint32_t bounce, uint32_t tailPaths) {
    uint32_t n[kQueueSegs];
    laneSegmentCounts(laneCounters(bv), bounce, kCntRays, n);
    const bool tail = batchTailLane(bv, n, tailPaths);
    const uint32_t tot = segmentTotals(n, !tail);
    if (__ballot(tot > 0) == 0) return;  // nothing for this kernel: no atomics
    BatchChunks bc;
    segmentLoop(
        bv.cursors + cursorWord(bounce, 1, 0), tot,
        [&](uint32_t seg) { bc = batchChunks(laneCounters(bv), bounce, kCntRays, seg, tail, kPartLevel); },
        [&](uint32_t seg, uint32_t chunk) {
            uint32_t r, count, base;
            batchLocate(bc, chunk, r, count, base);
            const BatchEntry& e = bv.entries[r];
            const FrameConsts g = *e.g;
            shadeChunk<GUIDE>(e.sc, e.pv, g, bounce, seg, count, base);
        });
}

// ------------------------------------------------------------------ shadow --
// An unoccluded shadow ray of path p (bounce `bounce`) adds its contribution
// to the path's radiance.
__device__ __forceinline__ void applyShadow(const PathView& pv, const FrameConsts& g, uint32_t bounce, uint32_t p,
                                            V3 c) {
    float4 L = pv.radiance[p];
    L.x += c.x;
    L.y += c.y;
    L.z += c.z;
    pv.radiance[p] = L;
    const uint32_t ts = g.guideTrainStride;
    if (pv.pathRec && p % ts == g.guideTrainPhase) {
        // NEE from a recorded vertex is not incident along its wi: move the
        // record's radiance snapshot past it.
        float4* r = pv.pathRec + (size_t(bounce) * ((pv.capacity + ts - 1) / ts) + p / ts) * 4;
        if (__float_as_uint(r[2].w) == g.guideStamp) r[1].w += luminance(c);
    }
}

// Any hit for items [base, base + 64) of segment `seg` of one request's
// shadow queue; unoccluded ones add their contribution to the path's radiance.
template <int BLOCK, int S>
__device__ __forceinline__ void shadowChunk(const SceneView& sc, const PathView& pv, const FrameConsts* __restrict__ gp,
                                            uint32_t bounce, uint32_t seg, uint32_t count, uint32_t base, uint2* lds) {
    const uint32_t i = base + laneId();
    const uint32_t k = seg * pv.segCap + i;  // shadow queue slot
    bool overflow = false;
    if (i < count) {
        const float4 o = pv.shadowO[k], d = pv.shadowD[k];
        SceneHit h;
        const bool occluded = traceSceneWide<true, S, BLOCK>(sc, gp->tlasWideRoot, gp->instanceCount,
                                                         make_float3(o.x, o.y, o.z), make_float3(d.x, d.y, d.z),
                                                         kEpsRay, o.w, lds, h, overflow);
        if (!occluded) applyShadow(pv, *gp, bounce, __float_as_uint(d.w), v3(pv.shadowC[k]));
        if (pv.neeStats) {
            // second moment and count of the stratum's contributions (an
            // occluded sample counts as zero)
            const uint2 x = pv.shadowX[k];
            if (x.x != ~0u) {
                const float c = occluded ? 0.0f : __uint_as_float(x.y);
                float* cellSums = reinterpret_cast<float*>(pv.neeStats + (x.x >> 1)) + ((x.x & 1u) ? 0 : 2);
                atomicAdd(cellSums, c * c);
                atomicAdd(cellSums + 1, 1.0f);
            }
        }
    }
    if (overflow) atomicOr(pv.counters + kStatusWord, 1u);
}

// Closest hits of ray queue `extendLevel` and any hits of shadow queue
// `shadowLevel` (either may be kNoLevel) over all requests in one launch.
// The two are independent -- shadow rays only add to path radiance, closest
// hits only write hit records -- so the slowest rays of one overlap with the
// other's work instead of each ending a launch of its own.  Chunks
// [0, extend chunks) of the shared cursor are closest hit, the rest any hit.
// Register budget: 6 waves per SIMD (up to 80 VGPRs).  Traversal waits on
// dependent node fetches, and the sixth wave hides more of that latency than
// the registers it costs; at 8 waves (64 VGPRs) the spills outweigh it.
// Shapes whose LDS stacks allow fewer waves per SIMD (160 KB of LDS per CU,
// S x 8 bytes per lane and a camera-ray packet stack per wave) get the
// registers of the waves they can hold.  The variant that also writes the
// guided query records (PREP) runs at most 5 waves (96 VGPRs): at 6 it
// spilled 27 VGPRs; on the guided room its launches took 467 us at 6 and at
// 4 waves (113 VGPRs, no spills), 454 us at 5 (13 spilled).
constexpr int traceWavesPerSimd(int block, int stack) {
    const int blocksPerCU = (160 * 1024) / (stack * 8 * block + (block / 64) * kPacketStack * 4);
    const int waves = blocksPerCU * (block / 64) / 4;
    return waves < 1 ? 1 : (waves > 6 ? 6 : waves);
}
template <int BLOCK, int S, bool PREP>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(PREP ? (traceWavesPerSimd(BLOCK, S) > 5 ? 5 : traceWavesPerSimd(BLOCK, S)) : traceWavesPerSimd(BLOCK, S)))) void kTraceBatch(BatchView bv, uint32_t extendLevel, uint32_t shadowLevel,
                                                     uint32_t tailPaths) {
    __shared__ uint2 ldsStack[S * BLOCK];
    __shared__ uint32_t packetStack[BLOCK / 64][kPacketStack];  // per wave, for camera rays
    uint32_t n[kQueueSegs];
    laneSegmentCounts(extendLevel != kNoLevel ? laneCounters(bv) : nullptr, extendLevel, kCntRays, n);
    // query records for guided requests whose level the tail kernel does not take
    const bool tail = PREP && batchTailLane(bv, n, tailPaths);
    uint32_t tot = segmentTotals(n, true);
    laneSegmentCounts(shadowLevel != kNoLevel ? laneCounters(bv) : nullptr, shadowLevel, kCntShadows, n);
    tot += segmentTotals(n, true);
    if (__ballot(tot > 0) == 0) return;  // levels the tail kernel finished: no atomics
    BatchChunks be, bs;
    segmentLoop(
        bv.cursors + (extendLevel != kNoLevel ? cursorWord(extendLevel, 0, 0) : cursorWord(shadowLevel, 2, 0)), tot,
        [&](uint32_t seg) {
            const uint32_t* cnt = laneCounters(bv);
            be = extendLevel != kNoLevel ? batchChunks(cnt, extendLevel, kCntRays, seg) : BatchChunks{ 0, 0, 0 };
            bs = shadowLevel != kNoLevel ? batchChunks(cnt, shadowLevel, kCntShadows, seg) : BatchChunks{ 0, 0, 0 };
        },
        [&](uint32_t seg, uint32_t chunk) {
            uint32_t r, count, base;
            if (chunk < be.total) {
                batchLocate(be, chunk, r, count, base);
                const BatchEntry& e = bv.entries[r];
                const bool prep = PREP && e.pv.guideIdx != nullptr && !__shfl(int(tail), int(r), 64);
                extendChunk<BLOCK, S, PREP>(e.sc, e.pv, e.g, extendLevel, seg, count, base, ldsStack + threadIdx.x,
                                            prep, packetStack[threadIdx.x >> 6]);
            } else {
                batchLocate(bs, chunk - be.total, r, count, base);
                const BatchEntry& e = bv.entries[r];
                shadowChunk<BLOCK, S>(e.sc, e.pv, e.g, shadowLevel, seg, count, base, ldsStack + threadIdx.x);
            }
        });
}

// ------------------------------------------------------------------- tail --
// The rest of the frame for the paths of ray queue `bounce` (whose closest
// hits the trace launch of this level found), one lane per path: shade, its
// shadow ray, then closest hit and shading of every further bounce until the
// path ends.  Per path this is the same sequence of operations as the
// per-level kernels (shade b, shadow b, closest hit b + 1, ...), and no
// floating-point contraction is applied in this file, so the result does not
// depend on which kernel shaded a vertex.  Used when few paths are left: a
// per-level launch then mostly waits for its slowest ray, while here each
// wave only waits for its own paths.  Ray and shadow-ray counts of the levels
// it covers go to the kCntTail* words (statistics; the queues stay empty).
//
// A vertex's shadow ray and its continuation are independent, so they are
// traced together: the shadow ray of a lane that also continues moves to a
// lane of the wave that has no ray of its own this round (its path ended, or
// ends here without a shadow ray), and one traversal serves closest-hit and
// any-hit lanes alike.  Deep levels hold few live paths per wave, so a
// level's wait becomes the longer of the two walks instead of their sum.
// Shadow rays for which no lane is free are traced by their own lane after.
// pairs: 128 words of LDS for this wave.
template <int BLOCK, int S>
__device__ __forceinline__ void tailChunk(const SceneView& sc, const PathView& pv, const FrameConsts& g,
                                          uint32_t bounce, uint32_t seg, uint32_t count, uint32_t base, uint2* lds,
                                          uint32_t* pairs) {
    const uint32_t lane = laneId();
    const uint32_t i = base + lane;
    const uint32_t k = seg * pv.segCap + i;  // queue slot
    bool overflow = false;
    bool alive = i < count;
    uint32_t p = 0, inst = kNoHit;
    V3 dir = v3(0.0f, 0.0f, 0.0f);
    float4 h = make_float4(1e30f, 0.0f, 0.0f, __uint_as_float(kNoHit));
    if (alive) {
        p = __float_as_uint(pv.rayO[bounce & 1][k].w);
        dir = v3(pv.rayD[bounce & 1][k]);
        h = pv.hit[k];
        inst = pv.hitInst[k];
    }
    for (uint32_t b = bounce; __ballot(alive) != 0; ++b) {
        ShadeOut o;
        o.emitRay = false;
        o.emitShadow = false;
        if (alive) o = shadePath<false>(sc, pv, g, b, p, dir, h, inst, kNoHit);
        waveAppend(pv.counters + cntWord(b, kCntTailShadows, seg), o.emitShadow);
        waveAppend(pv.counters + cntWord(b + 1, kCntTailRays, seg), o.emitRay);
        // Pair the r-th lane that needs a helper with the r-th free lane.
        const bool cont = o.emitRay;
        const bool need = o.emitShadow && cont;
        const bool freeLane = !cont && !o.emitShadow;
        const uint64_t needMask = __ballot(need), freeMask = __ballot(freeLane);
        const uint32_t needRank = __builtin_amdgcn_mbcnt_hi(uint32_t(needMask >> 32),
                                                            __builtin_amdgcn_mbcnt_lo(uint32_t(needMask), 0u));
        const uint32_t freeRank = __builtin_amdgcn_mbcnt_hi(uint32_t(freeMask >> 32),
                                                            __builtin_amdgcn_mbcnt_lo(uint32_t(freeMask), 0u));
        const uint32_t nNeed = uint32_t(__builtin_popcountll(needMask));
        const uint32_t nFree = uint32_t(__builtin_popcountll(freeMask));
        if (need) pairs[needRank] = lane;
        if (freeLane) pairs[64 + freeRank] = lane;
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        const bool helper = freeLane && freeRank < nNeed;
        const bool helped = need && needRank < nFree;
        const uint32_t partner = helper ? pairs[freeRank] : (helped ? pairs[64 + needRank] : lane);
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        const float pox = __shfl(o.shO.x, int(partner), 64), poy = __shfl(o.shO.y, int(partner), 64),
                    poz = __shfl(o.shO.z, int(partner), 64);
        const float pdx = __shfl(o.shD.x, int(partner), 64), pdy = __shfl(o.shD.y, int(partner), 64),
                    pdz = __shfl(o.shD.z, int(partner), 64);
        const float pmax = __shfl(o.shMax, int(partner), 64);
        // One walk for every lane with a ray: continuations (closest hit),
        // shadow rays of lanes that end here, and the shadow rays handed over.
        const bool ownShadow = o.emitShadow && !cont;
        const bool anyRay = ownShadow || helper;
        bool traced = false;
        SceneHit hh;
        if (cont || anyRay) {
            const float3 ro = cont ? make_float3(o.newO.x, o.newO.y, o.newO.z)
                                   : (ownShadow ? make_float3(o.shO.x, o.shO.y, o.shO.z) : make_float3(pox, poy, poz));
            const float3 rd = cont ? make_float3(o.newD.x, o.newD.y, o.newD.z)
                                   : (ownShadow ? make_float3(o.shD.x, o.shD.y, o.shD.z) : make_float3(pdx, pdy, pdz));
            const float tmax = cont ? 1e30f : (ownShadow ? o.shMax : pmax);
            traced = traceSceneWide<false, S, BLOCK, false>(sc, g.tlasWideRoot, g.instanceCount, ro, rd, kEpsRay, tmax,
                                                            lds, hh, overflow, anyRay);
        }
        const uint64_t occHelped = __ballot(helper && traced);
        bool occluded = ownShadow && traced;
        if (helped) occluded = ((occHelped >> partner) & 1u) != 0;
        if (need && !helped) {
            SceneHit sh;
            occluded = traceSceneWide<true, S, BLOCK, false>(sc, g.tlasWideRoot, g.instanceCount,
                                                             make_float3(o.shO.x, o.shO.y, o.shO.z),
                                                             make_float3(o.shD.x, o.shD.y, o.shD.z), kEpsRay, o.shMax,
                                                             lds, sh, overflow);
        }
        if (o.emitShadow && !occluded) applyShadow(pv, g, b, p, o.shC);
        if (cont) {
            h = traced ? make_float4(hh.t, hh.b1, hh.b2, __uint_as_float(hh.tri))
                       : make_float4(1e30f, 0.0f, 0.0f, __uint_as_float(kNoHit));
            inst = traced ? hh.inst : kNoHit;
            dir = o.newD;
        }
        alive = cont;
    }
    if (overflow) atomicOr(pv.counters + kStatusWord, 1u);
}

// Serves the requests batchChunks assigns to the tail at level `bounce`
// (kShadeBatch serves the others).  Held to at most 168 VGPRs (3 waves per SIMD,
// a few spills): measured faster than 2 waves without spills (1080p frame
// 4.85 against 5.02 ms).  Shapes whose LDS (S x 8 bytes per lane and 512
// bytes per wave of hand-over slots, 160 KB per CU) holds fewer waves get
// the registers of the waves they can hold.
constexpr int tailWavesPerSimd(int block, int stack) {
    const int blocksPerCU = (160 * 1024) / (stack * 8 * block + (block / 64) * 512);
    const int waves = blocksPerCU * (block / 64) / 4;
    return waves < 1 ? 1 : (waves > 3 ? 3 : waves);
}
template <int BLOCK, int S>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(tailWavesPerSimd(BLOCK, S)))) void kTailBatch(BatchView bv,
                                                                                            uint32_t bounce,
                                                                                            uint32_t tailPaths) {
    __shared__ uint2 ldsStack[S * BLOCK];
    __shared__ uint32_t pairLds[BLOCK / 64][128];  // per wave: shadow-ray hand-over
    uint32_t n[kQueueSegs];
    laneSegmentCounts(laneCounters(bv), bounce, kCntRays, n);
    const bool tail = batchTailLane(bv, n, tailPaths);
    const uint32_t tot = segmentTotals(n, tail);
    if (__ballot(tot > 0) == 0) return;
    BatchChunks bc;
    segmentLoop(
        bv.cursors + cursorWord(bounce, 3, 0), tot,
        [&](uint32_t seg) { bc = batchChunks(laneCounters(bv), bounce, kCntRays, seg, tail, kPartTail); },
        [&](uint32_t seg, uint32_t chunk) {
            uint32_t r, count, base;
            batchLocate(bc, chunk, r, count, base);
            const BatchEntry& e = bv.entries[r];
            const FrameConsts g = *e.g;
            tailChunk<BLOCK, S>(e.sc, e.pv, g, bounce, seg, count, base, ldsStack + threadIdx.x,
                                pairLds[threadIdx.x >> 6]);
        });
}

// -------------------------------------------------------------- accumulate --
// pathtrace.comp:219-228
__device__ __forceinline__ void accumulatePixel(const PathView& pv, const FrameConsts& g, uint32_t p) {
    V3 L = v3(pv.radiance[p]);
    if (g.fireflyClamp > 0.0f) {
        const float m = maxComp(L);
        if (m > g.fireflyClamp) L = L * (g.fireflyClamp / m);
    }
    if (!isfinite(L.x) || !isfinite(L.y) || !isfinite(L.z)) L = v3(0.0f, 0.0f, 0.0f);
    const float4 prev = g.frameIndex == 0 ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : pv.accum[p];
    pv.accum[p] = make_float4(prev.x + L.x, prev.y + L.y, prev.z + L.z, prev.w + 1.0f);
}

// ----------------------------------------------------------------- resolve --
// resolve.comp:13-55: exposure, optional ACES (Narkowicz 2015), clamp, sRGB.
__device__ __forceinline__ float aces(float x) {
    const float a = 2.51f, b = 0.03f, c = 2.43f, d = 0.59f, e = 0.14f;
    return clampf((x * (a * x + b)) / (x * (c * x + d) + e), 0.0f, 1.0f);
}
__device__ __forceinline__ float linearToSrgb(float c) {
    return c <= 0.0031308f ? c * 12.92f : 1.055f * powf(c, 1.0f / 2.4f) - 0.055f;
}
__device__ __forceinline__ uint32_t unorm8(float c) {
    return uint32_t(rintf(clampf(c, 0.0f, 1.0f) * 255.0f));
}

__device__ __forceinline__ void resolvePixel(const PathView& pv, const FrameConsts& g, uint32_t p) {
    const float4 acc = pv.accum[p];
    float c[3] = { 0.0f, 0.0f, 0.0f };
    if (acc.w > 0.0f) { c[0] = acc.x / acc.w; c[1] = acc.y / acc.w; c[2] = acc.z / acc.w; }
    uint32_t px = 0xFF000000u;
    for (int i = 0; i < 3; ++i) {
        float x = c[i] * g.exposure;
        if (g.tonemap == 1) x = aces(x);
        x = linearToSrgb(clampf(x, 0.0f, 1e4f));
        px |= unorm8(x) << (8 * i);
    }
    pv.outRgba8[p] = px;
}

// Accumulate and resolve of every active request of a batched frame.
__global__ __launch_bounds__(256) void kFinishBatch(BatchView bv) {
    if (blockIdx.x == 0 && threadIdx.x < bv.count) {
        // The frame's queue profile, for the host's choice of graph depth.
        const PathView& pv = bv.entries[threadIdx.x].pv;
        if (bv.entries[threadIdx.x].active && pv.levelRays) {
            for (uint32_t b = 0; b < kMaxBounceLevels; ++b) {
                uint32_t n = 0;
                for (uint32_t q = 0; q < kQueueSegs; ++q) n += pv.counters[cntWord(b, kCntRays, q)];
                pv.levelRays[b] = n;
            }
        }
    }
    for (uint32_t w = (blockIdx.x * blockDim.x + threadIdx.x) & ~63u; w < bv.pixelSlots; w += gridDim.x * blockDim.x) {
        const uint32_t r = batchRequestOfSlot(bv, w);
        const BatchEntry& e = bv.entries[r];
        if (!e.active) continue;
        const uint32_t p = w - e.pixelBase + laneId();
        if (p < e.pixels) {
            const FrameConsts& g = *e.g;
            accumulatePixel(e.pv, g, p);
            resolvePixel(e.pv, g, p);
        }
    }
}

// RGBA32F readback value: accum.rgb * (1 / spp), alpha 1 (Renderer.cpp:900-908).
__global__ __launch_bounds__(256) void kResolve32(PathView pv, uint32_t n, uint32_t spp) {
    const float inv = spp > 0 ? 1.0f / float(spp) : 0.0f;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        const float4 acc = pv.accum[p];
        pv.outRgba32[p] = make_float4(acc.x * inv, acc.y * inv, acc.z * inv, 1.0f);
    }
}

__global__ void kClearWords(uint32_t* words, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) words[i] = 0u;
}

// ---------------------------------------------------------------- launchers --
static uint32_t gridFor(uint32_t n, uint32_t block, uint32_t cap) {
    uint32_t g = (n + block - 1) / block;
    if (g > cap) g = cap;
    return g ? g : 1;
}

hipError_t launchClearWords(uint32_t* words, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(kClearWords, dim3(gridFor(n, 256, 1024)), dim3(256), 0, s, words, n);
    return hipGetLastError();
}

#define MRT_PT_SHAPES(X) X(64, 16) X(128, 8) X(128, 16) X(128, 32) X(256, 8) X(256, 16)

hipError_t launchResolve32(const PathView& pv, uint32_t pixels, uint32_t spp, hipStream_t s) {
    hipLaunchKernelGGL(kResolve32, dim3(gridFor(pixels, 256, 8192)), dim3(256), 0, s, pv, pixels, spp);
    return hipGetLastError();
}

// ------------------------------------------------------------ batched frame --
template <int B, int S>
static hipError_t traceBatchT(const BatchView& bv, uint32_t extendLevel, uint32_t shadowLevel, uint32_t grid,
                              bool prep, uint32_t tailPaths, hipStream_t s) {
    if (prep)
        hipLaunchKernelGGL((kTraceBatch<B, S, true>), dim3(grid), dim3(B), 0, s, bv, extendLevel, shadowLevel, tailPaths);
    else
        hipLaunchKernelGGL((kTraceBatch<B, S, false>), dim3(grid), dim3(B), 0, s, bv, extendLevel, shadowLevel, tailPaths);
    return hipGetLastError();
}

hipError_t launchClearBatch(const BatchView& bv, hipStream_t s) {
    const uint32_t lines = bv.count * kCounterLines + kCursorLines;
    hipLaunchKernelGGL(kClearBatch, dim3(gridFor(lines, 256, 1024)), dim3(256), 0, s, bv);
    return hipGetLastError();
}

hipError_t launchGenerateBatch(const BatchView& bv, hipStream_t s) {
    hipLaunchKernelGGL(kGenerateBatch, dim3(gridFor(bv.pixelSlots, 256, 8192)), dim3(256), 0, s, bv);
    return hipGetLastError();
}

hipError_t launchTraceBatch(const BatchView& bv, uint32_t extendLevel, uint32_t shadowLevel, const TraceShape& shape,
                            uint32_t tailPaths, bool guidePrep, hipStream_t s) {
    if (extendLevel == kNoLevel && shadowLevel == kNoLevel) return hipSuccess;
    if ((extendLevel != kNoLevel && extendLevel >= kMaxBounceLevels) ||
        (shadowLevel != kNoLevel && shadowLevel >= kMaxBounceLevels))
        return hipErrorInvalidValue;
#define X(B, S)                                                                                  \
    if (shape.block == B && shape.stack == S)                                                     \
        return traceBatchT<B, S>(bv, extendLevel, shadowLevel, shape.grid, guidePrep, tailPaths, s);
    MRT_PT_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launchShadeBatch(const BatchView& bv, uint32_t bounce, uint32_t tailPaths, bool guided, uint32_t grid,
                            hipStream_t s) {
    if (guided)
        hipLaunchKernelGGL(kShadeBatch<true>, dim3(grid), dim3(256), 0, s, bv, bounce, tailPaths);
    else
        hipLaunchKernelGGL(kShadeBatch<false>, dim3(grid), dim3(256), 0, s, bv, bounce, tailPaths);
    return hipGetLastError();
}

template <int B, int S>
static hipError_t tailBatchT(const BatchView& bv, uint32_t bounce, uint32_t tailPaths, uint32_t grid, hipStream_t s) {
    hipLaunchKernelGGL((kTailBatch<B, S>), dim3(grid), dim3(B), 0, s, bv, bounce, tailPaths);
    return hipGetLastError();
}

hipError_t launchTailBatch(const BatchView& bv, uint32_t bounce, uint32_t tailPaths, const TraceShape& shape,
                           hipStream_t s) {
    if (tailPaths == 0) return hipSuccess;
    if (bounce >= kMaxBounceLevels) return hipErrorInvalidValue;
#define X(B, S) if (shape.block == B && shape.stack == S) return tailBatchT<B, S>(bv, bounce, tailPaths, shape.grid, s);
    MRT_PT_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t batchTailBlocksPerCU(uint32_t block, uint32_t stack, int* blocksPerCU) {
#define X(B, S) \
    if (block == B && stack == S) return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocksPerCU, kTailBatch<B, S>, B, 0);
    MRT_PT_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launchFinishBatch(const BatchView& bv, hipStream_t s) {
    hipLaunchKernelGGL(kFinishBatch, dim3(gridFor(bv.pixelSlots, 256, 8192)), dim3(256), 0, s, bv);
    return hipGetLastError();
}

hipError_t batchShadeBlocksPerCU(bool guided, int* blocksPerCU) {
    return guided ? hipOccupancyMaxActiveBlocksPerMultiprocessor(blocksPerCU, kShadeBatch<true>, 256, 0)
                  : hipOccupancyMaxActiveBlocksPerMultiprocessor(blocksPerCU, kShadeBatch<false>, 256, 0);
}

hipError_t batchTraceBlocksPerCU(uint32_t block, uint32_t stack, bool guidePrep, int* blocksPerCU) {
#define X(B, S)                                                                                              \
    if (block == B && stack == S)                                                                            \
        return guidePrep ? hipOccupancyMaxActiveBlocksPerMultiprocessor(blocksPerCU, kTraceBatch<B, S, true>, B, 0)  \
                         : hipOccupancyMaxActiveBlocksPerMultiprocessor(blocksPerCU, kTraceBatch<B, S, false>, B, 0);
    MRT_PT_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

bool ptShapeSupported(uint32_t block, uint32_t stack) {
#define X(B, S) if (block == B && stack == S) return true;
    MRT_PT_SHAPES(X)
#undef X
    return false;
}


} // namespace dev
} // namespace mrt
