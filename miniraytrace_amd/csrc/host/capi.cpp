// The 27 C entry points of mini_raytrace.h plus the engine-level extensions
// of mini_raytrace_ext.h.  Conventions of the reference capi/src/capi.cpp:
// structs are accepted when non-NULL with structSize >= the size this
// library was built with; Add* return MRT_INVALID_ID on failure; void
// mutators ignore bad handles; exceptions never cross the boundary.
#include "capi_handles.h"

#include "common.h"
#include "engine.h"
#include "mini_raytrace.h"
#include "mini_raytrace_ext.h"

#include <cstddef>
#include <memory>

namespace {

using mrt::unwrap;
using mrt::validStruct;

// Runs a void mutator; failures are recorded as the last error only.
template <typename F>
void quiet(F&& f) {
    (void)mrt::guarded([&] {
        f();
        return MRT_SUCCESS;
    });
}

template <typename F>
uint32_t idCall(F&& f) {
    uint32_t id = MRT_INVALID_ID;
    (void)mrt::guarded([&] {
        id = f();
        return MRT_SUCCESS;
    });
    return id;
}

} // namespace

extern "C" {

MRT_API mrtResult mrtCreateEngine(const mrtEngineDesc* desc, mrtEngine** outEngine) {
    if (!validStruct(desc) || !outEngine) return MRT_ERROR_INVALID_ARGUMENT;
    return mrt::guarded([&] {
        auto w = std::make_unique<mrtEngine>();
        w->engine = std::make_unique<mrt::Engine>(*desc);
        *outEngine = w.release();
        return MRT_SUCCESS;
    });
}

MRT_API void mrtDestroyEngine(mrtEngine* engine) {
    quiet([&] { delete engine; });
}

MRT_API mrtMeshId mrtAddMesh(mrtEngine* e, const mrtMeshDesc* d) {
    mrt::Engine* eng = unwrap(e);
    if (!eng || !validStruct(d)) return MRT_INVALID_ID;
    return idCall([&] { return eng->scene().addMesh(*d); });
}

MRT_API void mrtUpdateMesh(mrtEngine* e, mrtMeshId id, const mrtMeshDesc* d) {
    mrt::Engine* eng = unwrap(e);
    if (eng && validStruct(d)) quiet([&] { eng->scene().updateMesh(id, *d); });
}

MRT_API void mrtRemoveMesh(mrtEngine* e, mrtMeshId id) {
    if (mrt::Engine* eng = unwrap(e)) quiet([&] { eng->scene().removeMesh(id); });
}

MRT_API mrtInstanceId mrtAddInstance(mrtEngine* e, mrtMeshId mesh, const float* xform3x4, mrtMaterialId mat) {
    mrt::Engine* eng = unwrap(e);
    if (!eng) return MRT_INVALID_ID;
    return idCall([&] { return eng->scene().addInstance(mesh, xform3x4, mat); });
}

MRT_API void mrtSetInstanceTransform(mrtEngine* e, mrtInstanceId id, const float* xform3x4) {
    if (mrt::Engine* eng = unwrap(e)) quiet([&] { eng->scene().setInstanceTransform(id, xform3x4); });
}

MRT_API void mrtRemoveInstance(mrtEngine* e, mrtInstanceId id) {
    if (mrt::Engine* eng = unwrap(e)) quiet([&] { eng->scene().removeInstance(id); });
}

MRT_API mrtMaterialId mrtAddMaterial(mrtEngine* e, const mrtMaterialDesc* d) {
    mrt::Engine* eng = unwrap(e);
    if (!eng || !validStruct(d)) return MRT_INVALID_ID;
    return idCall([&] { return eng->scene().addMaterial(*d); });
}

MRT_API void mrtUpdateMaterial(mrtEngine* e, mrtMaterialId id, const mrtMaterialDesc* d) {
    mrt::Engine* eng = unwrap(e);
    if (eng && validStruct(d)) quiet([&] { eng->scene().updateMaterial(id, *d); });
}

MRT_API void mrtRemoveMaterial(mrtEngine* e, mrtMaterialId id) {
    if (mrt::Engine* eng = unwrap(e)) quiet([&] { eng->scene().removeMaterial(id); });
}

MRT_API mrtTextureId mrtAddTexture(mrtEngine* e, const mrtTextureDesc* d) {
    mrt::Engine* eng = unwrap(e);
    if (!eng || !validStruct(d)) return MRT_INVALID_ID;
    return idCall([&] { return eng->scene().addTexture(*d); });
}

MRT_API void mrtRemoveTexture(mrtEngine* e, mrtTextureId id) {
    if (mrt::Engine* eng = unwrap(e)) quiet([&] { eng->scene().removeTexture(id); });
}

MRT_API mrtLightId mrtAddLight(mrtEngine* e, const mrtLightDesc* d) {
    mrt::Engine* eng = unwrap(e);
    if (!eng || !validStruct(d)) return MRT_INVALID_ID;
    return idCall([&] { return eng->scene().addLight(*d); });
}

MRT_API void mrtUpdateLight(mrtEngine* e, mrtLightId id, const mrtLightDesc* d) {
    mrt::Engine* eng = unwrap(e);
    if (eng && validStruct(d)) quiet([&] { eng->scene().updateLight(id, *d); });
}

MRT_API void mrtRemoveLight(mrtEngine* e, mrtLightId id) {
    if (mrt::Engine* eng = unwrap(e)) quiet([&] { eng->scene().removeLight(id); });
}

MRT_API void mrtSetEnvironment(mrtEngine* e, const mrtTextureDesc* hdri, float rotationRad, float intensity) {
    mrt::Engine* eng = unwrap(e);
    if (!eng) return;
    if (!hdri) {
        quiet([&] { eng->scene().clearEnvironment(); });
        return;
    }
    if (validStruct(hdri)) quiet([&] { eng->scene().setEnvironment(*hdri, rotationRad, intensity); });
}

MRT_API void mrtSetCamera(mrtEngine* e, const mrtCameraDesc* d) {
    mrt::Engine* eng = unwrap(e);
    if (eng && validStruct(d)) quiet([&] { eng->scene().setCamera(*d); });
}

MRT_API void mrtSetCameraEx(mrtEngine* e, const mrtCameraDescEx* d) {
    mrt::Engine* eng = unwrap(e);
    if (eng && validStruct(d)) quiet([&] { eng->scene().setCameraEx(*d); });
}

MRT_API mrtResult mrtCommitScene(mrtEngine* e, mrtCommitStats* outStats) {
    mrt::Engine* eng = unwrap(e);
    if (!eng) return MRT_ERROR_INVALID_HANDLE;
    return mrt::guarded([&] {
        mrtCommitStats st{};
        eng->commit(&st);
        if (outStats && outStats->structSize >= sizeof(mrtCommitStats)) {
            outStats->blasRebuilt = st.blasRebuilt;
            outStats->tlasRebuilt = st.tlasRebuilt;
            outStats->commitMs = st.commitMs;
        }
        return MRT_SUCCESS;
    });
}

MRT_API mrtResult mrtSetRenderSettings(mrtEngine* e, const mrtRenderSettings* d) {
    mrt::Engine* eng = unwrap(e);
    if (!eng || !validStruct(d)) return MRT_ERROR_INVALID_ARGUMENT;
    return mrt::guarded([&] {
        mrt::Settings s;
        s.width = d->width;
        s.height = d->height;
        s.maxBounces = d->maxBounces;
        s.sppLimit = d->sppLimit;
        s.exposureEV = d->exposureEV;
        s.fireflyClamp = d->fireflyClamp;
        s.tonemap = uint32_t(d->tonemap);
        s.denoise = d->denoise != 0;
        eng->setSettings(s);
        return MRT_SUCCESS;
    });
}

MRT_API mrtResult mrtRenderFrame(mrtEngine* e, mrtFrameInfo* outInfo) {
    mrt::Engine* eng = unwrap(e);
    if (!eng) return MRT_ERROR_INVALID_HANDLE;
    return mrt::guarded([&] {
        mrtFrameInfo info{};
        eng->renderFrame(&info);
        if (outInfo && outInfo->structSize >= sizeof(mrtFrameInfo)) {
            outInfo->spp = info.spp;
            outInfo->converged = info.converged;
            outInfo->denoised = info.denoised;
            outInfo->frameMs = info.frameMs;
        }
        return MRT_SUCCESS;
    });
}

MRT_API mrtResult mrtReadFramebuffer(mrtEngine* e, mrtPixelFormat fmt, void* dst, size_t dstSize) {
    mrt::Engine* eng = unwrap(e);
    if (!eng || !dst) return MRT_ERROR_INVALID_ARGUMENT;
    return mrt::guarded([&] {
        eng->readFramebuffer(fmt, dst, dstSize);
        return MRT_SUCCESS;
    });
}

MRT_API void mrtResetAccumulation(mrtEngine* e) {
    if (mrt::Engine* eng = unwrap(e)) eng->resetAccumulation();
}

MRT_API void mrtSetLogCallback(mrtLogFn fn, void* user) { mrt::setLogCallback(fn, user); }

MRT_API const char* mrtResultToString(mrtResult r) { return mrt::resultToString(r); }

MRT_API const char* mrtGetLastErrorMessage(void) { return mrt::lastErrorMessage().c_str(); }

// ---- extensions ----------------------------------------------------------------

MRT_API mrtResult mrtxSetPathOptions(mrtEngine* e, const mrtxPathOptions* o) {
    mrt::Engine* eng = unwrap(e);
    if (!eng) return MRT_ERROR_INVALID_HANDLE;
    if (!validStruct(o)) return MRT_ERROR_INVALID_ARGUMENT;
    return mrt::guarded([&] {
        eng->setPathOptions(*o);
        return MRT_SUCCESS;
    });
}

MRT_API mrtResult mrtxGetPathOptions(mrtEngine* e, mrtxPathOptions* o) {
    mrt::Engine* eng = unwrap(e);
    if (!eng) return MRT_ERROR_INVALID_HANDLE;
    if (!validStruct(o)) return MRT_ERROR_INVALID_ARGUMENT;
    eng->getPathOptions(*o);
    return MRT_SUCCESS;
}

MRT_API mrtResult mrtxSetProfiling(mrtEngine* e, int32_t enable) {
    mrt::Engine* eng = unwrap(e);
    if (!eng) return MRT_ERROR_INVALID_HANDLE;
    eng->setProfiling(enable != 0);
    return MRT_SUCCESS;
}

MRT_API mrtResult mrtxGetFrameStats(mrtEngine* e, mrtxFrameStats* out) {
    mrt::Engine* eng = unwrap(e);
    if (!eng) return MRT_ERROR_INVALID_HANDLE;
    if (!out || out->structSize < offsetof(mrtxFrameStats, commitMs)) return MRT_ERROR_INVALID_ARGUMENT;
    const mrt::FrameStats& s = eng->stats();
    out->frameMs = s.frameMs;
    for (int i = 0; i < 8; ++i) out->gpuMs[i] = s.stageMs[i];
    for (uint32_t b = 0; b < MRTX_MAX_BOUNCE_LEVELS; ++b) {
        out->rays[b] = s.rays[b];
        out->shadowRays[b] = s.shadowRays[b];
    }
    out->status = s.status;
    if (out->structSize < offsetof(mrtxFrameStats, graphTailLevel)) return MRT_SUCCESS;
    out->commitMs = s.commitMs;
    out->commitUploadBytes = s.commitUploadBytes;
    out->geometryRepacks = s.geometryRepacks;
    out->meshesWritten = s.meshesWritten;
    if (out->structSize < offsetof(mrtxFrameStats, graphTailLevel) + sizeof(uint32_t)) return MRT_SUCCESS;
    out->graphTailLevel = s.graphTailLevel;
    return MRT_SUCCESS;
}

MRT_API mrtResult mrtxExportScene(mrtEngine* e, mrtxSceneExport* io) {
    mrt::Engine* eng = unwrap(e);
    if (!eng) return MRT_ERROR_INVALID_HANDLE;
    // the fields from `tangents` on were appended: the shorter struct is accepted
    if (!io || io->structSize < offsetof(mrtxSceneExport, tangents)) return MRT_ERROR_INVALID_ARGUMENT;
    return mrt::guarded([&] {
        eng->exportScene(*io);
        return MRT_SUCCESS;
    });
}

} // extern "C"
