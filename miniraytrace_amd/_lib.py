"""ctypes binding of libmrt.so (the C ABI in include/mini_raytrace.h and
include/mini_raytrace_ext.h).

The library is loaded from this package directory.  There is no fallback:
if libmrt.so is missing or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libmrt.so"

# ---- enums (mini_raytrace.h) ------------------------------------------------
MRT_SUCCESS = 0
MRT_ERROR_UNKNOWN = 1
MRT_ERROR_VULKAN_INIT = 2
MRT_ERROR_OUT_OF_MEMORY = 3
MRT_ERROR_INVALID_ARGUMENT = 4
MRT_ERROR_INVALID_HANDLE = 5
MRT_ERROR_SHADER_LOAD = 6
MRT_ERROR_DEVICE_LOST = 7
MRT_INVALID_ID = 0

MRT_PIXEL_RGBA8_SRGB = 0
MRT_PIXEL_RGBA32F_LINEAR = 1
MRT_TEX_RGBA8_SRGB = 0
MRT_TEX_RGBA8_UNORM = 1
MRT_TEX_RGBA32F = 2
MRT_LIGHT_SUN = 0
MRT_LIGHT_POINT = 1
MRT_LIGHT_RECT = 2
MRT_TONEMAP_LINEAR = 0
MRT_TONEMAP_ACES = 1
MRT_CAMERA_PERSPECTIVE = 0
MRT_CAMERA_ORTHOGRAPHIC = 1
MRTX_NO_HIT = 0xFFFFFFFF


class _Sized(C.Structure):
    """Structs whose first field is structSize; filled in automatically."""

    def __init__(self, **kw):
        super().__init__(**kw)
        self.structSize = C.sizeof(self)


F3 = C.c_float * 3


class mrtEngineDesc(_Sized):
    _fields_ = [("structSize", C.c_size_t), ("enableValidation", C.c_int32),
                ("width", C.c_uint32), ("height", C.c_uint32), ("gpuIndex1", C.c_uint32)]


class mrtMeshDesc(_Sized):
    _fields_ = [("structSize", C.c_size_t), ("vertexCount", C.c_uint32),
                ("indexCount", C.c_uint32), ("positions", C.POINTER(C.c_float)),
                ("normals", C.POINTER(C.c_float)), ("uvs", C.POINTER(C.c_float)),
                ("indices", C.POINTER(C.c_uint32))]


class mrtMaterialDesc(_Sized):
    _fields_ = [("structSize", C.c_size_t), ("baseColor", F3), ("opacity", C.c_float),
                ("roughness", C.c_float), ("metallic", C.c_float),
                ("transmission", C.c_float), ("ior", C.c_float), ("emission", F3),
                ("baseColorTex", C.c_uint32), ("roughnessTex", C.c_uint32),
                ("metallicTex", C.c_uint32), ("normalTex", C.c_uint32),
                ("emissionTex", C.c_uint32)]


class mrtTextureDesc(_Sized):
    _fields_ = [("structSize", C.c_size_t), ("width", C.c_uint32), ("height", C.c_uint32),
                ("format", C.c_int), ("pixels", C.c_void_p), ("generateMips", C.c_int32)]


class mrtLightDesc(_Sized):
    _fields_ = [("structSize", C.c_size_t), ("type", C.c_int), ("direction", F3),
                ("position", F3), ("radiance", F3), ("angularRadius", C.c_float),
                ("corner", F3), ("edge0", F3), ("edge1", F3), ("twoSided", C.c_int32)]


class mrtCameraDesc(_Sized):
    _fields_ = [("structSize", C.c_size_t), ("position", F3), ("target", F3),
                ("up", F3), ("fovYDeg", C.c_float)]


class mrtCameraDescEx(_Sized):
    _fields_ = [("structSize", C.c_size_t), ("projection", C.c_int32), ("position", F3),
                ("forward", F3), ("up", F3), ("left", C.c_float), ("right", C.c_float),
                ("bottom", C.c_float), ("top", C.c_float)]


class mrtRenderSettings(_Sized):
    _fields_ = [("structSize", C.c_size_t), ("width", C.c_uint32), ("height", C.c_uint32),
                ("maxBounces", C.c_uint32), ("sppLimit", C.c_uint32),
                ("exposureEV", C.c_float), ("fireflyClamp", C.c_float),
                ("tonemap", C.c_int), ("denoise", C.c_int32)]


class mrtFrameInfo(_Sized):
    _fields_ = [("structSize", C.c_size_t), ("spp", C.c_uint32), ("converged", C.c_int32),
                ("denoised", C.c_int32), ("frameMs", C.c_float)]


class mrtCommitStats(_Sized):
    _fields_ = [("structSize", C.c_size_t), ("blasRebuilt", C.c_uint32),
                ("tlasRebuilt", C.c_int32), ("commitMs", C.c_float)]


# ---- ext structs (mini_raytrace_ext.h) ---------------------------------------
class mrtxQuerySceneInfo(_Sized):
    _fields_ = [("structSize", C.c_size_t), ("triCount", C.c_uint32),
                ("nodeCount", C.c_uint32), ("leafCount", C.c_uint32),
                ("maxDepth", C.c_uint32), ("buildMs", C.c_float),
                ("deviceBytes", C.c_uint64)]


class mrtxTraversalConfig(_Sized):
    _fields_ = [("structSize", C.c_size_t), ("blockSize", C.c_uint32),
                ("stackDepth", C.c_uint32), ("loopMode", C.c_uint32)]


class mrtxPathOptions(_Sized):
    _fields_ = [("structSize", C.c_size_t), ("seed", C.c_uint32), ("useGraph", C.c_uint32),
                ("extendBlock", C.c_uint32), ("extendStack", C.c_uint32),
                ("shadowBlock", C.c_uint32), ("shadowStack", C.c_uint32), ("tailPaths", C.c_uint32),
                ("cameraPerLane", C.c_uint32)]


TAIL_DISABLED = 0xFFFFFFFF

MRTX_MAX_BOUNCE_LEVELS = 18


class mrtxFrameStats(_Sized):
    _fields_ = [("structSize", C.c_size_t), ("frameMs", C.c_float), ("gpuMs", C.c_float * 8),
                ("rays", C.c_uint32 * MRTX_MAX_BOUNCE_LEVELS),
                ("shadowRays", C.c_uint32 * MRTX_MAX_BOUNCE_LEVELS), ("status", C.c_uint32),
                ("commitMs", C.c_float), ("commitUploadBytes", C.c_uint64), ("geometryRepacks", C.c_uint32),
                ("meshesWritten", C.c_uint32), ("graphTailLevel", C.c_uint32)]


class mrtxBatchRequest(C.Structure):
    _fields_ = [("engine", C.c_void_p), ("frames", C.c_uint32), ("framesDone", C.c_uint32)]


class mrtxBatchStats(_Sized):
    _fields_ = [("structSize", C.c_size_t), ("captureMs", C.c_float), ("launchMs", C.c_float),
                ("requests", C.c_uint32), ("frames", C.c_uint32), ("mode", C.c_uint32),
                ("extendBlock", C.c_uint32), ("extendStack", C.c_uint32), ("shadowBlock", C.c_uint32),
                ("shadowStack", C.c_uint32)]


MRTX_GUIDE_PARAM_COUNT = 20136
MRTX_GUIDE_OUTPUTS = 40
MRTX_GUIDE_NOALLOC_MAX_RECORDS = 8192
MRTX_TEXTURE_SLOTS = 64


class mrtxGuidingOptions(_Sized):
    _fields_ = [("structSize", C.c_size_t), ("enableSampling", C.c_uint32), ("enableTraining", C.c_uint32),
                ("selectProb", C.c_float), ("learningRate", C.c_float), ("batchSize", C.c_uint32),
                ("stepsPerFrame", C.c_uint32), ("minRoughness", C.c_float), ("seed", C.c_uint32),
                ("positionScale", C.c_float), ("warmupFrames", C.c_uint32), ("warmupStepsPerFrame", C.c_uint32)]


class mrtxGuidingStats(_Sized):
    _fields_ = [("structSize", C.c_size_t), ("steps", C.c_uint32), ("lastBatch", C.c_uint32),
                ("lastLoss", C.c_float), ("poolRecords", C.c_uint32), ("candidates", C.c_uint32),
                ("lastFrameRows", C.c_uint32), ("envBrightShare", C.c_float), ("envCellsLearned", C.c_uint32),
                ("envMeanSplit", C.c_float)]


class mrtxSceneExport(_Sized):
    _fields_ = [("structSize", C.c_size_t), ("nodeCount", C.c_uint32), ("triCount", C.c_uint32),
                ("vertexCount", C.c_uint32), ("instanceCount", C.c_uint32),
                ("materialCount", C.c_uint32), ("lightCount", C.c_uint32), ("tlasRoot", C.c_uint32),
                ("blasNodeCount", C.c_uint32), ("envWidth", C.c_uint32), ("envHeight", C.c_uint32),
                ("envIntegral", C.c_float), ("nodes", C.c_void_p), ("triangles", C.c_void_p),
                ("positions", C.c_void_p), ("normals", C.c_void_p), ("uvs", C.c_void_p),
                ("instances", C.c_void_p), ("materials", C.c_void_p), ("lights", C.c_void_p),
                ("envAlias", C.c_void_p), ("tangents", C.c_void_p), ("textureTable", C.c_void_p),
                ("tex8Count", C.c_uint32), ("tex32Count", C.c_uint32), ("tex8", C.c_void_p),
                ("tex32", C.c_void_p)]


LOG_FN = C.CFUNCTYPE(None, C.c_int, C.c_char_p, C.c_void_p)

_P = C.c_void_p
_U32 = C.c_uint32
_R = C.c_int  # mrtResult

# name -> (restype, argtypes)
_SIGS = {
    "mrtCreateEngine": (_R, [C.POINTER(mrtEngineDesc), C.POINTER(_P)]),
    "mrtDestroyEngine": (None, [_P]),
    "mrtAddMesh": (_U32, [_P, C.POINTER(mrtMeshDesc)]),
    "mrtUpdateMesh": (None, [_P, _U32, C.POINTER(mrtMeshDesc)]),
    "mrtRemoveMesh": (None, [_P, _U32]),
    "mrtAddInstance": (_U32, [_P, _U32, C.POINTER(C.c_float), _U32]),
    "mrtSetInstanceTransform": (None, [_P, _U32, C.POINTER(C.c_float)]),
    "mrtRemoveInstance": (None, [_P, _U32]),
    "mrtAddMaterial": (_U32, [_P, C.POINTER(mrtMaterialDesc)]),
    "mrtUpdateMaterial": (None, [_P, _U32, C.POINTER(mrtMaterialDesc)]),
    "mrtRemoveMaterial": (None, [_P, _U32]),
    "mrtAddTexture": (_U32, [_P, C.POINTER(mrtTextureDesc)]),
    "mrtRemoveTexture": (None, [_P, _U32]),
    "mrtAddLight": (_U32, [_P, C.POINTER(mrtLightDesc)]),
    "mrtUpdateLight": (None, [_P, _U32, C.POINTER(mrtLightDesc)]),
    "mrtRemoveLight": (None, [_P, _U32]),
    "mrtSetEnvironment": (None, [_P, C.POINTER(mrtTextureDesc), C.c_float, C.c_float]),
    "mrtSetCamera": (None, [_P, C.POINTER(mrtCameraDesc)]),
    "mrtSetCameraEx": (None, [_P, C.POINTER(mrtCameraDescEx)]),
    "mrtCommitScene": (_R, [_P, C.POINTER(mrtCommitStats)]),
    "mrtSetRenderSettings": (_R, [_P, C.POINTER(mrtRenderSettings)]),
    "mrtRenderFrame": (_R, [_P, C.POINTER(mrtFrameInfo)]),
    "mrtReadFramebuffer": (_R, [_P, C.c_int, _P, C.c_size_t]),
    "mrtResetAccumulation": (None, [_P]),
    "mrtSetLogCallback": (None, [LOG_FN, _P]),
    "mrtResultToString": (C.c_char_p, [C.c_int]),
    "mrtGetLastErrorMessage": (C.c_char_p, []),
    # ext: ray queries
    "mrtxBuildBvh": (_R, [_P, _U32, _U32, _P, C.c_uint64, _P, C.POINTER(C.c_uint64)]),
    "mrtxBuildBvhDevice": (_R, [_P, _U32, _U32, _P, C.c_uint64, _P, C.POINTER(C.c_uint64), _P]),
    "mrtxCreateQueryScene": (_R, [_P, _U32, _P, _U32, _U32, C.POINTER(_P)]),
    "mrtxDestroyQueryScene": (None, [_P]),
    "mrtxGetQuerySceneInfo": (_R, [_P, C.POINTER(mrtxQuerySceneInfo)]),
    "mrtxExportQueryBvh": (_R, [_P, _P, C.c_size_t, _P, C.c_size_t]),
    "mrtxSetTraversalConfig": (_R, [_P, C.POINTER(mrtxTraversalConfig)]),
    "mrtxGetTraversalConfig": (_R, [_P, C.POINTER(mrtxTraversalConfig)]),
    "mrtxTraceClosest": (_R, [_P, _P, _U32, _P, _P]),
    "mrtxTraceAny": (_R, [_P, _P, _U32, _P, _P]),
    "mrtxGetQueryStatus": (_R, [_P, C.POINTER(_U32)]),
    "mrtxAutotuneTraversal": (_R, [_P, _P, _U32, _P, C.POINTER(mrtxTraversalConfig)]),
    # ext: engine
    "mrtxSetPathOptions": (_R, [_P, C.POINTER(mrtxPathOptions)]),
    "mrtxGetPathOptions": (_R, [_P, C.POINTER(mrtxPathOptions)]),
    "mrtxSetProfiling": (_R, [_P, C.c_int32]),
    "mrtxGetFrameStats": (_R, [_P, C.POINTER(mrtxFrameStats)]),
    "mrtxExportScene": (_R, [_P, C.POINTER(mrtxSceneExport)]),
    # ext: batching
    "mrtxRenderBatch": (_R, [C.POINTER(mrtxBatchRequest), _U32, C.POINTER(mrtxBatchStats)]),
    "mrtxAutotunePath": (_R, [_P, C.POINTER(mrtxPathOptions)]),
    # ext: path guiding
    "mrtxSetGuiding": (_R, [_P, C.POINTER(mrtxGuidingOptions)]),
    "mrtxGetGuiding": (_R, [_P, C.POINTER(mrtxGuidingOptions)]),
    "mrtxGetGuidingStats": (_R, [_P, C.POINTER(mrtxGuidingStats)]),
    "mrtxGetGuideParams": (_R, [_P, _P, C.c_size_t]),
    "mrtxSetGuideParams": (_R, [_P, _P, C.c_size_t]),
    "mrtxCreateGuideNetwork": (_R, [C.c_int32, _U32, _U32, C.POINTER(_P)]),
    "mrtxDestroyGuideNetwork": (None, [_P]),
    "mrtxGuideNetworkSetParams": (_R, [_P, _P, C.c_size_t]),
    "mrtxGuideNetworkGetParams": (_R, [_P, _P, C.c_size_t]),
    "mrtxGuideNetworkGetAdam": (_R, [_P, _P, _P, C.c_size_t, C.POINTER(_U32)]),
    "mrtxGuideNetworkInfer": (_R, [_P, _P, _U32, _P]),
    "mrtxGuideNetworkGradient": (_R, [_P, _P, _U32, _P, C.c_size_t, C.POINTER(C.c_float)]),
    "mrtxGuideNetworkTrain": (_R, [_P, _P, _U32, C.c_float, C.POINTER(C.c_float)]),
    "mrtxGuideNetworkTrainNoAlloc": (_R, [_P, _P, _U32, C.c_float, _P, C.POINTER(C.c_float)]),
    "mrtxGuideNetworkSample": (_R, [_P, _P, _P, _U32, _P]),
    "mrtxGuideNetworkPdf": (_R, [_P, _P, _P, _U32, _P]),
}

# The 27 entry points of the reference ABI (capi/include/mini_raytrace.h:184-237).
ABI_SYMBOLS = [n for n in _SIGS if n.startswith("mrt") and not n.startswith("mrtx")]

_lib = None


class MrtError(RuntimeError):
    def __init__(self, code: int, what: str):
        self.code = code
        super().__init__(f"{what}: {result_name(code)} ({code}): {last_error()}")


def lib() -> C.CDLL:
    """Loads libmrt.so (once).  Raises if it is missing: there is no fallback."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing; build it with "
                               "`python -m miniraytrace_amd.build`")
        handle = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            try:
                fn = getattr(handle, name)
            except AttributeError:
                continue  # reported by tests/test_abi.py; calling it raises
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def result_name(code: int) -> str:
    return lib().mrtResultToString(code).decode()


def last_error() -> str:
    return lib().mrtGetLastErrorMessage().decode()


def check(code: int, what: str) -> None:
    if code != MRT_SUCCESS:
        raise MrtError(code, what)
