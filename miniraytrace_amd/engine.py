"""Python mirror of the MiniRaytrace C ABI (include/mini_raytrace.h).

`Engine` wraps one mrtEngine handle with the same operations, argument
meaning and error behaviour as the C functions: Add* return an id
(MRT_INVALID_ID = 0 on failure), void mutators silently ignore bad input,
mrtResult-returning calls raise MrtError on failure.  Everything goes
through libmrt.so; there is no Python-side rendering path.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L


def _f3(v) -> L.F3:
    return L.F3(*[float(x) for x in v])


class Engine:
    def __init__(self, width: int = 0, height: int = 0, gpu_index1: int = 0, validation: bool = False):
        desc = L.mrtEngineDesc(enableValidation=int(validation), width=width, height=height,
                               gpuIndex1=gpu_index1)
        h = C.c_void_p()
        L.check(L.lib().mrtCreateEngine(C.byref(desc), C.byref(h)), "mrtCreateEngine")
        self._h = h
        self._keep = []

    # -- lifecycle ---------------------------------------------------------------
    def close(self) -> None:
        if self._h:
            L.lib().mrtDestroyEngine(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    # -- meshes ------------------------------------------------------------------
    @staticmethod
    def _mesh_desc(positions, indices, normals=None, uvs=None):
        pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
        idx = np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1)
        d = L.mrtMeshDesc(vertexCount=pos.shape[0], indexCount=idx.size)
        d.positions = pos.ctypes.data_as(C.POINTER(C.c_float))
        d.indices = idx.ctypes.data_as(C.POINTER(C.c_uint32))
        keep = [pos, idx]
        if normals is not None:
            n = np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 3)
            d.normals = n.ctypes.data_as(C.POINTER(C.c_float))
            keep.append(n)
        if uvs is not None:
            u = np.ascontiguousarray(uvs, dtype=np.float32).reshape(-1, 2)
            d.uvs = u.ctypes.data_as(C.POINTER(C.c_float))
            keep.append(u)
        return d, keep

    def add_mesh(self, positions, indices, normals=None, uvs=None) -> int:
        d, keep = self._mesh_desc(positions, indices, normals, uvs)
        return L.lib().mrtAddMesh(self._h, C.byref(d))

    def update_mesh(self, mesh_id: int, positions, indices, normals=None, uvs=None) -> None:
        d, keep = self._mesh_desc(positions, indices, normals, uvs)
        L.lib().mrtUpdateMesh(self._h, mesh_id, C.byref(d))

    def remove_mesh(self, mesh_id: int) -> None:
        L.lib().mrtRemoveMesh(self._h, mesh_id)

    # -- instances -----------------------------------------------------------------
    @staticmethod
    def _xf(xform3x4):
        if xform3x4 is None:
            return None, None
        a = np.ascontiguousarray(xform3x4, dtype=np.float32).reshape(12)
        return a.ctypes.data_as(C.POINTER(C.c_float)), a

    def add_instance(self, mesh_id: int, xform3x4=None, material_id: int = 0) -> int:
        p, keep = self._xf(xform3x4)
        return L.lib().mrtAddInstance(self._h, mesh_id, p, material_id)

    def set_instance_transform(self, inst_id: int, xform3x4) -> None:
        p, keep = self._xf(xform3x4)
        L.lib().mrtSetInstanceTransform(self._h, inst_id, p)

    def remove_instance(self, inst_id: int) -> None:
        L.lib().mrtRemoveInstance(self._h, inst_id)

    # -- materials / textures -------------------------------------------------------
    @staticmethod
    def material_desc(base_color=(0.8, 0.8, 0.8), opacity=1.0, roughness=0.5, metallic=0.0,
                      transmission=0.0, ior=1.45, emission=(0, 0, 0), base_color_tex=0,
                      roughness_tex=0, metallic_tex=0, normal_tex=0, emission_tex=0):
        return L.mrtMaterialDesc(baseColor=_f3(base_color), opacity=opacity, roughness=roughness,
                                 metallic=metallic, transmission=transmission, ior=ior,
                                 emission=_f3(emission), baseColorTex=base_color_tex,
                                 roughnessTex=roughness_tex, metallicTex=metallic_tex,
                                 normalTex=normal_tex, emissionTex=emission_tex)

    def add_material(self, **kw) -> int:
        d = self.material_desc(**kw)
        return L.lib().mrtAddMaterial(self._h, C.byref(d))

    def update_material(self, mat_id: int, **kw) -> None:
        d = self.material_desc(**kw)
        L.lib().mrtUpdateMaterial(self._h, mat_id, C.byref(d))

    def remove_material(self, mat_id: int) -> None:
        L.lib().mrtRemoveMaterial(self._h, mat_id)

    def add_texture(self, pixels: np.ndarray, fmt: int = L.MRT_TEX_RGBA8_SRGB, mips: bool = False) -> int:
        px = np.ascontiguousarray(pixels)
        h, w = px.shape[:2]
        d = L.mrtTextureDesc(width=w, height=h, format=fmt, pixels=px.ctypes.data, generateMips=int(mips))
        return L.lib().mrtAddTexture(self._h, C.byref(d))

    def remove_texture(self, tex_id: int) -> None:
        L.lib().mrtRemoveTexture(self._h, tex_id)

    # -- lights / environment / camera -----------------------------------------------
    @staticmethod
    def light_desc(type=L.MRT_LIGHT_SUN, direction=(0, 1, 0), position=(0, 0, 0), radiance=(1, 1, 1),
                   angular_radius=0.00465, corner=(0, 0, 0), edge0=(1, 0, 0), edge1=(0, 0, 1), two_sided=False):
        return L.mrtLightDesc(type=type, direction=_f3(direction), position=_f3(position),
                              radiance=_f3(radiance), angularRadius=angular_radius, corner=_f3(corner),
                              edge0=_f3(edge0), edge1=_f3(edge1), twoSided=int(two_sided))

    def add_light(self, **kw) -> int:
        d = self.light_desc(**kw)
        return L.lib().mrtAddLight(self._h, C.byref(d))

    def update_light(self, light_id: int, **kw) -> None:
        d = self.light_desc(**kw)
        L.lib().mrtUpdateLight(self._h, light_id, C.byref(d))

    def remove_light(self, light_id: int) -> None:
        L.lib().mrtRemoveLight(self._h, light_id)

    def set_environment(self, rgba32f: np.ndarray | None, rotation: float = 0.0, intensity: float = 1.0,
                        fmt: int = L.MRT_TEX_RGBA32F) -> None:
        if rgba32f is None:
            L.lib().mrtSetEnvironment(self._h, None, rotation, intensity)
            return
        px = np.ascontiguousarray(rgba32f, dtype=np.float32 if fmt == L.MRT_TEX_RGBA32F else np.uint8)
        h, w = px.shape[:2]
        d = L.mrtTextureDesc(width=w, height=h, format=fmt, pixels=px.ctypes.data, generateMips=0)
        L.lib().mrtSetEnvironment(self._h, C.byref(d), rotation, intensity)

    def set_camera(self, position=(0, 0, 5), target=(0, 0, 0), up=(0, 1, 0), fov_y_deg=45.0) -> None:
        d = L.mrtCameraDesc(position=_f3(position), target=_f3(target), up=_f3(up), fovYDeg=fov_y_deg)
        L.lib().mrtSetCamera(self._h, C.byref(d))

    def set_camera_ex(self, projection=L.MRT_CAMERA_PERSPECTIVE, position=(0, 0, 5), forward=(0, 0, -1),
                      up=(0, 1, 0), left=-1.0, right=1.0, bottom=-1.0, top=1.0) -> None:
        d = L.mrtCameraDescEx(projection=projection, position=_f3(position), forward=_f3(forward),
                              up=_f3(up), left=left, right=right, bottom=bottom, top=top)
        L.lib().mrtSetCameraEx(self._h, C.byref(d))

    # -- render ----------------------------------------------------------------------
    def commit(self) -> dict:
        st = L.mrtCommitStats()
        L.check(L.lib().mrtCommitScene(self._h, C.byref(st)), "mrtCommitScene")
        return {"blasRebuilt": st.blasRebuilt, "tlasRebuilt": st.tlasRebuilt, "commitMs": st.commitMs}

    def set_render_settings(self, width=1280, height=720, max_bounces=8, spp_limit=1024, exposure_ev=0.0,
                            firefly_clamp=50.0, tonemap=L.MRT_TONEMAP_ACES, denoise=True) -> None:
        s = L.mrtRenderSettings(width=width, height=height, maxBounces=max_bounces, sppLimit=spp_limit,
                                exposureEV=exposure_ev, fireflyClamp=firefly_clamp, tonemap=tonemap,
                                denoise=int(denoise))
        L.check(L.lib().mrtSetRenderSettings(self._h, C.byref(s)), "mrtSetRenderSettings")
        self._size = (width, height)

    def render_frame(self) -> dict:
        info = L.mrtFrameInfo()
        L.check(L.lib().mrtRenderFrame(self._h, C.byref(info)), "mrtRenderFrame")
        return {"spp": info.spp, "converged": info.converged, "denoised": info.denoised,
                "frameMs": info.frameMs}

    def read_rgba32f(self, width: int, height: int) -> np.ndarray:
        out = np.empty((height, width, 4), dtype=np.float32)
        L.check(L.lib().mrtReadFramebuffer(self._h, L.MRT_PIXEL_RGBA32F_LINEAR, out.ctypes.data, out.nbytes),
                "mrtReadFramebuffer")
        return out

    def read_rgba8(self, width: int, height: int) -> np.ndarray:
        out = np.empty((height, width, 4), dtype=np.uint8)
        L.check(L.lib().mrtReadFramebuffer(self._h, L.MRT_PIXEL_RGBA8_SRGB, out.ctypes.data, out.nbytes),
                "mrtReadFramebuffer")
        return out

    def reset_accumulation(self) -> None:
        L.lib().mrtResetAccumulation(self._h)

    # -- extensions --------------------------------------------------------------------
    def set_path_options(self, seed=0, use_graph=True, extend=(0, 0), shadow=(0, 0), tail_paths=0,
                         packet_camera=True) -> None:
        """tail_paths: 0 = default, L.TAIL_DISABLED = never use the tail kernel.
        packet_camera: trace camera rays by wave packets (same hits either way)."""
        o = L.mrtxPathOptions(seed=seed, useGraph=int(use_graph), extendBlock=extend[0],
                              extendStack=extend[1], shadowBlock=shadow[0], shadowStack=shadow[1],
                              tailPaths=tail_paths, cameraPerLane=int(not packet_camera))
        L.check(L.lib().mrtxSetPathOptions(self._h, C.byref(o)), "mrtxSetPathOptions")

    def path_options(self) -> dict:
        o = L.mrtxPathOptions()
        L.check(L.lib().mrtxGetPathOptions(self._h, C.byref(o)), "mrtxGetPathOptions")
        return {"seed": o.seed, "use_graph": bool(o.useGraph), "extend": (o.extendBlock, o.extendStack),
                "shadow": (o.shadowBlock, o.shadowStack), "tail_paths": o.tailPaths,
                "packet_camera": not o.cameraPerLane}

    # -- path guiding (mini_raytrace_ext.h section 5) --
    def set_guiding(self, sampling: bool = True, training: bool = True, select_prob: float = 0.5,
                    learning_rate: float = 1e-3, batch_size: int = 65536, steps_per_frame: int = 4,
                    min_roughness: float = 0.2, seed: int = 1, position_scale: float = 1.0,
                    warmup_frames: int = 32, warmup_steps_per_frame: int = 16) -> None:
        """warmup_steps_per_frame = steps_per_frame: no warm-up (the network's
        first warmup_frames trained frames otherwise take that many steps)."""
        o = L.mrtxGuidingOptions(enableSampling=int(sampling), enableTraining=int(training),
                                 selectProb=select_prob, learningRate=learning_rate, batchSize=batch_size,
                                 stepsPerFrame=steps_per_frame, minRoughness=min_roughness, seed=seed,
                                 positionScale=position_scale, warmupFrames=warmup_frames,
                                 warmupStepsPerFrame=warmup_steps_per_frame)
        L.check(L.lib().mrtxSetGuiding(self._h, C.byref(o)), "mrtxSetGuiding")

    def guiding(self) -> dict:
        o = L.mrtxGuidingOptions()
        L.check(L.lib().mrtxGetGuiding(self._h, C.byref(o)), "mrtxGetGuiding")
        return {k: getattr(o, k) for k, _ in o._fields_ if k != "structSize"}

    def guiding_stats(self) -> dict:
        o = L.mrtxGuidingStats()
        L.check(L.lib().mrtxGetGuidingStats(self._h, C.byref(o)), "mrtxGetGuidingStats")
        return {k: getattr(o, k) for k, _ in o._fields_ if k != "structSize"}

    def guide_params(self) -> np.ndarray:
        out = np.empty(L.MRTX_GUIDE_PARAM_COUNT, np.float32)
        L.check(L.lib().mrtxGetGuideParams(self._h, out.ctypes.data, out.size), "mrtxGetGuideParams")
        return out

    def set_guide_params(self, params: np.ndarray) -> None:
        p = np.ascontiguousarray(params, np.float32)
        L.check(L.lib().mrtxSetGuideParams(self._h, p.ctypes.data, p.size), "mrtxSetGuideParams")

    def autotune_path(self) -> dict:
        """Chooses the closest-hit and shadow launch shapes (restarts accumulation)."""
        o = L.mrtxPathOptions()
        L.check(L.lib().mrtxAutotunePath(self._h, C.byref(o)), "mrtxAutotunePath")
        return {"extend": (o.extendBlock, o.extendStack), "shadow": (o.shadowBlock, o.shadowStack)}

    def set_profiling(self, on: bool) -> None:
        L.check(L.lib().mrtxSetProfiling(self._h, int(on)), "mrtxSetProfiling")

    def frame_stats(self) -> dict:
        s = L.mrtxFrameStats()
        L.check(L.lib().mrtxGetFrameStats(self._h, C.byref(s)), "mrtxGetFrameStats")
        names = ["generate", "extend", "shade", "shadow", "accumulate", "tail", "guide", "train"]
        return {"frameMs": s.frameMs, "gpuMs": dict(zip(names, list(s.gpuMs))),
                "rays": list(s.rays), "shadowRays": list(s.shadowRays), "status": s.status,
                "commitMs": s.commitMs, "commitUploadBytes": s.commitUploadBytes,
                "geometryRepacks": s.geometryRepacks, "meshesWritten": s.meshesWritten,
                "graphTailLevel": s.graphTailLevel}

    def export_scene(self) -> dict:
        """Packed device scene as numpy arrays (two calls: counts, then data)."""
        e = L.mrtxSceneExport()
        L.check(L.lib().mrtxExportScene(self._h, C.byref(e)), "mrtxExportScene")
        from .query import NODE_DTYPE
        out = {
            "nodes": np.zeros(e.nodeCount, dtype=NODE_DTYPE),
            "triangles": np.zeros((e.triCount, 4), dtype=np.uint32),
            "positions": np.zeros((e.vertexCount, 4), dtype=np.float32),
            "normals": np.zeros((e.vertexCount, 4), dtype=np.float32),
            "uvs": np.zeros((e.vertexCount, 2), dtype=np.float32),
            "instances": np.zeros((e.instanceCount, 36), dtype=np.float32),
            "materials": np.zeros((e.materialCount, 20), dtype=np.float32),
            "lights": np.zeros((e.lightCount, 20), dtype=np.float32),
            "envAlias": np.zeros((e.envWidth * e.envHeight, 2), dtype=np.uint32),
        }
        out["tangents"] = np.zeros((e.vertexCount, 4), dtype=np.float32)
        out["textureTable"] = np.zeros((L.MRTX_TEXTURE_SLOTS, 4), dtype=np.uint32)
        out["tex8"] = np.zeros((e.tex8Count, 4), dtype=np.uint8)
        out["tex32"] = np.zeros((e.tex32Count, 4), dtype=np.float32)
        for k, a in out.items():
            setattr(e, k, a.ctypes.data if a.size else None)
        L.check(L.lib().mrtxExportScene(self._h, C.byref(e)), "mrtxExportScene")
        out.update(tlasRoot=e.tlasRoot, blasNodeCount=e.blasNodeCount, envIntegral=e.envIntegral,
                   envWidth=e.envWidth, envHeight=e.envHeight)
        return out
