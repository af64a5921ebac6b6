"""Ray queries over one triangle set (mrtxQueryScene in mini_raytrace_ext.h).

Mirrors the reference's BVH build + closest-hit traversal pair
(core/src/accel/Bvh.cpp:137-156, core/shaders/intersect.glsl:48-86) as a
device-batched query: rays and hits live in device memory (torch tensors are
used here only as device allocations and for their stream handle).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L

NODE_DTYPE = np.dtype([("lo", "<f4", 3), ("leftOrPrim", "<u4"), ("hi", "<f4", 3), ("packed", "<u4")])


def _stream_handle(stream) -> int:
    if stream is None:
        import torch
        stream = torch.cuda.current_stream()
    return int(getattr(stream, "cuda_stream", stream) or 0)


class QueryScene:
    """BVH over triangles (positions (V,3) float32, indices (T,3) uint32)."""

    def __init__(self, positions: np.ndarray, indices: np.ndarray, gpu_index1: int = 0):
        lib = L.lib()
        pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
        idx = np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1, 3)
        handle = C.c_void_p()
        L.check(lib.mrtxCreateQueryScene(pos.ctypes.data, pos.shape[0], idx.ctypes.data,
                                         idx.shape[0], gpu_index1, C.byref(handle)),
                "mrtxCreateQueryScene")
        self._h = handle
        self.tri_count = idx.shape[0]

    @classmethod
    def from_soup(cls, verts9: np.ndarray, gpu_index1: int = 0) -> "QueryScene":
        v = np.ascontiguousarray(verts9, dtype=np.float32).reshape(-1, 3)
        idx = np.arange(v.shape[0], dtype=np.uint32).reshape(-1, 3)
        return cls(v, idx, gpu_index1)

    def close(self) -> None:
        if self._h:
            L.lib().mrtxDestroyQueryScene(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> dict:
        out = L.mrtxQuerySceneInfo()
        L.check(L.lib().mrtxGetQuerySceneInfo(self._h, C.byref(out)), "mrtxGetQuerySceneInfo")
        return {k: getattr(out, k) for k, _ in out._fields_ if k != "structSize"}

    def export_bvh(self) -> tuple[np.ndarray, np.ndarray]:
        info = self.info()
        nodes = np.zeros(info["nodeCount"], dtype=NODE_DTYPE)
        order = np.zeros(self.tri_count, dtype=np.uint32)
        L.check(L.lib().mrtxExportQueryBvh(self._h, nodes.ctypes.data, nodes.nbytes,
                                           order.ctypes.data if order.size else None, order.size),
                "mrtxExportQueryBvh")
        return nodes, order

    def set_config(self, block: int, stack: int, loop: int = 0) -> None:
        cfg = L.mrtxTraversalConfig(blockSize=block, stackDepth=stack, loopMode=loop)
        L.check(L.lib().mrtxSetTraversalConfig(self._h, C.byref(cfg)), "mrtxSetTraversalConfig")

    def config(self) -> tuple[int, int, int]:
        cfg = L.mrtxTraversalConfig()
        L.check(L.lib().mrtxGetTraversalConfig(self._h, C.byref(cfg)), "mrtxGetTraversalConfig")
        return cfg.blockSize, cfg.stackDepth, cfg.loopMode

    def trace_closest(self, rays, hits, stream=None) -> None:
        """rays: device tensor (n, 8) float32; hits: device tensor (n, 2) int32."""
        n = rays.shape[0]
        L.check(L.lib().mrtxTraceClosest(self._h, rays.data_ptr(), n, hits.data_ptr(),
                                         _stream_handle(stream)), "mrtxTraceClosest")

    def trace_any(self, rays, occluded, stream=None) -> None:
        n = rays.shape[0]
        L.check(L.lib().mrtxTraceAny(self._h, rays.data_ptr(), n, occluded.data_ptr(),
                                     _stream_handle(stream)), "mrtxTraceAny")

    def autotune(self, rays, hits) -> tuple[int, int, int]:
        cfg = L.mrtxTraversalConfig()
        L.check(L.lib().mrtxAutotuneTraversal(self._h, rays.data_ptr(), rays.shape[0],
                                              hits.data_ptr(), C.byref(cfg)),
                "mrtxAutotuneTraversal")
        return cfg.blockSize, cfg.stackDepth, cfg.loopMode

    def status(self) -> int:
        flags = C.c_uint32()
        L.check(L.lib().mrtxGetQueryStatus(self._h, C.byref(flags)), "mrtxGetQueryStatus")
        return flags.value
