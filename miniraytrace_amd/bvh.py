"""Host-side BVH build through the C ABI (mrtxBuildBvh)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .query import NODE_DTYPE


def build_bvh(bounds6: np.ndarray, max_leaf: int = 4) -> tuple[np.ndarray, np.ndarray]:
    """bounds6: (n, 6) float32 (lo xyz, hi xyz).  Returns (nodes, primOrder)."""
    b = np.ascontiguousarray(bounds6, dtype=np.float32).reshape(-1, 6)
    n = b.shape[0]
    cap = max(2 * n, 1)
    nodes = np.zeros(cap, dtype=NODE_DTYPE)
    order = np.zeros(max(n, 1), dtype=np.uint32)
    cnt = C.c_uint64()
    L.check(L.lib().mrtxBuildBvh(b.ctypes.data, n, max_leaf, nodes.ctypes.data, cap,
                                 order.ctypes.data, C.byref(cnt)), "mrtxBuildBvh")
    return nodes[: cnt.value].copy(), order[:n].copy()


def build_bvh_device(bounds6, max_leaf: int = 4, stream: int = 0, timing: dict | None = None):
    """mrtxBuildBvhDevice on a torch CUDA tensor of (n, 6) float32 bounds;
    returns (nodes as a host structured array, primOrder) after copying the
    device outputs back.  `timing`, when given, receives "ms": the wall time
    of the call itself (it returns with the outputs written)."""
    import time
    import torch
    b = bounds6.contiguous()
    n = b.shape[0]
    cap = max(2 * n - 1, 1)
    d_nodes = torch.empty((cap, 8), dtype=torch.int32, device=b.device)
    d_order = torch.empty(max(n, 1), dtype=torch.int32, device=b.device)
    cnt = C.c_uint64()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    L.check(L.lib().mrtxBuildBvhDevice(C.c_void_p(b.data_ptr()), n, max_leaf, C.c_void_p(d_nodes.data_ptr()), cap,
                                       C.c_void_p(d_order.data_ptr()), C.byref(cnt),
                                       C.c_void_p(stream) if stream else None), "mrtxBuildBvhDevice")
    if timing is not None:
        timing["ms"] = (time.perf_counter() - t0) * 1e3
    nodes = d_nodes[: cnt.value].cpu().numpy().view(NODE_DTYPE).reshape(-1)
    return nodes.copy(), d_order[:n].cpu().numpy().view(np.uint32).copy()
