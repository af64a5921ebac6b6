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
