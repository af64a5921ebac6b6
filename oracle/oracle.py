"""ctypes wrapper of oracle/_build/liboracle.so.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py, and only as the checker or the timed CPU
baseline.  The product package never imports this module.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"
NODE_DTYPE = np.dtype([("lo", "<f4", 3), ("leftOrPrim", "<u4"), ("hi", "<f4", 3), ("packed", "<u4")])

_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        h = C.CDLL(str(LIB_PATH))
        P, U32, U64, F, I = C.c_void_p, C.c_uint32, C.c_uint64, C.c_float, C.c_int
        sig = {
            "or_pcg4d": (None, [P]),
            "or_rng1": (F, [P]),
            "or_rng2": (None, [P, P]),
            "or_bvh_build": (C.c_int64, [P, U32, U32, P, U64, P]),
            "or_tri_bounds": (None, [P, U32, P]),
            "or_isect_mt": (F, [P, P, P, P, P, F]),
            "or_isect_wt": (I, [P, P, P, P, P, F, F, P, P, P]),
            "or_brute_closest": (None, [P, U32, P, U64, P, I]),
            "or_bvh_closest": (None, [P, P, P, P, U64, P, I]),
            "or_ref_traverse": (None, [P, P, P, P, U64, P, I]),
            "or_glsl_traverse": (None, [P, P, P, P, U64, P, P, P, I]),
            "or_brute_any": (None, [P, U32, P, U64, P, I]),
            "or_hardware_threads": (I, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


# ---- rng (core/shaders/random.glsl:6-31) -------------------------------------
def pcg4d(v) -> np.ndarray:
    a = np.array(v, dtype=np.uint32).copy()
    lib().or_pcg4d(a.ctypes.data)
    return a


def rng_draws(px: int, py: int, frame: int, kinds: str) -> list:
    """Draw a sequence from state (px, py, frame, 0): kinds is a string of
    '1' (rng1) and '2' (rng2)."""
    s = np.array([px, py, frame, 0], dtype=np.uint32)
    out = []
    for k in kinds:
        if k == "1":
            out.append(float(lib().or_rng1(s.ctypes.data)))
        else:
            o = np.zeros(2, dtype=np.float32)
            lib().or_rng2(s.ctypes.data, o.ctypes.data)
            out.append((float(o[0]), float(o[1])))
    return out


# ---- bvh (core/src/accel/Bvh.cpp:137-156) ------------------------------------
def tri_bounds(verts9) -> np.ndarray:
    v = _f32(verts9).reshape(-1, 9)
    out = np.empty((v.shape[0], 6), dtype=np.float32)
    lib().or_tri_bounds(v.ctypes.data, v.shape[0], out.ctypes.data)
    return out


def bvh_build(bounds6, max_leaf: int = 4) -> tuple[np.ndarray, np.ndarray]:
    b = _f32(bounds6).reshape(-1, 6)
    n = b.shape[0]
    cap = max(2 * n, 1)
    nodes = np.zeros(cap, dtype=NODE_DTYPE)
    order = np.zeros(max(n, 1), dtype=np.uint32)
    cnt = lib().or_bvh_build(b.ctypes.data, n, max_leaf, nodes.ctypes.data, cap, order.ctypes.data)
    if cnt < 0:
        raise RuntimeError("node capacity exceeded")
    return nodes[:cnt].copy(), order[:n].copy()


# ---- intersection -----------------------------------------------------------
def isect_mt(v0, v1, v2, o, d, tmax=1e30) -> float:
    a = [_f32(x) for x in (v0, v1, v2, o, d)]
    return float(lib().or_isect_mt(*(x.ctypes.data for x in a), tmax))


def isect_wt(v0, v1, v2, o, d, tmin=1e-4, tlimit=float("inf")):
    """Returns (t, b1, b2) or None."""
    a = [_f32(x) for x in (v0, v1, v2, o, d)]
    t, b1, b2 = C.c_float(), C.c_float(), C.c_float()
    ok = lib().or_isect_wt(*(x.ctypes.data for x in a), tmin, tlimit,
                           C.byref(t), C.byref(b1), C.byref(b2))
    return (t.value, b1.value, b2.value) if ok else None


# ---- ray queries ------------------------------------------------------------
def _hits(n: int) -> np.ndarray:
    return np.zeros((n, 2), dtype=np.uint32)


def brute_closest(verts9, rays, threads: int = 0) -> np.ndarray:
    v = _f32(verts9).reshape(-1, 9)
    r = _f32(rays).reshape(-1, 8)
    h = _hits(r.shape[0])
    lib().or_brute_closest(v.ctypes.data, v.shape[0], r.ctypes.data, r.shape[0], h.ctypes.data, threads)
    return h


def bvh_closest(nodes, order, verts9, rays, threads: int = 0) -> np.ndarray:
    v, r = _f32(verts9).reshape(-1, 9), _f32(rays).reshape(-1, 8)
    nd = np.ascontiguousarray(nodes)
    od = np.ascontiguousarray(order, dtype=np.uint32)
    h = _hits(r.shape[0])
    lib().or_bvh_closest(nd.ctypes.data, od.ctypes.data, v.ctypes.data, r.ctypes.data,
                         r.shape[0], h.ctypes.data, threads)
    return h


def ref_traverse(nodes, order, verts9, rays, threads: int = 0) -> np.ndarray:
    v, r = _f32(verts9).reshape(-1, 9), _f32(rays).reshape(-1, 8)
    nd = np.ascontiguousarray(nodes)
    od = np.ascontiguousarray(order, dtype=np.uint32)
    h = _hits(r.shape[0])
    lib().or_ref_traverse(nd.ctypes.data, od.ctypes.data, v.ctypes.data, r.ctypes.data,
                          r.shape[0], h.ctypes.data, threads)
    return h


def glsl_traverse(nodes, order, verts9, rays, threads: int = 0, want_hits=True, want_counts=False):
    """Reference GLSL traversal order; returns (hits|None, counts|None, (nodes, tris))."""
    v, r = _f32(verts9).reshape(-1, 9), _f32(rays).reshape(-1, 8)
    nd = np.ascontiguousarray(nodes)
    od = np.ascontiguousarray(order, dtype=np.uint32)
    n = r.shape[0]
    h = _hits(n) if want_hits else None
    c = np.zeros((n, 2), dtype=np.uint32) if want_counts else None
    sums = np.zeros(2, dtype=np.uint64)
    lib().or_glsl_traverse(nd.ctypes.data, od.ctypes.data, v.ctypes.data, r.ctypes.data, n,
                           h.ctypes.data if h is not None else None,
                           c.ctypes.data if c is not None else None, sums.ctypes.data, threads)
    return h, c, (int(sums[0]), int(sums[1]))


def brute_any(verts9, rays, threads: int = 0) -> np.ndarray:
    v, r = _f32(verts9).reshape(-1, 9), _f32(rays).reshape(-1, 8)
    out = np.zeros(r.shape[0], dtype=np.uint8)
    lib().or_brute_any(v.ctypes.data, v.shape[0], r.ctypes.data, r.shape[0], out.ctypes.data, threads)
    return out


def hardware_threads() -> int:
    return int(lib().or_hardware_threads())
