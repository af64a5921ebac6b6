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


# ---- path tracer (core/shaders/pathtrace.comp restated in FP64) -------------
class _PtScene(C.Structure):
    _fields_ = [("nodes", C.c_void_p), ("tlasRoot", C.c_uint32), ("instanceCount", C.c_uint32),
                ("triangles", C.c_void_p), ("positions", C.c_void_p), ("normals", C.c_void_p),
                ("instances", C.c_void_p), ("materials", C.c_void_p), ("lights", C.c_void_p),
                ("lightCount", C.c_uint32), ("env", C.c_void_p), ("envW", C.c_uint32), ("envH", C.c_uint32),
                ("envRotation", C.c_float), ("envIntensity", C.c_float), ("envAlias", C.c_void_p)]


class _PtCamera(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("right", C.c_float * 3), ("up", C.c_float * 3),
                ("forward", C.c_float * 3), ("frustum", C.c_float * 4), ("ortho", C.c_int32)]


class _PtParams(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("x0", C.c_uint32), ("y0", C.c_uint32),
                ("cw", C.c_uint32), ("ch", C.c_uint32), ("frame0", C.c_uint32), ("spp", C.c_uint32),
                ("maxBounces", C.c_uint32), ("fireflyClamp", C.c_float), ("seedMix", C.c_uint32),
                ("threads", C.c_int32), ("countFetches", C.c_int32)]


def _bind_pt(h):
    h.or_pt_render.argtypes = [C.POINTER(_PtScene), C.POINTER(_PtCamera), C.POINTER(_PtParams),
                               C.c_void_p, C.c_void_p]
    h.or_pt_render.restype = None
    h.or_camera_lookat.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_float, C.c_uint32,
                                   C.c_uint32, C.POINTER(_PtCamera)]
    h.or_env_cdf_probs.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p]
    h.or_env_cdf_probs.restype = C.c_double


def camera_lookat(position, target, up, fov_y_deg, width, height) -> _PtCamera:
    _bind_pt(lib())
    cam = _PtCamera()
    a = [_f32(x) for x in (position, target, up)]
    lib().or_camera_lookat(a[0].ctypes.data, a[1].ctypes.data, a[2].ctypes.data, fov_y_deg, width,
                           height, C.byref(cam))
    return cam


def env_cdf_probs(env_rgba: np.ndarray) -> tuple[np.ndarray, float]:
    """Per-texel probability of the reference CDF sampler and the solid-angle
    luminance integral (Renderer.cpp:620-652)."""
    _bind_pt(lib())
    e = _f32(env_rgba)
    h, w = e.shape[:2]
    prob = np.zeros(h * w, np.float64)
    integral = lib().or_env_cdf_probs(e.ctypes.data, w, h, prob.ctypes.data)
    return prob.reshape(h, w), float(integral)


def pt_render(exported: dict, env: np.ndarray | None, camera: _PtCamera, width: int, height: int,
              crop=(0, 0, None, None), frame0: int = 0, spp: int = 1, max_bounces: int = 8,
              firefly_clamp: float = 0.0, env_rotation: float = 0.0, env_intensity: float = 1.0,
              use_alias: bool = False, seed_mix: int = 0, threads: int = 0, count_fetches: bool = False):
    """Renders a crop with the FP64 restatement.  `exported` is the packed
    scene from Engine.export_scene().  Returns (sum of radiance over the spp
    frames, shape (ch, cw, 3), counts[6] or None)."""
    _bind_pt(lib())
    x0, y0, cw, ch = crop
    cw = width - x0 if cw is None else cw
    ch = height - y0 if ch is None else ch
    keep = {k: np.ascontiguousarray(v) for k, v in exported.items() if isinstance(v, np.ndarray)}
    sc = _PtScene()
    sc.nodes = keep["nodes"].ctypes.data
    sc.tlasRoot = exported["tlasRoot"]
    sc.instanceCount = keep["instances"].shape[0]
    sc.triangles = keep["triangles"].ctypes.data
    sc.positions = keep["positions"].ctypes.data
    sc.normals = keep["normals"].ctypes.data
    sc.instances = keep["instances"].ctypes.data if keep["instances"].size else None
    sc.materials = keep["materials"].ctypes.data
    sc.lights = keep["lights"].ctypes.data if keep["lights"].size else None
    sc.lightCount = keep["lights"].shape[0]
    env_arr = None
    if env is not None:
        env_arr = _f32(env)
        sc.env = env_arr.ctypes.data
        sc.envH, sc.envW = env_arr.shape[:2]
        sc.envRotation = env_rotation
        sc.envIntensity = env_intensity
        if use_alias:
            sc.envAlias = keep["envAlias"].ctypes.data
    p = _PtParams(width=width, height=height, x0=x0, y0=y0, cw=cw, ch=ch, frame0=frame0, spp=spp,
                  maxBounces=max_bounces, fireflyClamp=firefly_clamp, seedMix=seed_mix, threads=threads,
                  countFetches=int(count_fetches))
    out = np.zeros((ch, cw, 3), np.float64)
    counts = np.zeros(6, np.uint64)
    lib().or_pt_render(C.byref(sc), C.byref(camera), C.byref(p), out.ctypes.data, counts.ctypes.data)
    return out, (counts if count_fetches else None)
