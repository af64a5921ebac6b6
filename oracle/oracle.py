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
                ("envRotation", C.c_float), ("envIntensity", C.c_float), ("envAlias", C.c_void_p),
                ("uvs", C.c_void_p), ("tangents", C.c_void_p), ("textures", C.c_void_p),
                ("textureCount", C.c_uint32)]


class _Texture(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("format", C.c_uint32), ("pixels", C.c_void_p)]


class _BsdfParams(C.Structure):
    _fields_ = [("base", C.c_double * 3), ("roughness", C.c_double), ("metallic", C.c_double),
                ("transmission", C.c_double), ("ior", C.c_double), ("entering", C.c_int32)]


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
    h.or_bsdf_eval.argtypes = [C.POINTER(_BsdfParams), C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p]
    h.or_bsdf_eval.restype = None
    h.or_bsdf_sample.argtypes = [C.POINTER(_BsdfParams), C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p,
                                 C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    h.or_bsdf_sample.restype = None
    h.or_light_sample.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]
    h.or_light_sample.restype = None
    h.or_texture_sample.argtypes = [C.POINTER(_Texture), C.c_double, C.c_double, C.c_void_p]
    h.or_texture_sample.restype = None


def _bsdf(base, roughness, metallic=0.0, transmission=0.0, ior=1.5, entering=True) -> _BsdfParams:
    m = _BsdfParams()
    m.base[:] = [float(x) for x in base]
    m.roughness, m.metallic, m.transmission, m.ior = roughness, metallic, transmission, ior
    m.entering = int(bool(entering))
    return m


def bsdf_eval(material: dict, wo, wi):
    """(f (N, 3), combined lobe pdf (N,)) of bsdf.glsl:100-144 in the local
    frame for one wo and wi (N, 3).  material: base, roughness, metallic,
    transmission, ior, entering."""
    _bind_pt(lib())
    m = _bsdf(**material)
    wo = np.ascontiguousarray(wo, np.float64)
    wi = np.ascontiguousarray(np.atleast_2d(wi), np.float64)
    n = wi.shape[0]
    f = np.zeros((n, 3))
    pdf = np.zeros(n)
    lib().or_bsdf_eval(C.byref(m), wo.ctypes.data, n, wi.ctypes.data, f.ctypes.data, pdf.ctypes.data)
    return f, pdf


def bsdf_sample(material: dict, wo, u: np.ndarray):
    """bsdfSample (bsdf.glsl:147-192) for each row of u (pick, u0, u1,
    Fresnel choice).  Returns (wi (N, 3), f (N, 3), pdf (N,), branch (N,),
    valid (N,) bool)."""
    _bind_pt(lib())
    m = _bsdf(**material)
    wo = np.ascontiguousarray(wo, np.float64)
    u = np.ascontiguousarray(u, np.float64)
    n = u.shape[0]
    wi = np.zeros((n, 3))
    f = np.zeros((n, 3))
    pdf = np.zeros(n)
    branch = np.zeros(n, np.int32)
    valid = np.zeros(n, np.uint8)
    lib().or_bsdf_sample(C.byref(m), wo.ctypes.data, n, u.ctypes.data, wi.ctypes.data, f.ctypes.data,
                         pdf.ctypes.data, branch.ctypes.data, valid.ctypes.data)
    return wi, f, pdf, branch, valid.astype(bool)


def light_sample(light20, pick_count: int, p, u):
    """Analytic light sample (lights.glsl:70-134 after the pick): dict with
    wi, radiance, pdf, dist, delta, valid."""
    _bind_pt(lib())
    l = _f32(light20)
    p = np.ascontiguousarray(p, np.float64)
    u = np.ascontiguousarray(u, np.float64)
    out = np.zeros(10)
    lib().or_light_sample(l.ctypes.data, pick_count, p.ctypes.data, u.ctypes.data, out.ctypes.data)
    return {"wi": out[0:3], "radiance": out[3:6], "pdf": out[6], "dist": out[7], "delta": bool(out[8]),
            "valid": bool(out[9])}


def _texture_struct(tex: dict | None):
    t = _Texture()
    if tex is None:
        return t, None
    px = np.ascontiguousarray(tex["pixels"], np.float32 if tex["format"] == 2 else np.uint8)
    t.height, t.width = px.shape[:2]
    t.format = tex["format"]
    t.pixels = px.ctypes.data
    return t, px


def texture_sample(tex: dict, u: float, v: float) -> np.ndarray:
    """Level-0 bilinear lookup (REPEAT) decoded to linear RGBA.  tex: pixels
    (H, W, 4) uint8 or float32, format 0 sRGB8 / 1 UNORM8 / 2 RGBA32F."""
    _bind_pt(lib())
    t, keep = _texture_struct(tex)
    out = np.zeros(4)
    lib().or_texture_sample(C.byref(t), float(u), float(v), out.ctypes.data)
    del keep
    return out


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
    """Renders a crop with the FP64 restatement.  `exported` is a packed
    scene (tests/scene_pack.py, or Engine.export_scene()); optional keys
    "uvs" (V, 2), "tangents" (V, 4) and "textures" (list of slot dicts for
    texture_sample, None = empty slot).  Returns (sum of radiance over the
    spp frames, shape (ch, cw, 3), counts[6] or None)."""
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
    for k in ("uvs", "tangents"):
        if k in keep:
            keep[k] = np.ascontiguousarray(keep[k], np.float32)
            setattr(sc, k, keep[k].ctypes.data)
    texs = exported.get("textures") or []
    tex_arr = (_Texture * max(len(texs), 1))()
    tex_keep = []
    for i, t in enumerate(texs):
        tex_arr[i], px = _texture_struct(t)
        tex_keep.append(px)
    if texs:
        sc.textures = C.cast(tex_arr, C.c_void_p)
        sc.textureCount = len(texs)
    p = _PtParams(width=width, height=height, x0=x0, y0=y0, cw=cw, ch=ch, frame0=frame0, spp=spp,
                  maxBounces=max_bounces, fireflyClamp=firefly_clamp, seedMix=seed_mix, threads=threads,
                  countFetches=int(count_fetches))
    out = np.zeros((ch, cw, 3), np.float64)
    counts = np.zeros(6, np.uint64)
    lib().or_pt_render(C.byref(sc), C.byref(camera), C.byref(p), out.ctypes.data, counts.ctypes.data)
    return out, (counts if count_fetches else None)
