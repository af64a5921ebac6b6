"""Seeded synthetic workloads (libmrt_scenes.so, host-only C++).

Random numbers come from std::mt19937_64 with floats u = (x >> 40) * 2^-24.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

_PATH = Path(__file__).resolve().parent / "libmrt_scenes.so"
_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not _PATH.exists():
            raise RuntimeError(f"{_PATH} is missing; build with `python -m miniraytrace_amd.build`")
        h = C.CDLL(str(_PATH))
        h.mrts_soup.argtypes = [C.c_uint32, C.c_uint64, C.c_void_p]
        h.mrts_soup_rays.argtypes = [C.c_uint64, C.c_uint64, C.c_void_p]
        _bind(h)
        _lib = h
    return _lib


def _bind(h):
    P, U32, U64, F = C.c_void_p, C.c_uint32, C.c_uint64, C.c_float
    h.mrts_terrain.argtypes = [U32, U32, F, F, F, U64, P, P]
    h.mrts_terrain_height.argtypes = [F, F, F, U64]
    h.mrts_terrain_height.restype = F
    h.mrts_icosphere.argtypes = [U32, P, P]
    h.mrts_icosphere.restype = U32
    h.mrts_sky.argtypes = [U32, U32, F, F, P]
    h.mrts_scatter_on_terrain.argtypes = [U32, F, F, F, U64, F, F, U32, U64, P, P]
    h.mrts_materials.argtypes = [U32, U64, P]
    h.mrts_room.argtypes = [U32, F, F, F, F, F, U64, P, P, P]
    h.mrts_room.restype = U32
    h.mrts_displaced_sphere.argtypes = [U32, F, U64, P, P]
    h.mrts_displaced_sphere.restype = U32
    h.mrts_random_instances.argtypes = [U32, F, F, F, U32, U64, P, P]


def terrain(nx: int = 1000, nz: int = 500, sx: float = 20.0, sz: float = 10.0, amplitude: float = 0.2,
            seed: int = 1) -> tuple[np.ndarray, np.ndarray]:
    """Perlin heightfield: ((nx+1)(nz+1), 3) positions, (2 nx nz, 3) indices."""
    pos = np.empty(((nx + 1) * (nz + 1), 3), np.float32)
    idx = np.empty((nx * nz * 2, 3), np.uint32)
    lib().mrts_terrain(nx, nz, sx, sz, amplitude, seed, pos.ctypes.data, idx.ctypes.data)
    return pos, idx


def icosphere(level: int = 3) -> tuple[np.ndarray, np.ndarray]:
    nv, nf = 10 * 4 ** level + 2, 20 * 4 ** level
    pos = np.empty((nv, 3), np.float32)
    idx = np.empty((nf, 3), np.uint32)
    got = lib().mrts_icosphere(level, pos.ctypes.data, idx.ctypes.data)
    assert got == nv
    return pos, idx


def sky(width: int = 2048, height: int = 1024, sun_radius: float = np.radians(0.25),
        sun_scale: float = 5e4) -> np.ndarray:
    """Equirect RGBA32F sky (height, width, 4), sun disc of 0.5 deg diameter."""
    out = np.empty((height, width, 4), np.float32)
    lib().mrts_sky(width, height, float(sun_radius), float(sun_scale), out.ctypes.data)
    return out


def scatter_on_terrain(count, sx, sz, amplitude, terrain_seed, rmin, rmax, nmat, seed):
    xf = np.empty((count, 12), np.float32)
    mats = np.empty(count, np.uint32)
    lib().mrts_scatter_on_terrain(count, sx, sz, amplitude, terrain_seed, rmin, rmax, nmat, seed,
                                  xf.ctypes.data, mats.ctypes.data)
    return xf, mats


def materials(count: int = 3, seed: int = 1) -> np.ndarray:
    out = np.empty((count, 5), np.float32)
    lib().mrts_materials(count, seed, out.ctypes.data)
    return out


CONFIG2 = dict(nx=1000, nz=500, sx=20.0, sz=10.0, amplitude=0.2, spheres=200, sphere_level=3,
               rmin=0.15, rmax=0.45, sky=(2048, 1024), camera=dict(position=(0.0, 2.2, 7.5),
                                                                  target=(0.0, 0.0, 0.0), fov_y_deg=45.0))


def load_terrain_scene(engine, cfg: dict = CONFIG2, seed: int = 1, sky_img: np.ndarray | None = None) -> dict:
    """Terrain + instanced icospheres + GGX materials + sky into an Engine.

    Geometry: one terrain mesh (nx x nz quads) and one icosphere mesh
    instanced `spheres` times; 3 materials; the terrain uses material 0."""
    tpos, tidx = terrain(cfg["nx"], cfg["nz"], cfg["sx"], cfg["sz"], cfg["amplitude"], seed)
    spos, sidx = icosphere(cfg["sphere_level"])
    mats = materials(3, seed + 100)
    mat_ids = [engine.add_material(base_color=tuple(m[:3]), roughness=float(m[3]), metallic=float(m[4]))
               for m in mats]
    tm = engine.add_mesh(tpos, tidx)
    sm = engine.add_mesh(spos, sidx, normals=spos)
    inst = [engine.add_instance(tm, None, mat_ids[0])]
    xf, which = scatter_on_terrain(cfg["spheres"], cfg["sx"], cfg["sz"], cfg["amplitude"], seed,
                                   cfg["rmin"], cfg["rmax"], 3, seed + 200)
    for i in range(cfg["spheres"]):
        inst.append(engine.add_instance(sm, xf[i], mat_ids[int(which[i])]))
    if sky_img is None:
        sky_img = sky(*cfg["sky"])
    engine.set_environment(sky_img, 0.0, 1.0)
    engine.set_camera(**cfg["camera"])
    return {"triangles": tidx.shape[0] + cfg["spheres"] * sidx.shape[0], "instances": cfg["spheres"] + 1,
            "materials": mats, "camera": dict(cfg["camera"]), "terrain_mesh": tm, "sphere_mesh": sm,
            "instance_ids": inst, "sphere_xforms": xf, "material_ids": mat_ids}


def terrain_config(triangles: int) -> dict:
    """CONFIG2-style terrain-plus-spheres scene of about `triangles`
    triangles: spheres scale with the budget (200 at 1.256M, at least 1),
    the 2:1 terrain grid takes the rest."""
    cfg = dict(CONFIG2)
    spheres = int(min(200, max(1, round(200 * triangles / 1_256_000))))
    terrain_tris = max(2 * 4 * 2, triangles - spheres * 1280)
    nz = max(2, int(round((terrain_tris / 4) ** 0.5)))
    cfg.update(nx=2 * nz, nz=nz, spheres=spheres)
    return cfg


def batch_requests(count: int = 64, seed: int = 1000) -> list[dict]:
    """Seeded request list of the batched workload: resolution log-uniform
    from 256x256 to 1920x1080 (one draw scales both axes), scene size
    log-uniform in [10k, 2M] triangles, spp uniform in 1..16; request k uses
    scene seed seed + k."""
    import math
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        t = rng.random()
        w, h = int(round(256 * (1920 / 256) ** t)), int(round(256 * (1080 / 256) ** t))
        tris = int(round(math.exp(math.log(10_000) + rng.random() * (math.log(2_000_000) - math.log(10_000)))))
        out.append({"width": w, "height": h, "triangles": tris, "spp": int(rng.integers(1, 17)),
                    "seed": seed + k})
    return out


def room(n: int = 204, half=(2.0, 3.0, 2.0), window_frac: float = 0.05, jitter: float = 1e-3,
         seed: int = 1) -> tuple[np.ndarray, np.ndarray]:
    """Closed room [-hx,hx] x [0,hy] x [-hz,hz] of 6 n x n-quad walls with a
    window (windowFrac of the +z wall) cut out; interior vertices jittered."""
    hx, hy, hz = half
    pos = np.empty((6 * (n + 1) ** 2, 3), np.float32)
    idx = np.empty((6 * n * n * 2, 3), np.uint32)
    nv = C.c_uint32()
    nt = lib().mrts_room(n, hx, hy, hz, window_frac, jitter, seed, pos.ctypes.data, idx.ctypes.data, C.byref(nv))
    return pos[: nv.value], idx[:nt]


CONFIG4 = dict(n=204, half=(2.0, 3.0, 2.0), window_frac=0.05, jitter=1e-3, albedo=0.7, width=1280, height=720,
               camera=dict(position=(-1.6, 1.6, -1.7), target=(0.6, 0.7, 1.2), fov_y_deg=60.0))


def load_room_scene(engine, cfg: dict = CONFIG4, seed: int = 1, sky_img: np.ndarray | None = None) -> dict:
    """Diffuse room lit only by the sky through its window (config 4)."""
    pos, idx = room(cfg["n"], cfg["half"], cfg["window_frac"], cfg["jitter"], seed)
    mat = engine.add_material(base_color=(cfg["albedo"],) * 3, roughness=1.0, metallic=0.0)
    m = engine.add_mesh(pos, idx)
    engine.add_instance(m, None, mat)
    if sky_img is None:
        sky_img = sky()
    engine.set_environment(sky_img, 0.0, 1.0)
    engine.set_camera(**cfg["camera"])
    return {"triangles": idx.shape[0], "instances": 1, "camera": dict(cfg["camera"])}


def displaced_sphere(n: int = 45, amplitude: float = 0.15, seed: int = 1) -> tuple[np.ndarray, np.ndarray]:
    """Geodesic sphere of 20 n^2 triangles (n = 45: 40,500), radially
    displaced by seeded waves: (10 n^2 + 2, 3) positions, (20 n^2, 3) indices."""
    pos = np.empty((10 * n * n + 2, 3), np.float32)
    idx = np.empty((20 * n * n, 3), np.uint32)
    nv = lib().mrts_displaced_sphere(n, amplitude, seed, pos.ctypes.data, idx.ctypes.data)
    return pos[:nv], idx


def random_instances(count: int, box: float, smin: float, smax: float, nmat: int,
                     seed: int) -> tuple[np.ndarray, np.ndarray]:
    """(count, 12) row-major 3x4 transforms uniformly random in a box of edge
    `box` (rotation, scale in [smin, smax], translation) and material ids."""
    xf = np.empty((count, 12), np.float32)
    mats = np.empty(count, np.uint32)
    lib().mrts_random_instances(count, box, smin, smax, nmat, seed, xf.ctypes.data, mats.ctypes.data)
    return xf, mats


# Guided and batched together at scale: 256 instances of a 40,500-triangle
# displaced sphere (10.4M triangles instanced) in a 20^3 box, 3 GGX materials
# as in the terrain scene; requests alternate between this field and the
# window-lit room, half at 3840x2160 and half at 1920x1080, max depth 12.
CONFIG5 = dict(instances=256, sphere_n=45, box=20.0, smin=0.5, smax=1.5, requests=16, max_depth=12,
               resolutions=((3840, 2160), (1920, 1080)),
               camera=dict(position=(0.0, 4.0, 26.0), target=(0.0, 0.0, 0.0), fov_y_deg=50.0))


def load_sphere_field(engine, cfg: dict = CONFIG5, seed: int = 1, sky_img: np.ndarray | None = None) -> dict:
    """The instanced displaced-sphere field of CONFIG5 into an Engine."""
    pos, idx = displaced_sphere(cfg["sphere_n"], 0.15, seed)
    mats = materials(3, seed + 100)
    mat_ids = [engine.add_material(base_color=tuple(m[:3]), roughness=float(m[3]), metallic=float(m[4]))
               for m in mats]
    mesh = engine.add_mesh(pos, idx)
    xf, which = random_instances(cfg["instances"], cfg["box"], cfg["smin"], cfg["smax"], 3, seed + 7)
    for k in range(cfg["instances"]):
        engine.add_instance(mesh, xf[k], mat_ids[int(which[k])])
    if sky_img is None:
        sky_img = sky()
    engine.set_environment(sky_img, 0.0, 1.0)
    engine.set_camera(**cfg["camera"])
    return {"triangles": cfg["instances"] * idx.shape[0], "unique_triangles": idx.shape[0],
            "instances": cfg["instances"], "camera": dict(cfg["camera"])}


def soup(count: int = 10_000, seed: int = 1) -> np.ndarray:
    """Triangle soup in the unit cube: (count, 9) float32 (v0, v1, v2)."""
    out = np.empty((count, 9), dtype=np.float32)
    lib().mrts_soup(count, seed, out.ctypes.data)
    return out


def soup_rays(n: int = 1_000_000, seed: int = 2) -> np.ndarray:
    """Rays from a radius-2 sphere around (0.5,0.5,0.5) towards points in the
    unit cube: (n, 8) float32 (ox oy oz tmin dx dy dz tmax)."""
    out = np.empty((n, 8), dtype=np.float32)
    lib().mrts_soup_rays(n, seed, out.ctypes.data)
    return out
