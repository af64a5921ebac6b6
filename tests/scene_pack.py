"""Independent CPU packer of the scene the path tracer consumes, for the
oracle path tracer and for checking the engine's own packing.

`SceneDesc` records what a caller passes through the C ABI (mrtAddMesh,
mrtAddInstance, mrtAddMaterial, mrtAddTexture, mrtAddLight,
mrtSetEnvironment and their update/remove calls) and

* `load(engine)` replays those calls on a libmrt engine (the ids the engine
  returns must be the ones this description assigned);
* `pack()` restates the reference's packing in numpy, without libmrt:
  - mesh attributes, Scene.cpp:13-92: positions copied; area-weighted
    normals when none are given and tangents when uvs are given, both in
    float32 with glm's operation order, so the results are bit-exact;
  - geometry, Renderer.cpp:333-446: meshes in id order, each mesh's BLAS
    built by the oracle's binned-SAH builder (oracle/oracle.cpp, which is
    pinned against the reference rules), triangles rewritten in leaf order
    with global vertex indices, nodes rebased (inner + node base, leaf +
    triangle base), missing uvs 0 and tangents (1, 0, 0, 1);
  - instances, Renderer.cpp:448-508: instances in id order, world boxes of
    the BLAS root boxes (8 corners, MathTypes.hpp:38-48), TLAS with leaf size
    1, instances in TLAS leaf order, worldToObject = inverse(objectToWorld)
    in float64 rounded to float32, material id -> packed index (0 = default);
  - materials, Renderer.cpp:510-550: index 0 default grey, then materials in
    id order with the reference's clamps and texture id -> slot (-1 = none);
  - lights, Renderer.cpp:552-603: environment first when present, then
    lights in id order in the GpuLight encoding;
  - textures, Scene.cpp:159-204: slots assigned in add order, freed slots
    reused last-freed first, a removed slot empty.
The one packed input not restated here is the environment's alias table:
the oracle path tracer either builds the reference CDFs itself or takes the
engine's alias table, whose texel distribution tests/test_pt_gpu.py pins
against the reference CDF.
"""
from __future__ import annotations

import numpy as np

from oracle import oracle as orc

F = np.float32
MAX_TEXTURE_SLOTS = 64  # the reference's kMaxTextures

# mini_raytrace.h enums
LIGHT_SUN, LIGHT_POINT, LIGHT_RECT = 0, 1, 2
TEX_SRGB8, TEX_UNORM8, TEX_RGBA32F = 0, 1, 2
# GpuLight types (GpuTypes.hpp:72-83)
GPU_ENV, GPU_SUN, GPU_POINT, GPU_RECT = 0, 1, 2, 3


def _f(x):
    return np.asarray(x, F)


def _dot(a, b):
    """glm::dot of float32 vec3 rows: (x + y) + z."""
    return (a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1]) + a[..., 2] * b[..., 2]


def _cross(a, b):
    """glm::cross(a, b) in float32."""
    return np.stack([a[..., 1] * b[..., 2] - b[..., 1] * a[..., 2],
                     a[..., 2] * b[..., 0] - b[..., 2] * a[..., 0],
                     a[..., 0] * b[..., 1] - b[..., 0] * a[..., 1]], -1)


def _normalize(v):
    """glm::normalize: v * (1 / sqrt(dot(v, v))) in float32."""
    return v * (F(1) / np.sqrt(_dot(v, v)))[..., None]


def _accumulate(n_vertices, idx, vals):
    """acc[idx[t, k]] += vals[t] in triangle order, float32."""
    acc = np.zeros((n_vertices, 3), F)
    flat = idx.reshape(-1)
    rep = np.repeat(vals, 3, axis=0)
    for c in range(3):
        col = acc[:, c].copy()
        np.add.at(col, flat, rep[:, c])
        acc[:, c] = col
    return acc


def compute_normals(pos, idx):
    """Scene.cpp:13-27: area-weighted vertex normals, (0, 1, 0) when the
    accumulated normal is shorter than 1e-12."""
    p0, p1, p2 = pos[idx[:, 0]], pos[idx[:, 1]], pos[idx[:, 2]]
    acc = _accumulate(pos.shape[0], idx, _cross(p1 - p0, p2 - p0))
    length = np.sqrt(_dot(acc, acc))
    ok = length > F(1e-12)
    out = np.tile(_f([0, 1, 0]), (pos.shape[0], 1))
    out[ok] = _normalize(acc[ok])
    return out


def compute_tangents(pos, idx, nrm, uv):
    """Scene.cpp:31-60: per-face tangents accumulated, Gram-Schmidt against
    the normal, handedness in w."""
    i0, i1, i2 = idx[:, 0], idx[:, 1], idx[:, 2]
    e1, e2 = pos[i1] - pos[i0], pos[i2] - pos[i0]
    d1, d2 = uv[i1] - uv[i0], uv[i2] - uv[i0]
    det = d1[:, 0] * d2[:, 1] - d2[:, 0] * d1[:, 1]
    keep = np.abs(det) >= F(1e-12)
    r = np.zeros_like(det)
    r[keep] = F(1) / det[keep]
    T = (e1 * d2[:, 1:2] - e2 * d1[:, 1:2]) * r[:, None]
    B = (e2 * d1[:, 0:1] - e1 * d2[:, 0:1]) * r[:, None]
    T[~keep] = 0
    B[~keep] = 0
    # skipped faces add nothing; adding +0.0 leaves every sum unchanged
    tan = _accumulate(pos.shape[0], idx, T)
    bit = _accumulate(pos.shape[0], idx, B)
    t = tan - nrm * _dot(nrm, tan)[:, None]
    short = np.sqrt(_dot(t, t)) < F(1e-9)
    if short.any():
        n = nrm[short]
        alt = np.where((np.abs(n[:, 0]) < F(0.9))[:, None], _cross(n, np.tile(_f([1, 0, 0]), (n.shape[0], 1))),
                       _cross(n, np.tile(_f([0, 1, 0]), (n.shape[0], 1))))
        t[short] = alt
    t = _normalize(t)
    handed = np.where(_dot(_cross(nrm, t), bit) < F(0), F(-1), F(1))
    return np.concatenate([t, handed[:, None]], axis=1)


def mat4_from_rows(xf):
    """MathTypes.hpp:51-57: row-major 3x4 -> 4x4 (NULL = identity)."""
    m = np.eye(4, dtype=F)
    if xf is not None:
        m[:3, :] = _f(xf).reshape(3, 4)
    return m


def transform_box(lo, hi, m):
    """MathTypes.hpp:38-48 with glm's mat4 * vec4 association:
    (c0 x + c1 y) + (c2 z + c3 w)."""
    corners = np.array([[hi[0] if i & 1 else lo[0], hi[1] if i & 2 else lo[1], hi[2] if i & 4 else lo[2]]
                        for i in range(8)], F)
    w = (m[:3, 0] * corners[:, 0:1] + m[:3, 1] * corners[:, 1:2]) + (m[:3, 2] * corners[:, 2:3] + m[:3, 3] * F(1))
    return w.min(0), w.max(0)


def default_material():
    m = np.zeros(20, F)
    m[0:4] = (0.8, 0.8, 0.8, 1.0)
    m[4:8] = (0, 0, 0, 0.5)
    m[8:12] = (0, 0, 1.45, 0)
    m[12:17] = np.array([-1] * 5, np.int32).view(F)
    return m


class SceneDesc:
    """A scene as a sequence of C-ABI calls (ids from one counter, from 1)."""

    def __init__(self):
        self.next_id = 1
        self.meshes = {}      # id -> dict(positions, indices, normals, uvs)
        self.instances = {}   # id -> (mesh, xform3x4 or None, material id)
        self.materials = {}   # id -> material kwargs (Engine.material_desc names)
        self.lights = {}      # id -> light kwargs (Engine.light_desc names)
        self.textures = {}    # id -> slot
        self.slots = []       # slot -> dict(pixels, format) or None (removed)
        self.free_slots = []
        self.env = None       # (pixels, rotation, intensity)
        self.camera = None    # set_camera keyword arguments
        self.calls = []       # replay log

    def _id(self):
        i = self.next_id
        self.next_id += 1
        return i

    # -- ABI mirror ------------------------------------------------------------------------
    def add_mesh(self, positions, indices, normals=None, uvs=None):
        i = self._id()
        self.meshes[i] = dict(positions=_f(positions).reshape(-1, 3),
                              indices=np.asarray(indices, np.uint32).reshape(-1, 3),
                              normals=None if normals is None else _f(normals).reshape(-1, 3),
                              uvs=None if uvs is None else _f(uvs).reshape(-1, 2))
        self.calls.append(("add_mesh", i, (positions, indices), dict(normals=normals, uvs=uvs)))
        return i

    def update_mesh(self, mesh_id, positions, indices, normals=None, uvs=None):
        if mesh_id in self.meshes:
            self.meshes[mesh_id] = dict(positions=_f(positions).reshape(-1, 3),
                                        indices=np.asarray(indices, np.uint32).reshape(-1, 3),
                                        normals=None if normals is None else _f(normals).reshape(-1, 3),
                                        uvs=None if uvs is None else _f(uvs).reshape(-1, 2))
        self.calls.append(("update_mesh", None, (mesh_id, positions, indices), dict(normals=normals, uvs=uvs)))

    def remove_mesh(self, mesh_id):
        if self.meshes.pop(mesh_id, None) is not None:
            self.instances = {k: v for k, v in self.instances.items() if v[0] != mesh_id}
        self.calls.append(("remove_mesh", None, (mesh_id,), {}))

    def add_instance(self, mesh_id, xform3x4=None, material_id=0):
        if mesh_id not in self.meshes:
            return 0
        i = self._id()
        self.instances[i] = (mesh_id, None if xform3x4 is None else _f(xform3x4).reshape(12), material_id)
        self.calls.append(("add_instance", i, (mesh_id, xform3x4, material_id), {}))
        return i

    def set_instance_transform(self, inst_id, xform3x4):
        if inst_id in self.instances and xform3x4 is not None:
            m, _, mat = self.instances[inst_id]
            self.instances[inst_id] = (m, _f(xform3x4).reshape(12), mat)
        self.calls.append(("set_instance_transform", None, (inst_id, xform3x4), {}))

    def remove_instance(self, inst_id):
        self.instances.pop(inst_id, None)
        self.calls.append(("remove_instance", None, (inst_id,), {}))

    def add_material(self, **kw):
        i = self._id()
        self.materials[i] = dict(kw)
        self.calls.append(("add_material", i, (), dict(kw)))
        return i

    def update_material(self, mat_id, **kw):
        if mat_id in self.materials:
            self.materials[mat_id] = dict(kw)
        self.calls.append(("update_material", None, (mat_id,), dict(kw)))

    def add_texture(self, pixels, fmt=TEX_SRGB8):
        i = self._id()
        tex = dict(pixels=np.ascontiguousarray(pixels), format=fmt)
        if self.free_slots:
            slot = self.free_slots.pop()
            self.slots[slot] = tex
        else:
            slot = len(self.slots)
            self.slots.append(tex)
        self.textures[i] = slot
        self.calls.append(("add_texture", i, (pixels, fmt), {}))
        return i

    def remove_texture(self, tex_id):
        slot = self.textures.pop(tex_id, None)
        if slot is not None:
            self.slots[slot] = None
            self.free_slots.append(slot)
        self.calls.append(("remove_texture", None, (tex_id,), {}))

    def add_light(self, **kw):
        i = self._id()
        self.lights[i] = dict(kw)
        self.calls.append(("add_light", i, (), dict(kw)))
        return i

    def update_light(self, light_id, **kw):
        if light_id in self.lights:
            self.lights[light_id] = dict(kw)
        self.calls.append(("update_light", None, (light_id,), dict(kw)))

    def remove_light(self, light_id):
        self.lights.pop(light_id, None)
        self.calls.append(("remove_light", None, (light_id,), {}))

    def set_environment(self, pixels, rotation=0.0, intensity=1.0):
        self.env = None if pixels is None else (_f(pixels), float(rotation), float(intensity))
        self.calls.append(("set_environment", None, (pixels,), dict(rotation=rotation, intensity=intensity)))

    def set_camera(self, **kw):
        self.camera = dict(kw)
        self.calls.append(("set_camera", None, (), dict(kw)))

    # -- replay ------------------------------------------------------------------------------
    def load(self, engine):
        """Replays every call on a miniraytrace_amd Engine; returns it."""
        for name, expect_id, args, kw in self.calls:
            got = getattr(engine, name)(*args, **kw)
            if expect_id is not None:
                assert got == expect_id, f"{name} returned id {got}, the description assigned {expect_id}"
        return engine

    # -- packing -------------------------------------------------------------------------------
    def _mesh_arrays(self, m):
        pos, idx = m["positions"], m["indices"]
        nrm = m["normals"] if m["normals"] is not None else compute_normals(pos, idx)
        if m["uvs"] is not None:
            uv = m["uvs"]
            tan = compute_tangents(pos, idx, nrm, uv)
        else:
            uv = np.zeros((pos.shape[0], 2), F)
            tan = np.tile(_f([1, 0, 0, 1]), (pos.shape[0], 1))
        return pos, idx, nrm, uv, tan

    def _material_row(self, kw):
        def slot(tid):
            s = self.textures.get(int(tid or 0))
            return -1 if s is None or s >= MAX_TEXTURE_SLOTS else s
        g = default_material()
        base = kw.get("base_color", (0.8, 0.8, 0.8))
        g[0:3] = base
        g[3] = kw.get("opacity", 1.0)
        g[4:7] = kw.get("emission", (0, 0, 0))
        g[7] = np.clip(F(kw.get("roughness", 0.5)), 0, 1)
        g[8] = np.clip(F(kw.get("metallic", 0.0)), 0, 1)
        g[9] = np.clip(F(kw.get("transmission", 0.0)), 0, 1)
        g[10] = max(F(kw.get("ior", 1.45)), F(1))
        g[11] = 0
        g[12:17] = np.array([slot(kw.get(k, 0)) for k in
                             ("base_color_tex", "roughness_tex", "metallic_tex", "normal_tex", "emission_tex")],
                            np.int32).view(F)
        return g

    def _light_row(self, kw):
        g = np.zeros(20, F)
        t = kw.get("type", LIGHT_SUN)
        rad = _f(kw.get("radiance", (1, 1, 1)))
        if t == LIGHT_SUN:
            d = _normalize(_f(kw.get("direction", (0, 1, 0))))
            g[0:3] = d
            g[3] = np.cos(np.maximum(F(kw.get("angular_radius", 0.00465)), F(1e-4)), dtype=F)
            g[4:7] = rad
            gt = GPU_SUN
        elif t == LIGHT_POINT:
            g[0:3] = kw.get("position", (0, 0, 0))
            g[4:7] = rad
            gt = GPU_POINT
        else:
            e0, e1 = _f(kw.get("edge0", (1, 0, 0))), _f(kw.get("edge1", (0, 0, 1)))
            c = _cross(e0, e1)
            g[0:3] = kw.get("corner", (0, 0, 0))
            g[3] = np.sqrt(_dot(c, c))
            g[4:7] = e0
            g[8:11] = e1
            g[12:15] = rad
            g[15] = 1.0 if kw.get("two_sided", False) else 0.0
            gt = GPU_RECT
        g[16] = np.array([gt], np.int32).view(F)[0]
        return g

    def pack(self) -> dict:
        pos_l, nrm_l, uv_l, tan_l, tri_l, node_l = [], [], [], [], [], []
        blas = {}  # mesh id -> (global root, root lo, root hi, local nodes, local order)
        nv = nt = nn = 0
        for mid in sorted(self.meshes):
            pos, idx, nrm, uv, tan = self._mesh_arrays(self.meshes[mid])
            if idx.shape[0]:
                bounds = orc.tri_bounds(pos[idx].reshape(-1, 9))
                nodes, order = orc.bvh_build(bounds, 4)
            else:
                nodes, order = np.zeros(0, orc.NODE_DTYPE), np.zeros(0, np.uint32)
            tris = np.zeros((idx.shape[0], 4), np.uint32)
            tris[:, :3] = idx[order] + nv
            glob = nodes.copy()
            inner = (glob["packed"] >> 2) == 0
            glob["leftOrPrim"][inner] += np.uint32(nn)
            glob["leftOrPrim"][~inner] += np.uint32(nt)
            lo = nodes["lo"][0] if len(nodes) else None
            hi = nodes["hi"][0] if len(nodes) else None
            blas[mid] = (nn, lo, hi, nodes, order)
            pos4 = np.zeros((pos.shape[0], 4), F)
            pos4[:, :3] = pos
            nrm4 = np.zeros((pos.shape[0], 4), F)
            nrm4[:, :3] = nrm
            pos_l.append(pos4)
            nrm_l.append(nrm4)
            uv_l.append(uv)
            tan_l.append(tan)
            tri_l.append(tris)
            node_l.append(glob)
            nv += pos.shape[0]
            nt += idx.shape[0]
            nn += len(nodes)
        tlas_root = nn

        mat_index = {}
        mats = [default_material()]
        for mid in sorted(self.materials):
            mat_index[mid] = len(mats)
            mats.append(self._material_row(self.materials[mid]))

        flat, wb = [], []
        for iid in sorted(self.instances):
            mesh, xf, mat = self.instances[iid]
            b = blas[mesh]
            if b[1] is None or not b[1][0] <= b[2][0]:
                continue
            m = mat4_from_rows(xf)
            flat.append((iid, mesh, m, mat))
            lo, hi = transform_box(b[1], b[2], m)
            wb.append(np.concatenate([lo, hi]))
        if flat:
            tl_nodes, tl_order = orc.bvh_build(np.array(wb, F), 1)
        else:
            tl_nodes, tl_order = np.zeros(1, orc.NODE_DTYPE), np.zeros(0, np.uint32)
            tl_nodes["lo"], tl_nodes["hi"] = np.finfo(F).max, -np.finfo(F).max
            tl_nodes["packed"] = 0
        tl_nodes = tl_nodes.copy()
        inner = (tl_nodes["packed"] >> 2) == 0
        if flat:
            tl_nodes["leftOrPrim"][inner] += np.uint32(tlas_root)
        inst_rows, inst_ids = [], []
        for k in tl_order:
            iid, mesh, m, mat = flat[k]
            row = np.zeros(36, F)
            row[0:16] = m.T.reshape(-1)
            row[16:32] = np.linalg.inv(m.astype(np.float64)).astype(F).T.reshape(-1)
            row[32] = np.array([blas[mesh][0]], np.uint32).view(F)[0]
            row[33] = np.array([mat_index.get(mat, 0)], np.uint32).view(F)[0]
            inst_rows.append(row)
            inst_ids.append(iid)

        lights = []
        if self.env is not None:
            lights.append(np.zeros(20, F))
        for lid in sorted(self.lights):
            lights.append(self._light_row(self.lights[lid]))

        cat = lambda l, shape, dt: np.concatenate(l) if l else np.zeros(shape, dt)
        return {
            "nodes": np.concatenate(node_l + [tl_nodes]),
            "tlasRoot": tlas_root,
            "triangles": cat(tri_l, (0, 4), np.uint32),
            "positions": cat(pos_l, (0, 4), F),
            "normals": cat(nrm_l, (0, 4), F),
            "uvs": cat(uv_l, (0, 2), F),
            "tangents": cat(tan_l, (0, 4), F),
            "instances": np.array(inst_rows, F).reshape(-1, 36),
            "materials": np.array(mats, F),
            "lights": np.array(lights, F).reshape(-1, 20),
            "textures": [None if s is None else dict(pixels=s["pixels"], format=s["format"]) for s in self.slots],
            "envAlias": np.zeros((0, 2), np.uint32),
            # bookkeeping for comparisons with the engine's export
            "blas": {mid: (b[0], b[3], b[4]) for mid, b in blas.items()},
            "instanceIds": inst_ids,
        }

    def render_oracle(self, camera, width, height, alias=None, **kw):
        """oracle.pt_render on pack(); alias = the engine's alias table when
        the environment is to be sampled as the engine samples it."""
        sc = self.pack()
        env = None
        if self.env is not None:
            env, rot, inten = self.env
            kw.setdefault("env_rotation", rot)
            kw.setdefault("env_intensity", inten)
            if alias is not None:
                sc["envAlias"] = alias
                kw["use_alias"] = True
        return orc.pt_render(sc, env, camera, width, height, **kw)


# ---- compatibility helpers for single-mesh scenes -------------------------------------------
def material(base=(0.8, 0.8, 0.8), opacity=1.0, emission=(0, 0, 0), roughness=0.5, metallic=0.0,
             transmission=0.0, ior=1.45):
    """Engine.material_desc keyword arguments."""
    return dict(base_color=base, opacity=opacity, emission=emission, roughness=roughness, metallic=metallic,
                transmission=transmission, ior=ior)


def pack(positions, indices, normals=None, materials=(), instances=((None, 0),), env=True):
    """One mesh; instances: (xform3x4 or None, material index into `materials`
    + 1, 0 = default); env: an environment light slot (pixels supplied to the
    renderer separately)."""
    d = SceneDesc()
    mesh = d.add_mesh(positions, indices, normals=normals)
    mids = [d.add_material(**m) for m in materials]
    for xf, k in instances:
        d.add_instance(mesh, xf, mids[k - 1] if k else 0)
    if env:
        d.set_environment(np.ones((1, 1, 4), F))
    return d.pack()
