"""The incremental-scene half of the C ABI on a live engine (reference
capi/src/capi.cpp:93-316, core/src/scene/Scene.cpp:100-320,
core/src/engine/Engine.cpp:44-60, tests/commit_tests.cpp:54-111).

For every mutator -- mrtUpdateMesh, mrtRemoveMesh, mrtSetInstanceTransform,
mrtRemoveInstance, mrtUpdateMaterial, mrtAddLight, mrtUpdateLight,
mrtRemoveLight, mrtSetEnvironment(NULL) -- an engine edited after a first
commit and render must:
* report the reference's commit statistics (a transform, instance or
  material edit rebuilds the TLAS only, a mesh edit rebuilds that mesh's
  BLAS only, a light edit rebuilds neither);
* render bit for bit what a fresh engine given the final scene renders;
* match the FP64 oracle sample for sample on the final scene, packed
  independently (tests/scene_pack.py).
mrtSetLogCallback delivers the library's messages with the user pointer and
stops after a NULL callback.  RGBA8 readback equals the resolve pass
restated in oracle/resolve.py (resolve.comp:24-56) to one code, for both
tone maps, several exposures and values above 1.
"""
import ctypes as C

import numpy as np
import pytest

from tests.conftest import cpu_threads
from tests.scene_pack import SceneDesc

pytestmark = pytest.mark.gpu

W, H, BOUNCES, SPP = 40, 32, 3, 2


def _print(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def gpu(torch_gpu):
    return torch_gpu


def _xf(t=(0, 0, 0), s=1.0, ang=0.0):
    c, sn = np.cos(ang), np.sin(ang)
    return np.array([[c * s, 0, sn * s, t[0]], [0, s, 0, t[1]], [-sn * s, 0, c * s, t[2]]], np.float32).reshape(12)


def base_scene():
    """Returns (desc, ids): floor, two sphere instances, three materials,
    sun + point lights, sky."""
    from miniraytrace_amd import scenes
    d = SceneDesc()
    ids = {}
    ids["m_floor"] = d.add_material(base_color=(0.7, 0.7, 0.7), roughness=0.8)
    ids["m_red"] = d.add_material(base_color=(0.8, 0.2, 0.1), roughness=0.4)
    ids["m_gold"] = d.add_material(base_color=(0.9, 0.7, 0.3), roughness=0.3, metallic=1.0)
    pos = np.array([[-4, 0, -4], [4, 0, -4], [4, 0, 4], [-4, 0, 4]], np.float32)
    ids["floor"] = d.add_mesh(pos, [0, 2, 1, 0, 3, 2])
    v, f = scenes.icosphere(2)
    ids["sphere"] = d.add_mesh(v, f, normals=v)
    d.add_instance(ids["floor"], None, ids["m_floor"])
    ids["i_a"] = d.add_instance(ids["sphere"], _xf((-0.6, 0.5, 0), 0.5), ids["m_red"])
    ids["i_b"] = d.add_instance(ids["sphere"], _xf((0.7, 0.4, -0.3), 0.4), ids["m_gold"])
    ids["sun"] = d.add_light(type=0, direction=(0.3, 1.0, 0.4), radiance=(8.0, 7.0, 6.0), angular_radius=0.02)
    ids["point"] = d.add_light(type=1, position=(0.0, 1.5, 1.0), radiance=(2.0, 2.0, 2.0))
    d.set_environment(scenes.sky(32, 16), 0.0, 0.5)
    d.set_camera(position=(0.0, 1.2, 3.0), target=(0.0, 0.3, 0.0), fov_y_deg=45.0)
    return d, ids


def _engine(desc):
    from miniraytrace_amd.engine import Engine
    e = Engine(width=W, height=H)
    e.set_render_settings(width=W, height=H, max_bounces=BOUNCES, spp_limit=64, firefly_clamp=0.0, denoise=False)
    desc.load(e)
    return e


def _render(e, frames=SPP):
    for _ in range(frames):
        e.render_frame()
    return e.read_rgba32f(W, H)


def _fresh(desc):
    e = _engine(desc)
    img = _render(e)
    e.close()
    return img


def _oracle_check(oracle, desc, img, alias):
    c = desc.camera
    cam = oracle.camera_lookat(c["position"], c["target"], (0, 1, 0), c["fov_y_deg"], W, H)
    ref, _ = desc.render_oracle(cam, W, H, alias=alias, spp=SPP, max_bounces=BOUNCES, threads=cpu_threads())
    ref /= SPP
    g = img[..., :3].astype(np.float64)
    rel = np.abs(g - ref).max(-1) / np.maximum(np.abs(ref).max(-1), 1e-3)
    agree = (rel < 1e-3).mean()
    _print(f"  oracle agreement {agree:.4f}")
    assert agree >= 0.99


def _edit_case(oracle, edit, expect_blas, expect_tlas):
    """First commit + render, apply `edit(desc_live, engine, ids)` to the
    live engine (the description records the same calls), then check commit
    statistics, equality with a fresh engine and the oracle."""
    d, ids = base_scene()
    e = _engine(d)
    e.commit()
    _render(e)
    n_before = len(d.calls)
    edit(d, ids)
    # replay only the new calls on the live engine
    for name, _, args, kw in d.calls[n_before:]:
        getattr(e, name)(*args, **kw)
    st = e.commit()
    _print(f"  commit after edit: {st}")
    assert st["blasRebuilt"] == expect_blas and st["tlasRebuilt"] == expect_tlas
    img = _render(e)
    alias = e.export_scene()["envAlias"] if d.env is not None else None
    e.close()
    fresh = _fresh(_final_copy(d))
    assert np.array_equal(img.view(np.uint32), fresh.view(np.uint32)), "edited engine differs from a fresh one"
    _oracle_check(oracle, d, img, alias)
    return d


def _final_copy(d):
    """A description holding only the final state, built with fresh calls
    in id order (ids may differ from the edited scene's; images do not
    depend on id values, only on their order)."""
    f = SceneDesc()
    tex = {}
    for tid, slot in sorted(d.textures.items()):
        t = d.slots[slot]
        tex[tid] = f.add_texture(t["pixels"], t["format"])
    mats = {}
    for mid in sorted(d.materials):
        kw = dict(d.materials[mid])
        for k in ("base_color_tex", "roughness_tex", "metallic_tex", "normal_tex", "emission_tex"):
            if kw.get(k):
                kw[k] = tex[kw[k]]
        mats[mid] = f.add_material(**kw)
    meshes = {}
    for mid in sorted(d.meshes):
        m = d.meshes[mid]
        meshes[mid] = f.add_mesh(m["positions"], m["indices"].reshape(-1), normals=m["normals"], uvs=m["uvs"])
    for iid in sorted(d.instances):
        mesh, xf, mat = d.instances[iid]
        f.add_instance(meshes[mesh], xf, mats.get(mat, 0))
    for lid in sorted(d.lights):
        f.add_light(**d.lights[lid])
    if d.env is not None:
        f.set_environment(*d.env)
    f.set_camera(**d.camera)
    return f


def test_set_instance_transform_rebuilds_tlas_only(gpu, oracle):
    _print("case: mrtSetInstanceTransform")
    _edit_case(oracle, lambda d, ids: d.set_instance_transform(ids["i_a"], _xf((-0.2, 0.7, 0.4), 0.6, 0.5)), 0, 1)


def test_remove_instance_rebuilds_tlas_only(gpu, oracle):
    _print("case: mrtRemoveInstance")
    _edit_case(oracle, lambda d, ids: d.remove_instance(ids["i_b"]), 0, 1)


def test_update_material_rebuilds_tlas_only(gpu, oracle):
    _print("case: mrtUpdateMaterial")
    _edit_case(oracle, lambda d, ids: d.update_material(ids["m_red"], base_color=(0.1, 0.3, 0.9), roughness=0.2,
                                                          metallic=0.5), 0, 1)


def test_update_mesh_rebuilds_that_blas(gpu, oracle):
    from miniraytrace_amd import scenes
    _print("case: mrtUpdateMesh")
    v, f = scenes.icosphere(3)
    v = v * (1 + 0.15 * np.sin(5 * v[:, :1]))
    _edit_case(oracle, lambda d, ids: d.update_mesh(ids["sphere"], v.astype(np.float32), f), 1, 1)


def test_remove_mesh_cascades_to_instances(gpu, oracle):
    _print("case: mrtRemoveMesh")
    d = _edit_case(oracle, lambda d, ids: d.remove_mesh(ids["sphere"]), 0, 1)
    assert len(d.instances) == 1


def test_add_light(gpu, oracle):
    _print("case: mrtAddLight")
    _edit_case(oracle, lambda d, ids: d.add_light(type=2, corner=(-0.5, 1.8, -0.5), edge0=(1.0, 0, 0),
                                                    edge1=(0, 0, 0.6), radiance=(3.0, 3.0, 3.0)), 0, 0)


def test_update_light(gpu, oracle):
    _print("case: mrtUpdateLight")
    _edit_case(oracle, lambda d, ids: d.update_light(ids["point"], type=1, position=(0.8, 1.0, 0.6),
                                                       radiance=(5.0, 1.0, 1.0)), 0, 0)


def test_remove_light(gpu, oracle):
    _print("case: mrtRemoveLight")
    _edit_case(oracle, lambda d, ids: d.remove_light(ids["sun"]), 0, 0)


def test_clear_environment(gpu, oracle):
    _print("case: mrtSetEnvironment(NULL)")
    d = _edit_case(oracle, lambda d, ids: d.set_environment(None), 0, 0)
    assert d.env is None


def test_light_edit_export_counts(gpu):
    _print("case: light and environment edits in the packed light table")
    d, ids = base_scene()
    e = _engine(d)
    assert e.export_scene()["lights"].shape[0] == 3  # environment + 2 lights
    e.remove_light(ids["point"])
    e.set_environment(None)
    ex = e.export_scene()
    assert ex["lights"].shape[0] == 1 and ex["envWidth"] == 0
    assert ex["lights"][0, 16].view(np.int32) == 1  # the sun
    e.close()


def test_log_callback(gpu):
    from miniraytrace_amd import _lib as L
    from miniraytrace_amd.engine import Engine
    _print("case: mrtSetLogCallback")
    got = []

    @L.LOG_FN
    def cb(level, msg, user):
        got.append((level, msg.decode(), user))

    marker = C.c_void_p(0x5EED)
    L.lib().mrtSetLogCallback(cb, marker)
    try:
        Engine(width=16, height=16).close()
        with pytest.raises(L.MrtError):
            Engine(width=16, height=16, gpu_index1=999)
        assert any(lv == 1 and "engine on HIP device" in m for lv, m, _ in got), got
        assert any(lv == 3 and "out of range" in m for lv, m, _ in got), got
        assert all(u == 0x5EED for _, _, u in got)
        n = len(got)
        L.lib().mrtSetLogCallback(L.LOG_FN(), None)
        Engine(width=16, height=16).close()
        assert len(got) == n
    finally:
        L.lib().mrtSetLogCallback(L.LOG_FN(), None)


@pytest.mark.parametrize("tonemap,ev", [(1, 0.0), (1, -1.5), (0, 0.0), (0, 2.0)])
def test_rgba8_matches_resolve_restatement(gpu, tonemap, ev):
    from miniraytrace_amd import scenes
    from oracle import resolve as R
    _print(f"case: RGBA8 resolve tonemap={tonemap} exposureEV={ev}")
    d, _ = base_scene()
    # emissive panels spanning dark to far above 1 (exact values at depth 0)
    for k, em in enumerate((0.001, 0.02, 0.3, 1.0, 4.0, 60.0)):
        m = d.add_material(base_color=(0, 0, 0), emission=(em, em * 0.5, em * 0.25))
        pos = np.array([[0, 0, 0], [0.25, 0, 0], [0.25, 0.25, 0], [0, 0.25, 0]], np.float32)
        d.add_instance(d.add_mesh(pos, [0, 1, 2, 0, 2, 3]), _xf((-0.8 + 0.27 * k, 0.9, 0.5)), m)
    from miniraytrace_amd.engine import Engine
    e = Engine(width=W, height=H)
    e.set_render_settings(width=W, height=H, max_bounces=BOUNCES, spp_limit=8, firefly_clamp=0.0, denoise=False,
                          tonemap=tonemap, exposure_ev=ev)
    d.load(e)
    for _ in range(3):
        e.render_frame()
    lin = e.read_rgba32f(W, H)
    out8 = e.read_rgba8(W, H)
    e.close()
    ref = R.resolve_rgba8(lin[..., :3], ev, tonemap)
    diff = np.abs(out8.astype(int) - ref.astype(int))
    _print(f"  max code difference {diff.max()}, codes off by one: {(diff > 0).mean():.4f}; "
           f"linear range {lin[..., :3].min():.4f} .. {lin[..., :3].max():.1f}")
    assert diff.max() <= 1
    assert np.all(out8[..., 3] == 255)
    assert lin[..., :3].max() > 1.0 and (out8[..., :3] == 255).any() and (out8[..., :3] < 20).any()
