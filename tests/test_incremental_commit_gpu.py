"""Incremental commit with growth (reference core/src/render/Renderer.cpp:
333-508, tests/commit_tests.cpp:54-111): an edited mesh is rewritten alone,
in place when its new data fits the ranges it holds in the vertex, triangle
and node arrays, else appended behind the used ranges; the TLAS goes into a
region reserved behind the BLAS nodes, and instances outgrowing it force a
full repack, as do appends beyond the allocated capacity and holes over half
the used space.

Each case checks the commit counters of mrtxGetFrameStats (repacks, meshes
written alone, bytes uploaded), that the edited engine renders bit for bit
what a fresh engine given the final scene renders (the layouts differ, the
images may not), and, for the in-place and appended cases, the FP64 oracle
sample for sample on the final scene packed independently.
"""
import numpy as np
import pytest

from tests.test_abi_live_gpu import _engine, _final_copy, _fresh, _oracle_check, _render, _xf, base_scene

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def gpu(torch_gpu):
    return torch_gpu


def _with_terrain(d, ids):
    """A 7,200-triangle terrain below the floor, reaching past it: the
    arrays then hold room (a quarter of the used size) for a grown sphere."""
    from miniraytrace_amd import scenes
    pos, idx = scenes.terrain(60, 60, 16.0, 16.0, 0.3, 3)
    pos = pos.copy()
    pos[:, 1] -= 0.4
    ids["terrain"] = d.add_mesh(pos, idx.reshape(-1))
    ids["m_terrain"] = d.add_material(base_color=(0.3, 0.5, 0.3), roughness=0.9)
    d.add_instance(ids["terrain"], None, ids["m_terrain"])


def _run(oracle, edits, prepare=None, check_oracle=True):
    """First commit + render; then each (edit, expect) in turn: the edit is
    applied to the description and replayed on the live engine, the commit's
    counter deltas are compared with expect = (repacks, meshes written), and
    the image with a fresh engine's.  Returns the list of commit stats."""
    d, ids = base_scene()
    if prepare:
        prepare(d, ids)
    e = _engine(d)
    e.commit()
    first = e.frame_stats()
    _render(e)
    out = []
    prev = first
    for edit, expect in edits:
        n_before = len(d.calls)
        edit(d, ids)
        for name, _, args, kw in d.calls[n_before:]:
            getattr(e, name)(*args, **kw)
        e.commit()
        st = e.frame_stats()
        delta = (st["geometryRepacks"] - prev["geometryRepacks"], st["meshesWritten"] - prev["meshesWritten"])
        _print(f"  commit: {st['commitMs']:.2f} ms, {st['commitUploadBytes']} B uploaded, "
               f"repacks +{delta[0]}, meshes written +{delta[1]}")
        assert delta == expect, f"expected (repacks, meshes written) {expect}, got {delta}"
        out.append((st, first))
        prev = st
        img = _render(e)
        alias = e.export_scene()["envAlias"] if d.env is not None else None
        fresh = _fresh(_final_copy(d))
        assert np.array_equal(img.view(np.uint32), fresh.view(np.uint32)), "edited engine differs from a fresh one"
        if check_oracle:
            _oracle_check(oracle, d, img, alias)
    e.close()
    return out


def _deformed_sphere(level, amp):
    from miniraytrace_amd import scenes
    v, f = scenes.icosphere(level)
    v = v * (1 + amp * np.sin(5 * v[:, :1]))
    return v.astype(np.float32), f


def test_same_size_mesh_edit_in_place(gpu, oracle):
    """Same vertex and triangle counts: the mesh is rewritten in its ranges;
    the upload is that mesh, the TLAS and the instances, not the scene."""
    _print("case: mesh edit in place")
    v, f = _deformed_sphere(2, 0.2)
    (st, first), = _run(oracle, [(lambda d, ids: d.update_mesh(ids["sphere"], v, f), (0, 1))], _with_terrain)
    assert st["commitUploadBytes"] < 0.25 * first["commitUploadBytes"]


def test_grown_mesh_appended(gpu, oracle):
    """A mesh that outgrows its ranges moves behind the used ranges (its old
    ranges become a hole); then a new mesh is appended too, and the export
    holds both behind the TLAS region."""
    _print("case: grown mesh appended, new mesh appended")
    v, f = _deformed_sphere(3, 0.15)
    cube_v, cube_f = _cube()

    def add_cube(d, ids):
        ids["cube"] = d.add_mesh(cube_v, cube_f)
        d.add_instance(ids["cube"], _xf((0.9, 0.25, 0.6), 0.25, 0.3), ids["m_red"])

    _run(oracle, [(lambda d, ids: d.update_mesh(ids["sphere"], v, f), (0, 1)), (add_cube, (0, 1))], _with_terrain)


def _cube():
    v = np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)], np.float32)
    f = np.array([0, 1, 3, 0, 3, 2, 4, 6, 7, 4, 7, 5, 0, 4, 5, 0, 5, 1,
                  2, 3, 7, 2, 7, 6, 0, 2, 6, 0, 6, 4, 1, 5, 7, 1, 7, 3], np.uint32)
    return v, f


def test_appended_layout_in_export(gpu):
    """After an append the appended mesh's BLAS root lies behind the TLAS
    region, its triangles behind the others, and instances point at it."""
    _print("case: export after an append")
    d, ids = base_scene()
    _with_terrain(d, ids)
    e = _engine(d)
    e.commit()
    ex0 = e.export_scene()
    v, f = _deformed_sphere(3, 0.15)
    e.update_mesh(ids["sphere"], v, f)
    d.update_mesh(ids["sphere"], v, f)
    e.commit()
    ex = e.export_scene()
    e.close()
    roots = ex["instances"][:, 32].view(np.uint32)
    assert ex["tlasRoot"] == ex0["tlasRoot"]
    assert roots.max() > ex["tlasRoot"] and ex["triangles"].shape[0] > ex0["triangles"].shape[0]
    assert ex["nodes"].shape[0] > ex0["nodes"].shape[0]


def test_tlas_capacity_outgrown_repacks(gpu, oracle):
    """Instances beyond the TLAS region reserved at the last repack force a
    full repack (Renderer.cpp:484-489)."""
    _print("case: TLAS capacity outgrown")

    def add_instances(d, ids):
        for k in range(4):
            d.add_instance(ids["sphere"], _xf((-1.2 + 0.8 * k, 0.15, 1.0), 0.15), ids["m_gold"])

    _run(oracle, [(add_instances, (1, 0))], check_oracle=False)


def test_instances_within_capacity_no_repack(gpu):
    """One added instance fits the reserved TLAS region: no repack."""
    _print("case: TLAS within capacity")
    _run(None, [(lambda d, ids: d.add_instance(ids["sphere"], _xf((0.0, 0.2, 1.0), 0.2), ids["m_gold"]), (0, 0))],
         check_oracle=False)


def test_holes_over_half_repack(gpu):
    """Removing the terrain leaves most of the arrays unused: that commit
    repacks."""
    _print("case: holes over half the used space")
    _run(None, [(lambda d, ids: d.remove_mesh(ids["terrain"]), (1, 0))], _with_terrain, check_oracle=False)


def test_append_beyond_capacity_repacks(gpu):
    """A grown mesh that does not fit the arrays' spare room repacks."""
    from miniraytrace_amd import scenes
    _print("case: append beyond capacity")
    pos, idx = scenes.terrain(70, 70, 16.0, 16.0, 0.3, 4)
    pos = pos.copy()
    pos[:, 1] -= 0.4
    _run(None, [(lambda d, ids: d.update_mesh(ids["terrain"], pos, idx.reshape(-1)), (1, 0))], _with_terrain,
         check_oracle=False)
