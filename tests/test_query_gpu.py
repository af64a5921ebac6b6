"""GPU parity of the closest-hit / any-hit traversal kernels (config 1:
10,000-triangle seeded soup, 1,000,000 rays), called through the C ABI
(mrtxTraceClosest / mrtxTraceAny in libmrt.so).

Bar: identical triangle ids and bit-identical t against the oracle's
watertight brute force and BVH traversal (ties resolved to the lowest id);
against the reference's own CPU traversal (tests/bvh_tests.cpp:49-73,
Moeller-Trumbore) ids agree except on edge-grazing / sliver rays, counted.
"""
import numpy as np
import pytest

from tests.conftest import cpu_threads

pytestmark = pytest.mark.gpu

N_RAYS = 1_000_000


def _print(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def c1(torch_gpu, oracle):
    from miniraytrace_amd import scenes
    from miniraytrace_amd.query import QueryScene
    torch = torch_gpu
    verts = scenes.soup(10_000, 1)
    rays = scenes.soup_rays(N_RAYS, 2)
    qs = QueryScene.from_soup(verts)
    nodes, order = qs.export_bvh()
    d_rays = torch.from_numpy(rays).cuda()
    return dict(torch=torch, verts=verts, rays=rays, qs=qs, nodes=nodes, order=order, d_rays=d_rays)


def _trace(c, qs=None, rays=None):
    torch = c["torch"]
    qs = qs or c["qs"]
    d_rays = c["d_rays"] if rays is None else torch.from_numpy(np.ascontiguousarray(rays)).cuda()
    hits = torch.empty((d_rays.shape[0], 2), dtype=torch.int32, device="cuda")
    qs.trace_closest(d_rays, hits)
    torch.cuda.synchronize()
    assert qs.status() == 0
    return hits.cpu().numpy().view(np.uint32)


def test_builder_matches_oracle(c1, oracle):
    _print("case: c1 builder parity")
    n_o, o_o = oracle.bvh_build(oracle.tri_bounds(c1["verts"]), 4)
    assert n_o.tobytes() == c1["nodes"].tobytes()
    assert np.array_equal(o_o, c1["order"])


def test_closest_exact_vs_oracle(c1, oracle):
    _print("case: c1 closest 1M rays vs oracle bvh_closest")
    gpu = _trace(c1)
    ref = oracle.bvh_closest(c1["nodes"], c1["order"], c1["verts"], c1["rays"], cpu_threads())
    assert np.array_equal(gpu[:, 0], ref[:, 0])
    assert np.array_equal(gpu[:, 1], ref[:, 1])  # t bit-identical
    assert (gpu[:, 0] != 0xFFFFFFFF).mean() > 0.8


def test_closest_exact_vs_brute_force(c1, oracle):
    _print("case: c1 closest vs brute force (first 20k rays)")
    gpu = _trace(c1)[:20_000]
    brute = oracle.brute_closest(c1["verts"], c1["rays"][:20_000], cpu_threads())
    assert np.array_equal(gpu, brute)


def test_closest_vs_reference_traversal(c1, oracle):
    """Reference CPU traversal (Moeller-Trumbore): ids agree except on a
    handful of edge / sliver rays; t agrees to the reference test's 1e-4."""
    _print("case: c1 closest vs reference MT traversal")
    gpu = _trace(c1)
    ref = oracle.ref_traverse(c1["nodes"], c1["order"], c1["verts"], c1["rays"], cpu_threads())
    mismatch = (gpu[:, 0] != ref[:, 0]).sum()
    _print(f"id mismatches vs Moeller-Trumbore reference: {mismatch} / {N_RAYS}")
    assert mismatch <= 50
    same = (gpu[:, 0] == ref[:, 0]) & (gpu[:, 0] != 0xFFFFFFFF)
    tg = gpu[same, 1].view(np.float32).astype(np.float64)
    tr = ref[same, 1].view(np.float32).astype(np.float64)
    rel = np.abs(tg - tr) / tr
    _print(f"t rel diff: max {rel.max():.3e}, >1e-6: {(rel > 1e-6).sum()}")
    assert rel.max() < 1e-4
    assert (rel > 1e-6).mean() < 0.01


@pytest.mark.parametrize("block,stack,loop", [(64, 8, 0), (64, 32, 0), (128, 16, 0), (256, 8, 0), (256, 32, 0),
                                              (64, 8, 1), (128, 16, 1), (256, 32, 1), (64, 8, 2), (128, 8, 2),
                                              (256, 32, 2), (128, 16, 3), (256, 8, 3)])
def test_all_launch_shapes_identical(c1, oracle, block, stack, loop):
    _print(f"case: c1 shape block={block} stack={stack} loop={loop}")
    base = _trace(c1)
    c1["qs"].set_config(block, stack, loop)
    try:
        other = _trace(c1)
    finally:
        c1["qs"].set_config(128, 16, 0)
    assert np.array_equal(base, other)


@pytest.mark.parametrize("loop", [0, 1, 2, 3])
def test_any_hit_vs_brute(c1, oracle, loop):
    _print(f"case: c1 any-hit with finite tmax, loop {loop}")
    torch = c1["torch"]
    rays = c1["rays"][:50_000].copy()
    rng = np.random.default_rng(4)
    rays[:, 7] = rng.uniform(0.5, 3.0, size=rays.shape[0]).astype(np.float32)
    d = torch.from_numpy(rays).cuda()
    occ = torch.empty(rays.shape[0], dtype=torch.int32, device="cuda")
    c1["qs"].set_config(128, 16, loop)
    try:
        c1["qs"].trace_any(d, occ)
        torch.cuda.synchronize()
        got = occ.cpu().numpy()
        exp = oracle.brute_any(c1["verts"], rays, cpu_threads())
        assert np.array_equal(got.astype(np.uint8), exp)
        # closest with the same tmax agrees with occlusion
        hits = _trace(c1, rays=rays)
        assert np.array_equal(hits[:, 0] != 0xFFFFFFFF, exp.astype(bool))
    finally:
        c1["qs"].set_config(128, 16, 0)


@pytest.mark.parametrize("loop", [0, 2])
def test_edge_cases(c1, oracle, loop):
    """Also a 100-primitive leaf, which the four-wide walk pushes as a
    refetch entry."""
    _print(f"case: edge cases (empty, single, coincident ties, axis rays), loop {loop}")
    from miniraytrace_amd.query import QueryScene
    torch = c1["torch"]
    # empty scene
    qs = QueryScene(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32))
    qs.set_config(128, 8, loop)
    rays = c1["rays"][:1000]
    h = _trace(c1, qs=qs, rays=rays)
    assert np.all(h[:, 0] == 0xFFFFFFFF)
    qs.close()
    # one triangle hit head-on along an axis (direction components exactly 0)
    tri = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]], np.float32)
    qs = QueryScene.from_soup(tri)
    qs.set_config(128, 8, loop)
    r = np.array([[0.25, 0.25, 5, 1e-4, 0, 0, -1, np.inf],
                  [0.8, 0.8, 5, 1e-4, 0, 0, -1, np.inf],
                  [0.25, 0.25, 5, 1e-4, 0, 0, -1, 4.0],
                  [0.0, 0.0, 5, 1e-4, 0, 0, -1, np.inf]], np.float32)
    h = _trace(c1, qs=qs, rays=r)
    exp = oracle.brute_closest(tri, r, 1)
    assert np.array_equal(h, exp)
    assert h[0, 0] == 0 and h[1, 0] == 0xFFFFFFFF and h[2, 0] == 0xFFFFFFFF
    qs.close()
    # 100 coincident triangles (one 100-prim leaf): the tie goes to id 0
    same = np.tile(tri, (100, 1))
    same[37] += np.float32(0.0)
    qs = QueryScene.from_soup(same)
    qs.set_config(128, 8, loop)
    h = _trace(c1, qs=qs, rays=r)
    assert h[0, 0] == 0 and h[0, 1] == np.float32(5.0).view(np.uint32)
    assert np.array_equal(h, oracle.brute_closest(same, r, 1))
    qs.close()
    # duplicated soup: every tie must resolve to the lower copy
    dup = np.concatenate([c1["verts"][:2000], c1["verts"][:2000]])
    qs = QueryScene.from_soup(dup)
    qs.set_config(128, 8, loop)
    rr = c1["rays"][:20000]
    h = _trace(c1, qs=qs, rays=rr)
    assert np.array_equal(h, oracle.brute_closest(dup, rr, cpu_threads()))
    assert np.all((h[:, 0] < 2000) | (h[:, 0] == 0xFFFFFFFF))
    qs.close()
