"""GPU binned-SAH BVH builder (mrtxBuildBvhDevice) against the CPU builder
(mrtxBuildBvh, which tests/test_oracle_bvh.py pins against the oracle's
restatement of core/src/accel/Bvh.cpp) and against the oracle builder
itself at 1M triangles: identical node records and primitive order, byte
for byte.

Cases: sizes on both sides of the 512-primitive threshold between the
level-by-level and the one-wavefront-per-subtree phases (and of the
1024-slot chunks of the first), random soups, the 1,000,000-triangle
terrain, coincident centroids (leaf by
degenerate extent), and boxes so large that SAH prefers an oversized leaf,
which needs the reference's median split and so is built on the CPU by the
same call.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def gpu(torch_gpu):
    return torch_gpu


def _soup_bounds(n, seed, size=0.03):
    rng = np.random.default_rng(seed)
    c = rng.random((n, 3), dtype=np.float32)
    v = c[:, None, :] + (rng.random((n, 3, 3), dtype=np.float32) * 2 - 1) * size
    return np.concatenate([v.min(1), v.max(1)], 1).astype(np.float32)


def _mesh_bounds(pos, idx):
    t = pos[idx.reshape(-1, 3)]
    return np.concatenate([t.min(1), t.max(1)], 1).astype(np.float32)


@pytest.fixture(scope="module")
def fallbacks():
    """Messages of builds that went to the CPU (the median split), through
    mrtSetLogCallback."""
    import ctypes as C
    from miniraytrace_amd import _lib as L
    seen = []

    def on_log(level, msg, user):
        text = msg.decode()
        if "median split" in text:
            seen.append(text)

    cb = L.LOG_FN(on_log)
    L.lib().mrtSetLogCallback(cb, None)
    yield seen
    L.lib().mrtSetLogCallback(L.LOG_FN(), None)


def _check(torch, bounds, max_leaf=4, label="", fallbacks=None, on_gpu=True):
    from miniraytrace_amd.bvh import build_bvh, build_bvh_device
    d = torch.from_numpy(bounds).cuda()
    timing = {}
    before = len(fallbacks) if fallbacks is not None else 0
    gn, go = build_bvh_device(d, max_leaf, timing=timing)
    if fallbacks is not None:
        assert (len(fallbacks) == before) == on_gpu, "built on the GPU" if on_gpu else "built on the CPU"
    t_gpu = timing["ms"] * 1e-3
    t0 = time.perf_counter()
    cn, co = build_bvh(bounds, max_leaf)
    t_cpu = time.perf_counter() - t0
    _print(f"{label}: {bounds.shape[0]} prims, {len(gn)} nodes; GPU {t_gpu * 1e3:.1f} ms, CPU {t_cpu * 1e3:.1f} ms")
    assert len(gn) == len(cn)
    assert gn.tobytes() == cn.tobytes(), f"{int((gn != cn).sum())} nodes differ"
    assert np.array_equal(go, co)
    return gn, go


@pytest.mark.parametrize("n", [1, 3, 4, 5, 17, 512, 513, 1000, 1024, 1025, 2049, 40_000])
def test_random_soups(gpu, fallbacks, n):
    _check(gpu, _soup_bounds(n, 100 + n), label=f"soup {n}", fallbacks=fallbacks)


def test_leaf_sizes(gpu, fallbacks):
    b = _soup_bounds(5000, 7)
    for leaf in (1, 2, 8):
        _check(gpu, b, leaf, label=f"soup 5000, max leaf {leaf}", fallbacks=fallbacks)


def test_terrain_million_matches_oracle(gpu, fallbacks, oracle):
    from miniraytrace_amd import scenes
    pos, idx = scenes.terrain(1000, 500, 20.0, 10.0, 0.2, 1)
    b = _mesh_bounds(pos, idx)
    gn, go = _check(gpu, b, label="terrain 1M", fallbacks=fallbacks)
    t0 = time.perf_counter()
    on, oo = oracle.bvh_build(b, 4)
    _print(f"oracle builder {time.perf_counter() - t0:.1f} s")
    assert gn.tobytes() == on.tobytes()
    assert np.array_equal(go, oo)


def test_degenerate_inputs(gpu, fallbacks):
    # every centroid equal: one leaf of all primitives (extent below 1e-8)
    same = np.tile(np.array([[0, 0, 0, 1, 1, 1]], np.float32), (3000, 1))
    gn, _ = _check(gpu, same, label="coincident centroids", fallbacks=fallbacks)
    assert len(gn) == 1 and (gn[0]["packed"] >> 2) == 3000
    # clusters of coincident centroids among spread ones
    b = _soup_bounds(6000, 3)
    b[:2500] = b[0]
    _check(gpu, b, label="clustered centroids", fallbacks=fallbacks)
    # boxes much larger than the centroid spread: SAH prefers a leaf, the
    # reference takes the median split (built on the CPU by the same call)
    rng = np.random.default_rng(5)
    c = rng.random((4000, 3), dtype=np.float32)
    huge = np.concatenate([c - 50.0, c + 50.0], 1).astype(np.float32)
    _check(gpu, huge, label="median-split fallback", fallbacks=fallbacks, on_gpu=False)


@pytest.mark.parametrize("n", [400, 40_000])
def test_signed_zero_ties(gpu, fallbacks, n):
    """Bounds that tie between -0.0 and +0.0 keep the first zero in slot
    order, in the one-wavefront subtrees and in the level phase."""
    from tests.test_oracle_bvh import signed_zero_bounds, zero_signs
    gn, _ = _check(gpu, signed_zero_bounds(n, 20 + n), label=f"signed zeros {n}", fallbacks=fallbacks)
    neg, pos = zero_signs(gn)
    assert neg > 0 and pos > 0


def test_engine_commit_uses_identical_bvh(gpu):
    """A 1M-triangle mesh commits through the GPU builder; the packed BLAS
    equals the CPU builder's over the same triangle bounds."""
    from miniraytrace_amd import scenes
    from miniraytrace_amd.bvh import build_bvh
    from miniraytrace_amd.engine import Engine
    pos, idx = scenes.terrain(1000, 500, 20.0, 10.0, 0.2, 1)
    e = Engine(width=64, height=64)
    m = e.add_mesh(pos, idx)
    e.add_instance(m, None, 0)
    st = e.commit()
    _print(f"commit of 1M triangles: {st['commitMs']:.1f} ms")
    ex = e.export_scene()
    e.close()
    cn, co = build_bvh(_mesh_bounds(pos, idx), 4)
    root = 1  # behind the unused node that puts child pairs on 64-byte lines
    got = ex["nodes"][root:root + len(cn)].copy()
    inner = (got["packed"] >> 2) == 0
    got["leftOrPrim"][inner] -= np.uint32(root)
    assert got.tobytes() == cn.tobytes()
    assert np.array_equal(ex["triangles"][:, :3], idx.reshape(-1, 3)[co])
