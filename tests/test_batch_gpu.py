"""Batched rendering (mrtxRenderBatch) and the path-tracer autotuner on the GPU.

* A batch of requests with uneven resolutions, scenes and sample counts
  leaves every engine with exactly the image that the same number of
  mrtRenderFrame calls produces (bit for bit, RGBA32F), spp limits included.
* The autotuner picks supported launch shapes and does not change images.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def gpu(torch_gpu):
    return torch_gpu


def _engine(req, sky, spp_limit=1 << 20):
    from miniraytrace_amd import scenes
    from miniraytrace_amd.engine import Engine
    e = Engine(width=req["width"], height=req["height"])
    e.set_render_settings(width=req["width"], height=req["height"], max_bounces=8, spp_limit=spp_limit,
                          firefly_clamp=0.0, denoise=False)
    scenes.load_terrain_scene(e, scenes.terrain_config(req["triangles"]), seed=req["seed"], sky_img=sky)
    e.commit()
    return e


# Path-tracer launch shapes (threads per block, LDS stack entries) the kernels are built for.
SUPPORTED = {(64, 16), (128, 8), (128, 16), (128, 32), (256, 8), (256, 16)}

REQS = [dict(width=64, height=48, triangles=6_000, spp=3, seed=1000),
        dict(width=97, height=41, triangles=40_000, spp=1, seed=1001),
        dict(width=128, height=72, triangles=15_000, spp=5, seed=1002),
        dict(width=32, height=32, triangles=9_000, spp=4, seed=1003)]


def test_batch_equals_one_at_a_time(gpu):
    from miniraytrace_amd import scenes
    from miniraytrace_amd.batch import render_batch
    _print("case: batch vs one at a time, 4 requests + spp-limit clamp")
    sky = scenes.sky(256, 128)
    limits = [1 << 20, 1 << 20, 1 << 20, 2]  # the last request stops at 2 spp
    batch = [_engine(r, sky, lim) for r, lim in zip(REQS, limits)]
    single = [_engine(r, sky, lim) for r, lim in zip(REQS, limits)]
    st = render_batch(batch, [r["spp"] for r in REQS])
    _print("batch stats:", st)
    assert st["mode"] == 0 and st["framesDone"] == [3, 1, 5, 2]
    for e, r in zip(single, REQS):
        for _ in range(r["spp"]):
            e.render_frame()
    for i, (a, b, r) in enumerate(zip(batch, single, REQS)):
        ia = a.read_rgba32f(r["width"], r["height"])
        ib = b.read_rgba32f(r["width"], r["height"])
        assert np.array_equal(ia, ib), i
        assert np.isfinite(ia).all() and ia[..., :3].mean() > 0
    # a second batch call continues the accumulation like further frames do
    st2 = render_batch(batch, [2, 2, 2, 2])
    assert st2["framesDone"] == [2, 2, 2, 0]
    for e, n in zip(single, [2, 2, 2, 0]):
        for _ in range(n):
            e.render_frame()
    for a, b, r in zip(batch, single, REQS):
        assert np.array_equal(a.read_rgba32f(r["width"], r["height"]), b.read_rgba32f(r["width"], r["height"]))


def test_batch_with_guided_request_equals_one_at_a_time(gpu):
    """A request with guiding training on joins the merged batch (its query
    records, and training after each frame, run inside the shared graph);
    training without guided sampling leaves images as plain, so they still
    match one-at-a-time rendering."""
    from miniraytrace_amd import scenes
    from miniraytrace_amd.batch import render_batch
    _print("case: batch with a guided (training-only) request")
    sky = scenes.sky(256, 128)
    reqs = REQS[:3]
    batch = [_engine(r, sky) for r in reqs]
    single = [_engine(r, sky) for r in reqs]
    for e in (batch[1], single[1]):
        e.set_guiding(sampling=False, training=True, batch_size=1024, steps_per_frame=1)
    st = render_batch(batch, [r["spp"] for r in reqs])
    assert st["mode"] == 0 and st["framesDone"] == [r["spp"] for r in reqs]
    for e, r in zip(single, reqs):
        for _ in range(r["spp"]):
            e.render_frame()
    for a, b, r in zip(batch, single, reqs):
        assert np.array_equal(a.read_rgba32f(r["width"], r["height"]), b.read_rgba32f(r["width"], r["height"]))
    assert batch[1].guiding_stats()["steps"] == single[1].guiding_stats()["steps"] > 0


def test_batch_guided_sampling_runs_merged(gpu):
    """Guided sampling in the merged batch: every request finishes its frames
    with finite images whose mean matches plain rendering (unbiased) within
    the noise of a few frames."""
    from miniraytrace_amd import scenes
    from miniraytrace_amd.batch import render_batch
    _print("case: batch with guided sampling")
    sky = scenes.sky(256, 128)
    r = REQS[2]
    g1, g2, plain = _engine(r, sky), _engine(r, sky), _engine(r, sky)
    for e in (g1, g2):
        e.set_guiding(sampling=True, training=True, batch_size=4096, steps_per_frame=2)
    st = render_batch([g1, g2, plain], [32, 32, 32])
    assert st["mode"] == 0 and st["framesDone"] == [32, 32, 32]
    imgs = [e.read_rgba32f(r["width"], r["height"])[..., :3].astype(np.float64) for e in (g1, g2, plain)]
    for im in imgs:
        assert np.isfinite(im).all() and im.mean() > 0
    assert abs(imgs[0].mean() / imgs[2].mean() - 1.0) < 0.05
    assert abs(imgs[1].mean() / imgs[2].mean() - 1.0) < 0.05


def test_batch_chain_mode_equals_one_at_a_time(gpu):
    """More requests than one merged batch holds (64) run as per-request graph
    chains in one launch; images equal one-at-a-time rendering."""
    from miniraytrace_amd import scenes
    from miniraytrace_amd.batch import render_batch
    from miniraytrace_amd.engine import Engine
    _print("case: batch chain mode (65 requests)")
    sky = scenes.sky(64, 32)

    def small(k):
        e = Engine(width=16, height=16)
        e.set_render_settings(width=16, height=16, max_bounces=4, spp_limit=1 << 20, firefly_clamp=0.0,
                              denoise=False)
        scenes.load_terrain_scene(e, scenes.terrain_config(2_000), seed=2000 + k, sky_img=sky)
        e.commit()
        return e

    batch = [small(k) for k in range(65)]
    single = [small(k) for k in range(65)]
    st = render_batch(batch, [2] * 65)
    assert st["mode"] == 1 and st["framesDone"] == [2] * 65
    for e in single:
        for _ in range(2):
            e.render_frame()
    for a, b in zip(batch, single):
        assert np.array_equal(a.read_rgba32f(16, 16), b.read_rgba32f(16, 16))
    for e in batch + single:
        e.close()


def test_batch_argument_checks(gpu):
    from miniraytrace_amd import _lib as L, scenes
    from miniraytrace_amd.batch import render_batch
    e = _engine(REQS[0], scenes.sky(64, 32))
    with pytest.raises(L.MrtError):
        render_batch([e, e], [1, 1])  # duplicate engine
    assert render_batch([], [])["requests"] == 0


def test_autotune_keeps_images(gpu):
    from miniraytrace_amd import scenes
    _print("case: autotune")
    sky = scenes.sky(256, 128)
    r = REQS[2]
    a, b = _engine(r, sky), _engine(r, sky)
    chosen = a.autotune_path()
    _print("chosen shapes:", chosen)
    assert tuple(chosen["extend"]) in SUPPORTED and tuple(chosen["shadow"]) in SUPPORTED
    for e in (a, b):
        for _ in range(4):
            e.render_frame()
    assert np.array_equal(a.read_rgba32f(r["width"], r["height"]), b.read_rgba32f(r["width"], r["height"]))


@pytest.mark.parametrize("tail_paths", [1 << 30, 3000])
def test_tail_kernel_equals_per_level(gpu, tail_paths):
    """The tail kernel (one lane per path for the rest of the frame) gives
    the same image, bit for bit, and the same per-level ray counts as the
    per-level launches: from the first level (tail_paths above the frame's
    paths) and from a level in the middle of the frame."""
    from miniraytrace_amd import _lib as L, scenes
    _print(f"case: tail kernel vs per-level launches, tail_paths={tail_paths}")
    sky = scenes.sky(256, 128)
    r = REQS[2]
    a, b = _engine(r, sky), _engine(r, sky)
    a.set_path_options(tail_paths=tail_paths)
    b.set_path_options(tail_paths=L.TAIL_DISABLED)
    for e in (a, b):
        for _ in range(3):
            e.render_frame()
        e.set_profiling(True)
        e.render_frame()
    sa, sb = a.frame_stats(), b.frame_stats()
    _print("rays per level:", list(sa["rays"])[:10], "tail ms", sa["gpuMs"]["tail"], "vs", sb["gpuMs"]["tail"])
    assert list(sa["rays"]) == list(sb["rays"]) and list(sa["shadowRays"]) == list(sb["shadowRays"])
    assert sa["gpuMs"]["tail"] > 0.0 and sb["gpuMs"]["tail"] >= 0.0
    ia, ib = a.read_rgba32f(r["width"], r["height"]), b.read_rgba32f(r["width"], r["height"])
    assert np.array_equal(ia, ib)
    assert np.isfinite(ia).all() and ia[..., :3].mean() > 0
