"""Frame graphs whose depth follows the frame's queue profile.

A captured plain frame ends with a tail launch that takes every path left at
the first level whose queue the previous frame held at most the tail
threshold in (engine.cpp fittedTailLevel); the per-level launches after it
would find empty queues.  The tail shades and traces exactly as the
per-level kernels do, so the image must not change: frames replayed from
such a graph are compared bit for bit with frames enqueued launch by launch
(no graph, every level).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NONE = 0xFFFFFFFF


def _print(*a):
    print(*a, flush=True)


def _engine(W, H, frames, tail, use_graph):
    from miniraytrace_amd import scenes
    from miniraytrace_amd.engine import Engine
    cfg = dict(scenes.CONFIG2)
    cfg.update(nx=120, nz=60, spheres=24, sphere_level=2, sky=(256, 128))
    e = Engine(width=W, height=H)
    e.set_render_settings(width=W, height=H, max_bounces=8, spp_limit=frames, firefly_clamp=0.0, denoise=False)
    scenes.load_terrain_scene(e, cfg, seed=1, sky_img=scenes.sky(256, 128))
    e.commit()
    e.set_path_options(seed=1, use_graph=use_graph, tail_paths=tail)
    return e


def test_fitted_graph_depth_renders_the_same_image(torch_gpu):
    W, H, F = 320, 180, 6
    # Queue profile of this scene, to put the tail threshold at level 2.
    e = _engine(W, H, 1, 0, False)
    e.set_profiling(True)
    e.render_frame()
    rays = e.frame_stats()["rays"]
    e.close()
    tail = int(rays[2])
    _print(f"case: fitted graph depth, {W}x{H}, rays per level {rays[:6]}, tail threshold {tail}")
    assert rays[1] > tail + tail // 8 and tail > 0

    out = {}
    for use_graph in (False, True):
        e = _engine(W, H, F, tail, use_graph)
        levels = []
        for _ in range(F):
            e.render_frame()
            levels.append(e.frame_stats()["graphTailLevel"])
        out[use_graph] = (e.read_rgba32f(W, H).copy(), levels)
        e.close()
    eager, graph = out[False], out[True]
    _print(f"graph tail levels per frame: {graph[1]}")
    assert eager[1] == [NONE] * F
    # the first frame has no queue profile yet; the graph captured after it ends at level 2
    assert graph[1][0] == NONE and all(lv == 2 for lv in graph[1][1:])
    assert np.array_equal(graph[0].view(np.uint32), eager[0].view(np.uint32))


def test_graph_depth_follows_the_queue_profile(torch_gpu):
    """The camera turns to the sky: level 1 empties, the graph ending at level
    2 no longer fits (the level before its last holds far fewer rays than the
    threshold), the next frame captures one ending at level 1; images still
    equal the launch-by-launch ones frame for frame."""
    from miniraytrace_amd import scenes
    W, H = 320, 180
    e = _engine(W, H, 1, 0, False)
    e.set_profiling(True)
    e.render_frame()
    tail = int(e.frame_stats()["rays"][2])
    e.close()
    cam = dict(scenes.CONFIG2["camera"])
    sky_cam = dict(cam, target=(cam["position"][0], cam["position"][1] + 10.0, cam["position"][2] - 0.5))
    _print(f"case: graph depth after a camera turn, {W}x{H}, tail threshold {tail}")
    runs = {}
    for use_graph in (False, True):
        e = _engine(W, H, 1 << 20, tail, use_graph)
        levels, images = [], []
        for f in range(7):
            if f == 3:
                e.set_camera(**sky_cam)
            e.render_frame()
            st = e.frame_stats()
            levels.append(st["graphTailLevel"])
            images.append(e.read_rgba32f(W, H).copy())
        runs[use_graph] = (levels, images)
        e.close()
    levels = runs[True][0]
    _print(f"graph tail levels per frame: {levels}")
    assert runs[False][0] == [NONE] * 7
    assert levels[0] == NONE and levels[1] == levels[2] == levels[3] == 2
    assert levels[5] == levels[6] and levels[5] in (0, 1)
    for f in range(7):
        assert np.array_equal(runs[True][1][f].view(np.uint32), runs[False][1][f].view(np.uint32)), f"frame {f}"
