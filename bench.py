#!/usr/bin/env python3
"""Benchmark driver.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Workloads:
  pt         (default) plain progressive path tracing at 1920x1080 on a
             1.256M-triangle terrain scene; one step = one frame (1 spp).
             The default run appends the traversal, batch, guided,
             convergence, (shortened) scale, bvh and edit objects below.
  traversal  closest-hit queries: 10,000-triangle seeded soup, 1,000,000
             rays (one step = one mrtxTraceClosest launch over all rays).
  guided     equal-RMSE wall time of guided vs plain path tracing on a room
             lit through a window (1280x720, 499k triangles).
  batch      64 uneven requests rendered by one batched graph launch vs one
             frame call at a time (images must be identical).
  scale      guided and batched together: 16 requests at 4K / 1080p over a
             10.4M-triangle instanced field and the room, depth 12, one
             guiding network per request (8 frames per request; the default
             run appends a shortened run: 4 frames, MSE at 16 spp).
  converge   RMSE of the plain 1080p terrain frame's centre crop against
             the FP64 oracle at every power-of-two spp up to 1024.
  bvh        GPU BLAS build of the 1M-triangle terrain vs the threaded CPU
             builder (identical nodes), and the engine commit of that mesh.
  edit       edit-to-frame latency on the 1080p terrain scene: an instance
             moved, the sphere mesh and the terrain deformed, a mesh added.

Prints ONE JSON line.  Inputs are resident in device memory before the timed
region; the timed region is bracketed by device synchronisation.  The
roofline block prices the dominant kernel with ALGORITHMIC bytes (reference
traversal order, counted by the oracle in the cpu_baseline leg) divided by
its average launch time measured with HIP events on the launch stream.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Msamples/s at 1920x1080 + equal-RMSE wall time, guided vs plain, one MI355X"


def _log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _measured_traffic(kernel: str, key: str):
    """HBM bytes of `kernel` from the committed rocprofv3 counter summary, or
    None when that exact kernel instantiation was not profiled."""
    try:
        with open(os.path.join(ROOT, "profiles", "hbm_traffic.json")) as f:
            rec = json.load(f)["kernels"].get(kernel)
    except (OSError, ValueError, KeyError):
        return None
    return None if rec is None else rec.get(key)


# -------------------------------------------------------------------- CPU leg --
# Every use of the oracle (oracle/, the CPU restatement) in this file is in
# the functions of this section, and runs only with the CPU baseline on (not
# --no-cpu-baseline): the timed CPU baselines, the reference-order node and
# triangle counts that price the rooflines' algorithmic bytes, and the FP64
# reference images that RMSE figures are measured against.  No GPU result is
# computed by it.
def _cpu_camera(camera, W, H):
    from oracle import oracle as orc
    return orc.camera_lookat(camera["position"], camera["target"], camera.get("up", (0, 1, 0)),
                             camera["fov_y_deg"], W, H)


def _centre_crop(W, H, cw, ch):
    cw, ch = min(cw, W), min(ch, H)
    return ((W - cw) // 2, (H - ch) // 2, cw, ch)


def _cpu_fetch_counts(exported, sky, camera, W, H, crop, max_bounces) -> np.ndarray:
    """Reference-order fetch counts of one frame's paths on a crop (the
    traversal order of intersect.glsl:48-178): closest queries, nodes
    popped, triangles tested, then the same three for shadow queries."""
    from oracle import oracle as orc
    _, counts = orc.pt_render(exported, sky, _cpu_camera(camera, W, H), W, H, crop=crop, frame0=0, spp=1,
                              max_bounces=max_bounces, threads=orc.hardware_threads(), count_fetches=True)
    return counts


def _cpu_fp64_reference(exported, sky, camera, W, H, crop, spp, max_bounces) -> np.ndarray:
    """FP64 crop image of the reference estimator (reference CDF environment
    sampling) over spp frames numbered from 2^20, so its samples are
    independent of every GPU frame."""
    from oracle import oracle as orc
    img, _ = orc.pt_render(exported, sky, _cpu_camera(camera, W, H), W, H, crop=crop, frame0=1 << 20, spp=spp,
                           max_bounces=max_bounces, threads=orc.hardware_threads())
    return img / spp


def _trace_bytes(counts, rays, shadows) -> dict:
    """Algorithmic bytes of one frame's trace launches: per closest-hit query
    a 32-byte ray read and a 20-byte hit record write, per shadow query a
    48-byte shadow-ray read, and for both 32 B per node popped and 64 B per
    triangle tested in the reference order; per-query figures from a crop,
    scaled to the frame's queue sizes (rays, shadows: per bounce level)."""
    q, nn, nt, qs, ns, ts = (int(x) for x in counts)
    per_query = 32 + 20 + 32 * nn / max(q, 1) + 64 * nt / max(q, 1)
    per_shadow = 48 + 32 * ns / max(qs, 1) + 64 * ts / max(qs, 1)
    return {"bytes_per_frame": per_query * sum(rays) + per_shadow * sum(shadows),
            "bytes_per_query": per_query, "bytes_per_shadow_query": per_shadow,
            "nodes_per_query": nn / max(q, 1), "tris_per_query": nt / max(q, 1),
            "nodes_per_shadow_query": ns / max(qs, 1), "tris_per_shadow_query": ts / max(qs, 1)}


def _profiled_rays(e):
    """Per-level ray and shadow-ray counts of one profiled frame; the
    engine's accumulation restarts afterwards."""
    e.set_profiling(True)
    e.render_frame()
    st = e.frame_stats()
    e.set_profiling(False)
    e.reset_accumulation()
    return [int(x) for x in st["rays"]], [int(x) for x in st["shadowRays"]]


def _rmse(a, b) -> float:
    return float(np.sqrt(((a - b) ** 2).mean()))


# ----------------------------------------------------------------- traversal --
def traversal_workload(args, torch) -> dict:
    from miniraytrace_amd import scenes
    from miniraytrace_amd.query import QueryScene

    n_tri, n_rays = 10_000, 1_000_000
    verts = scenes.soup(n_tri, 1)
    rays = scenes.soup_rays(n_rays, 2)
    qs = QueryScene.from_soup(verts)
    info = qs.info()
    d_rays = torch.from_numpy(rays).cuda()
    hits = torch.empty((n_rays, 2), dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream()
    if args.autotune:
        blk, stk, loop = qs.autotune(d_rays, hits)
    else:
        blk, stk, loop = qs.config()
    _log(f"traversal: {info['nodeCount']} nodes, depth {info['maxDepth']}, shape {blk}x{stk} loop {loop}")

    for _ in range(args.warmup):
        qs.trace_closest(d_rays, hits, stream)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        qs.trace_closest(d_rays, hits, stream)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = ev0.elapsed_time(ev1) / args.steps
    assert qs.status() == 0, "traversal stack overflow"
    gpu_hits = hits.cpu().numpy().view(np.uint32)

    out = {
        "workload": "closest-hit queries, 10k-triangle soup, 1M rays",
        "rays_per_step": n_rays,
        "ms_per_step": wall * 1e3 / args.steps,
        "kernel_ms": kernel_ms,
        "value_mrays_s": n_rays / (kernel_ms * 1e-3) / 1e6,
        "shape": {"block": blk, "lds_stack": stk,
                  "loop": ["step", "while-while", "four-wide step", "four-wide while-while"][loop]},
        "bvh": {k: info[k] for k in ("nodeCount", "maxDepth", "leafCount", "buildMs")},
    }

    if not args.no_cpu_baseline:
        from oracle import oracle as orc
        threads = orc.hardware_threads()
        nodes, order = qs.export_bvh()
        # algorithmic fetch counts in the reference traversal order
        href, _, (nn, nt) = orc.glsl_traverse(nodes, order, verts, rays, threads, want_hits=True)
        bytes_alg = 32 * n_rays + 8 * n_rays + 32 * nn + 64 * nt
        achieved = bytes_alg / (kernel_ms * 1e-3) / 1e9
        kname = f"{'kTraceClosestWide' if loop >= 2 else 'kTraceClosest'}<{blk}, {stk}, {loop}>"
        out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                           "traffic": _measured_traffic(kname, "bytes_per_launch"),
                           "kernel": kname,
                           "bytes_per_launch": bytes_alg,
                           "nodes_per_ray": nn / n_rays, "tris_per_ray": nt / n_rays}
        # parity spot check against the oracle's exact traversal
        check = orc.bvh_closest(nodes, order, verts, rays[:100_000], threads)
        out["parity_100k_exact"] = bool(np.array_equal(check, gpu_hits[:100_000]))
        # CPU baseline: the reference traversal (bvh_tests.cpp:49-73 restated),
        # all host threads, repeated to >= 10 s of CPU time.
        reps, t_cpu = 0, 0.0
        while t_cpu < 10.0 / max(1, threads) * 4 or reps < 2:
            t0 = time.perf_counter()
            orc.ref_traverse(nodes, order, verts, rays, threads)
            t_cpu += time.perf_counter() - t0
            reps += 1
            if reps >= 20:
                break
        t1 = time.perf_counter()
        orc.ref_traverse(nodes, order, verts, rays[:100_000], 1)
        one = 100_000 / (time.perf_counter() - t1)
        out["cpu_baseline"] = {"value": n_rays * reps / t_cpu / 1e6, "unit": "Mrays/s",
                               "cores": threads, "kind": "port",
                               "sample": f"{reps} x 1M config-1 rays, reference traversal "
                                         f"(Moeller-Trumbore, bvh_tests.cpp:49-73 restated)",
                               "single_thread_mrays_s": one / 1e6}
    qs.close()
    return out


def pt_workload(args, torch) -> dict:
    """Plain progressive path tracing at 1920x1080, 1 spp per frame, max depth 8:
    seeded terrain of 1,000,000 triangles + 200 instanced icospheres (1,280
    triangles each), 3 GGX materials, procedural 2048x1024 sky with a 0.5 deg
    sun.  One step = one mrtRenderFrame (1 sample for each of the 2,073,600
    pixels), through the C ABI."""
    from miniraytrace_amd import scenes
    from miniraytrace_amd.engine import Engine

    W, H = 1920, 1080
    e = Engine(width=W, height=H)
    total = args.warmup + args.steps + 8
    e.set_render_settings(width=W, height=H, max_bounces=8, spp_limit=total, firefly_clamp=0.0, denoise=False)
    t0 = time.perf_counter()
    sky = scenes.sky()
    info = scenes.load_terrain_scene(e, scenes.CONFIG2, seed=1, sky_img=sky)
    st = e.commit()
    _log(f"pt: scene {info['triangles']} triangles, {info['instances']} instances, commit "
         f"{st['commitMs']:.0f} ms (setup {time.perf_counter() - t0:.1f} s)")
    for _ in range(args.warmup):
        e.render_frame()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        e.render_frame()
    wall = time.perf_counter() - t0
    msamples = W * H * args.steps / wall / 1e6

    # Profiling pass: per-stage GPU time (HIP events on the engine's stream)
    # and queue sizes per bounce.
    e.set_profiling(True)
    prof = []
    for _ in range(3):
        e.render_frame()
        prof.append(e.frame_stats())
    e.set_profiling(False)
    stage = {k: float(np.mean([p["gpuMs"][k] for p in prof])) for k in prof[0]["gpuMs"]}
    ext = e.path_options()["extend"]
    rays = [int(x) for x in prof[-1]["rays"]]
    shadows = [int(x) for x in prof[-1]["shadowRays"]]
    out = {
        "workload": "plain progressive path tracing 1920x1080, 1 spp/frame, max depth 8, "
                    "terrain 1M tris + 200 icospheres (1.256M tris), 2048x1024 sky",
        "frames": args.steps,
        "ms_per_frame": wall * 1e3 / args.steps,
        "msamples_s": msamples,
        "gpu_stage_ms": stage,
        "rays_per_bounce": rays,
        "shadow_rays_per_bounce": shadows,
        "status": prof[-1]["status"],
        "triangles": info["triangles"],
    }

    if not args.no_cpu_baseline:
        from oracle import oracle as orc
        threads = orc.hardware_threads()
        exp = e.export_scene()
        cam = _cpu_camera(scenes.CONFIG2["camera"], W, H)
        # Bounded CPU sample: a 96x54 crop of the frame centre, 1 spp,
        # FP64 restatement with reference-order fetch counting.
        cw, ch = 96, 54
        crop = _centre_crop(W, H, cw, ch)
        t1 = time.perf_counter()
        counts = _cpu_fetch_counts(exp, sky, scenes.CONFIG2["camera"], W, H, crop, 8)
        t_count = time.perf_counter() - t1
        tb = _trace_bytes(counts, rays, shadows)
        frame_queries, frame_shadows = sum(rays), sum(shadows)
        bytes_alg = tb["bytes_per_frame"]
        # the tail kernel traces the last few paths' rays (and shades them)
        trace_ms = stage["extend"] + stage["shadow"] + stage["tail"]
        achieved = bytes_alg / (trace_ms * 1e-3) / 1e9
        kname = f"kTraceBatch<{ext[0]}, {ext[1]}>"
        out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": achieved / HBM_PEAK_GBS,
                           "traffic": _measured_traffic(kname, "bytes_per_frame"),
                           "kernel": f"{kname} (closest hit + any hit, all trace launches of one frame, "
                                     "with the tail kernel)",
                           "bytes_per_frame": bytes_alg, "trace_ms_per_frame": trace_ms,
                           "queries_per_frame": frame_queries, "shadow_queries_per_frame": frame_shadows,
                           "nodes_per_query": tb["nodes_per_query"], "tris_per_query": tb["tris_per_query"],
                           "nodes_per_shadow_query": tb["nodes_per_shadow_query"],
                           "tris_per_shadow_query": tb["tris_per_shadow_query"],
                           "crop_queries_per_path": int(counts[0]) / (cw * ch)}
        # CPU baseline: the FP64 path tracer timed on the same crop without counting.
        reps, t_cpu = 0, 0.0
        while t_cpu < 10.0 and reps < 64:
            t1 = time.perf_counter()
            orc.pt_render(exp, sky, cam, W, H, crop=crop, frame0=reps, spp=1, max_bounces=8, threads=threads)
            t_cpu += time.perf_counter() - t1
            reps += 1
        out["cpu_baseline"] = {"value": cw * ch * reps / t_cpu / 1e6, "unit": "Msamples/s", "cores": threads,
                               "kind": "port",
                               "sample": f"{reps} frames of a {cw}x{ch} centre crop, FP64 restatement of "
                                         "pathtrace.comp (no CPU path tracer exists in the reference)",
                               "count_pass_s": t_count}
    e.close()
    return out


# ------------------------------------------------------------------ convergence --
def convergence_workload(args, torch, sky) -> dict:
    """Convergence of the plain tracer on the config-2 frame (1920x1080, the
    1.256M-triangle terrain scene, depth 8, no firefly clamp): RMSE of the
    192x108 centre crop at spp = 1, 2, 4, ..., args.converge_spp against the
    FP64 oracle's crop at 4096 samples (reference CDF environment sampling,
    independent frames), with the fit rmse^2 = sigma^2 / spp + sigma^2 / 4096
    (the reference's own noise) and each checkpoint's ratio to it.  The sun
    (5e4 x the zenith radiance) makes single samples heavy-tailed, so at
    high spp a checkpoint's RMSE moves with the few sun paths it holds."""
    from miniraytrace_amd import scenes
    from miniraytrace_amd.engine import Engine
    W, H = 1920, 1080
    top = args.converge_spp
    e = Engine(width=W, height=H)
    e.set_render_settings(width=W, height=H, max_bounces=8, spp_limit=top, firefly_clamp=0.0, denoise=False)
    info = scenes.load_terrain_scene(e, scenes.CONFIG2, seed=1, sky_img=sky)
    e.commit()
    x0, y0, cw, ch = crop = _centre_crop(W, H, 192, 108)
    checkpoints = [1 << k for k in range(top.bit_length()) if (1 << k) <= top]
    crops, done = {}, 0
    t_gpu = 0.0
    for s_ in checkpoints:
        t_gpu += _timed_frames(e, s_ - done, torch)
        done = s_
        crops[s_] = e.read_rgba32f(W, H)[y0:y0 + ch, x0:x0 + cw, :3].astype(np.float64)
    exp = e.export_scene()
    e.close()
    ref_spp = 4096
    t0 = time.perf_counter()
    ref = _cpu_fp64_reference(exp, sky, info["camera"], W, H, crop, ref_spp, 8)
    t_ref = time.perf_counter() - t0
    rmse = np.array([_rmse(crops[s_], ref) for s_ in checkpoints])
    x = 1.0 / np.array(checkpoints, np.float64) + 1.0 / ref_spp
    sigma2 = float((rmse ** 2 * x).sum() / (x * x).sum())
    model = np.sqrt(sigma2 * x)
    out = {
        "workload": "plain PT convergence, 1920x1080 terrain 1.256M tris, 192x108 centre crop vs FP64 at 4096 spp",
        "spp": checkpoints,
        "rmse": [float(v) for v in rmse],
        "model_rmse": [float(v) for v in model],
        "rmse_over_model": [float(v) for v in rmse / model],
        "rmse_sqrt_spp": [float(v * np.sqrt(s_)) for v, s_ in zip(rmse, checkpoints)],
        "sigma": float(np.sqrt(sigma2)),
        "reference_spp": ref_spp,
        "reference_rmse_floor": float(np.sqrt(sigma2 / ref_spp)),
        "gpu_s": t_gpu,
        "reference_cpu_s": t_ref,
    }
    ratio = rmse / model
    worst = int(np.abs(ratio - 1).argmax())
    _log("convergence: RMSE " + ", ".join(f"{s_}:{v:.4g}" for s_, v in zip(checkpoints, rmse)) +
         f"; sigma {out['sigma']:.4f}; RMSE / (1/sqrt(spp) model) from {ratio.min():.3f} to {ratio.max():.3f} "
         f"(farthest from 1 at {checkpoints[worst]} spp); FP64 reference {t_ref:.1f} s")
    return out


# ------------------------------------------------------------- guided vs plain --
def _room_engine(W, H, seed, sky):
    from miniraytrace_amd import scenes
    from miniraytrace_amd.engine import Engine
    e = Engine(width=W, height=H)
    e.set_render_settings(width=W, height=H, max_bounces=8, spp_limit=1 << 30, firefly_clamp=0.0, denoise=False)
    info = scenes.load_room_scene(e, scenes.CONFIG4, seed=1, sky_img=sky)
    e.set_path_options(seed=seed)
    e.commit()
    return e, info


# Guiding network FLOPs (78 -> 64 -> 64 -> 64 -> 64 -> 40): forward 2 * sum of
# in x out = 39,680 per query; backward 39,680 for the weight gradients plus
# 2 * (3 * 64 * 64 + 64 * 40) = 29,696 for the deltas (no input gradient).
MLP_FWD_FLOP = 2 * (78 * 64 + 3 * 64 * 64 + 64 * 40)
MLP_TRAIN_FLOP = MLP_FWD_FLOP + MLP_FWD_FLOP + 2 * (3 * 64 * 64 + 64 * 40)
MFMA_BF16_PEAK_TFLOPS = 2500.0


def _mlp_roofline(rows, infer_ms, train_samples, train_ms) -> dict:
    """MFMA view of one guided frame: the network inference (all levels'
    query rows) and the training steps, against the dense BF16 peak.  The
    stage times include the launches' non-MFMA work (encoding, decoding,
    record collection, Adam), so these are lower bounds on MFMA use."""
    infer = rows * MLP_FWD_FLOP / (infer_ms * 1e-3) / 1e12 if infer_ms > 0 else None
    train = train_samples * MLP_TRAIN_FLOP / (train_ms * 1e-3) / 1e12 if train_ms > 0 else None
    return {"bound": "mfma", "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
            "infer_rows_per_frame": rows, "infer_flop_per_row": MLP_FWD_FLOP, "infer_ms": infer_ms,
            "infer_achieved": infer, "infer_frac": infer / MFMA_BF16_PEAK_TFLOPS if infer else None,
            "train_samples_per_frame": train_samples, "train_flop_per_sample": MLP_TRAIN_FLOP, "train_ms": train_ms,
            "train_achieved": train, "train_frac": train / MFMA_BF16_PEAK_TFLOPS if train else None}


def _fastest_tail(e, seed, torch, per_cu=(640, 960, 1280, 1920)) -> int:
    """Sets the engine's tail threshold to the candidate (rays per compute
    unit) with the least time over 8 frames and returns it.  The tail kernel
    changes speed only, never the image; the trial frames are discarded."""
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    best = None
    for c in per_cu:
        e.set_path_options(seed=seed, tail_paths=c * cus)
        _timed_frames(e, 2, torch)
        t = _timed_frames(e, 8, torch)
        if best is None or t < best[0]:
            best = (t, c)
    e.set_path_options(seed=seed, tail_paths=best[1] * cus)
    e.reset_accumulation()
    return best[1]


def _timed_frames(e, n, torch) -> float:
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(n):
        e.render_frame()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def guided_workload(args, torch) -> dict:
    """Equal-RMSE wall time, guided vs plain: a closed diffuse room (6 walls of
    204 x 204 quads, 499,392 triangles, albedo 0.7) lit only by the sky and
    its 0.5 degree sun through a window of 5 % of one wall, 1280x720, max
    depth 8, no firefly clamp.  Plain renders `spp` frames at its fastest tail
    threshold for the scene (chosen over 8-frame trials); guided starts from
    an untrained network, trains online (batch 65,536, 2 Adam steps per frame)
    and samples the learned mixture with probability 0.5, and is timed until
    its RMSE reaches plain's.  RMSE is taken against an independent plain
    render of 2 x spp samples (the unbiased estimator the oracle pins, so a
    bias of the guided estimator would show, not cancel); its own error adds
    equally to both sides.  With the CPU leg on, plain's spp-sample image and
    the last guided checkpoint are also compared with the FP64 oracle on a
    64x36 centre crop (4096 samples)."""
    from miniraytrace_amd import scenes
    W, H = scenes.CONFIG4["width"], scenes.CONFIG4["height"]
    spp = args.guided_spp
    sky = scenes.sky()
    plain, info = _room_engine(W, H, 11, sky)
    # The plain tracer runs at its fastest tail threshold for this scene (the
    # default is tuned on the terrain); guided requests hand levels to the
    # unguided tail by their own rule (a quarter of their pixels).
    plain_tail = _fastest_tail(plain, 11, torch)
    _log(f"guided: room {info['triangles']} triangles, {W}x{H}, plain target {spp} spp, "
         f"plain tail threshold {plain_tail} rays per CU")
    plain.render_frame()
    plain.reset_accumulation()
    t_plain = _timed_frames(plain, spp, torch)
    img_p = plain.read_rgba32f(W, H)[..., :3].astype(np.float64)
    plain.close()

    ref, _ = _room_engine(W, H, 99, sky)
    ref.set_path_options(seed=99, tail_paths=plain_tail * torch.cuda.get_device_properties(0).multi_processor_count)
    n_ref = 2 * spp
    t_ref = _timed_frames(ref, n_ref, torch)
    img_r = ref.read_rgba32f(W, H)[..., :3].astype(np.float64)
    exp = ref.export_scene() if not args.no_cpu_baseline else None
    ref.close()
    mse_p = float(((img_p - img_r) ** 2).mean())

    g, _ = _room_engine(W, H, 12, sky)
    g.set_guiding(sampling=True, training=True)
    warm = args.guided_warmup
    # frames that train from the untrained network with their samples
    # discarded (0 = every guided sample counts); their time is counted
    t_g = _timed_frames(g, warm, torch) if warm else 0.0
    g.reset_accumulation()
    # Render until guided has both reached plain's error and run plain's wall
    # time (at most 2 x spp frames); checkpoints every 16 frames (readbacks untimed).
    curve, done, chunk = [], 0, 16
    while done < 2 * spp:
        t_g += _timed_frames(g, chunk, torch)
        done += chunk
        img_g = g.read_rgba32f(W, H)[..., :3].astype(np.float64)
        curve.append((done, t_g, float(((img_g - img_r) ** 2).mean())))
        if curve[-1][2] <= mse_p and t_g >= t_plain:
            break
    img_g_last = g.read_rgba32f(W, H)[..., :3].astype(np.float64)
    g.set_profiling(True)
    g.render_frame()
    prof = g.frame_stats()
    st = g.guiding_stats()
    opts = g.guiding()
    g.close()

    def crossing(key_idx, target, out_idx):
        """Power-law interpolation of column out_idx where column key_idx meets target."""
        for a, b in zip(curve, curve[1:]):
            if (a[key_idx] - target) * (b[key_idx] - target) <= 0 and a[key_idx] != b[key_idx]:
                f = np.log(target / a[key_idx]) / np.log(b[key_idx] / a[key_idx])
                return float(a[out_idx] * (b[out_idx] / a[out_idx]) ** f)
        return None

    t_equal_rmse = crossing(2, mse_p, 1) if curve[0][2] > mse_p else curve[0][1]
    mse_equal_time = crossing(1, t_plain, 2)
    out = {
        "workload": f"room {info['triangles']} triangles, 5 % window, sky + 0.5 deg sun, {W}x{H}, depth 8",
        "plain_spp": spp,
        "plain_tail_per_cu": plain_tail,
        "plain_s": t_plain,
        "plain_ms_per_frame": t_plain * 1e3 / spp,
        "plain_rmse": mse_p ** 0.5,
        "guided_warmup_frames": warm,
        "guided_ms_per_frame": curve[-1][1] * 1e3 / (warm + curve[-1][0]),
        "guided_s_to_plain_rmse": t_equal_rmse,
        "time_ratio": (t_equal_rmse / t_plain) if t_equal_rmse else None,
        "guided_rmse_at_plain_time": mse_equal_time ** 0.5 if mse_equal_time else None,
        "equal_time_rmse_ratio": (mse_equal_time / mse_p) ** 0.5 if mse_equal_time else None,
        "reference_spp": n_ref,
        "reference_s": t_ref,
        "curve_frames_s_rmse": [(n, t, m ** 0.5) for n, t, m in curve],
        "guided_stage_ms": prof["gpuMs"],
        "train_steps": st["steps"],
        "last_loss": st["lastLoss"],
        "mlp": _mlp_roofline(st["lastFrameRows"], prof["gpuMs"]["guide"], st["lastBatch"] * opts["stepsPerFrame"],
                             prof["gpuMs"]["train"]),
    }
    if exp is not None:
        crop = _centre_crop(W, H, 64, 36)
        x0, y0, cw, ch = crop
        t0 = time.perf_counter()
        fp64 = _cpu_fp64_reference(exp, sky, info["camera"], W, H, crop, 4096, 8)
        out["fp64_crop"] = {
            "crop": list(crop), "reference_spp": 4096, "cpu_s": time.perf_counter() - t0,
            "plain_rmse": _rmse(img_p[y0:y0 + ch, x0:x0 + cw], fp64),
            "plain_reference_rmse": _rmse(img_r[y0:y0 + ch, x0:x0 + cw], fp64),
            "guided_rmse_last_checkpoint": _rmse(img_g_last[y0:y0 + ch, x0:x0 + cw], fp64),
            "guided_frames_last_checkpoint": curve[-1][0],
        }
        _log(f"guided: FP64 crop RMSE plain {spp} spp {out['fp64_crop']['plain_rmse']:.4g}, plain reference "
             f"{out['fp64_crop']['plain_reference_rmse']:.4g}, guided {curve[-1][0]} spp "
             f"{out['fp64_crop']['guided_rmse_last_checkpoint']:.4g}")
    _log(f"guided: plain {spp} spp in {t_plain:.2f} s (RMSE {out['plain_rmse']:.4g}); guided reaches it in "
         f"{t_equal_rmse} s -> ratio {out['time_ratio']}; equal-time RMSE ratio {out['equal_time_rmse_ratio']}")
    return out


# -------------------------------------------------------------------- batched --
def batch_workload(args, torch) -> dict:
    """Batched plain rendering: 64 seeded requests (resolution log-uniform
    from 256x256 to 1920x1080, terrain-plus-spheres scenes of 10k-2M
    triangles, 1-16 spp, depth 8), launch shapes picked per request by the
    autotuner, all rendered by one mrtxRenderBatch call (one graph launch).
    The same engines then render the same frames one mrtRenderFrame call at a
    time; the images must be identical."""
    from miniraytrace_amd import scenes
    from miniraytrace_amd.batch import render_batch
    from miniraytrace_amd.engine import Engine
    reqs = scenes.batch_requests(args.batch_requests, 1000)
    sky = scenes.sky()
    t0 = time.perf_counter()
    engines = []
    for r in reqs:
        e = Engine(width=r["width"], height=r["height"])
        e.set_render_settings(width=r["width"], height=r["height"], max_bounces=8, spp_limit=1 << 20,
                              firefly_clamp=0.0, denoise=False)
        r["camera"] = scenes.load_terrain_scene(e, scenes.terrain_config(r["triangles"]), seed=r["seed"],
                                                sky_img=sky)["camera"]
        e.commit()
        engines.append(e)
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    shapes = [e.autotune_path() for e in engines]
    t_tune = time.perf_counter() - t0
    _log(f"batch: {len(reqs)} requests, setup {t_setup:.1f} s, autotune {t_tune:.1f} s")
    ray_counts = [_profiled_rays(e) for e in engines]
    counts = None
    if not args.no_cpu_baseline:
        t0 = time.perf_counter()
        counts = [_cpu_fetch_counts(e.export_scene(), sky, r["camera"], r["width"], r["height"],
                                    _centre_crop(r["width"], r["height"], 32, 18), 8)
                  for e, r in zip(engines, reqs)]
        _log(f"batch: reference-order fetch counts of {len(reqs)} crops in {time.perf_counter() - t0:.1f} s")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = render_batch(engines, [r["spp"] for r in reqs])
    t_batch = time.perf_counter() - t0
    imgs = [e.read_rgba32f(r["width"], r["height"]) for e, r in zip(engines, reqs)]
    for e in engines:
        e.reset_accumulation()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for e, r in zip(engines, reqs):
        for _ in range(r["spp"]):
            e.render_frame()
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t0
    identical = all(np.array_equal(e.read_rgba32f(r["width"], r["height"]), im)
                    for e, r, im in zip(engines, reqs, imgs))
    for e in engines:
        e.close()
    samples = sum(r["width"] * r["height"] * r["spp"] for r in reqs)
    shape_counts = {}
    for sh in shapes:
        k = f"extend {sh['extend'][0]}x{sh['extend'][1]} / shadow {sh['shadow'][0]}x{sh['shadow'][1]}"
        shape_counts[k] = shape_counts.get(k, 0) + 1
    out = {
        "workload": f"{len(reqs)} requests, 256x256..1920x1080, 10k..2M triangles, 1..16 spp, depth 8",
        "requests": len(reqs),
        "samples": samples,
        "triangles_total": sum(r["triangles"] for r in reqs),
        "batch_s": t_batch,
        "batch_capture_ms": st["captureMs"],
        "batch_launch_ms": st["launchMs"],
        "msamples_s_batch": samples / t_batch / 1e6,
        "sequential_s": t_seq,
        "msamples_s_sequential": samples / t_seq / 1e6,
        "speedup_vs_sequential": t_seq / t_batch,
        "identical_to_sequential": identical,
        "autotune_s": t_tune,
        "setup_s": t_setup,
        "shapes": shape_counts,
    }
    if counts is not None:
        per_req = [_trace_bytes(c, rr, ss)["bytes_per_frame"] for c, (rr, ss) in zip(counts, ray_counts)]
        total = sum(b * r["spp"] for b, r in zip(per_req, reqs))
        achieved = total / (st["launchMs"] * 1e-3) / 1e9
        out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                           "kernel": "the whole batched graph (all requests' trace, shade, tail, generate and "
                                     "resolve launches): a lower bound for the trace kernels",
                           "bytes": total, "launch_ms": st["launchMs"],
                           "bytes_per_request_frame_min": min(per_req), "bytes_per_request_frame_max": max(per_req)}
    _log(f"batch: {samples / 1e6:.1f} Msamples in {t_batch:.2f} s (capture {st['captureMs']:.0f} ms) vs "
         f"sequential {t_seq:.2f} s; identical {identical}"
         + (f"; {out['roofline']['achieved']:.0f} GB/s algorithmic over the graph" if "roofline" in out else ""))
    return out


def copy_workload(args, torch) -> dict:
    """Plain device copy of a known size (1 GiB), used to calibrate the HBM
    byte counters (FETCH_SIZE / WRITE_SIZE) against known traffic."""
    n = 1 << 28  # 2^28 float32 = 1 GiB
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(args.warmup):
        dst.copy_(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dst.copy_(src)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    return {"workload": "device copy 1 GiB", "ms_per_step": dt * 1e3,
            "bytes_read": 4 * n, "bytes_written": 4 * n, "gbs": 8 * n / dt / 1e9}


# ------------------------------------------------------- guided + batched at scale --
def _scale_engine(k, sky, seed):
    """Request k of the scale workload: scene by parity (sphere field / room),
    resolution by (k // 2) % 2 (3840x2160 / 1920x1080), max depth 12."""
    from miniraytrace_amd import scenes
    from miniraytrace_amd.engine import Engine
    cfg = scenes.CONFIG5
    W, H = cfg["resolutions"][(k // 2) % 2]
    e = Engine(width=W, height=H)
    e.set_render_settings(width=W, height=H, max_bounces=cfg["max_depth"], spp_limit=1 << 30,
                          firefly_clamp=0.0, denoise=False)
    if k % 2 == 0:
        info = scenes.load_sphere_field(e, cfg, seed=1 + k // 2, sky_img=sky)
    else:
        info = scenes.load_room_scene(e, scenes.CONFIG4, seed=1, sky_img=sky)
    e.set_path_options(seed=seed)
    e.commit()
    return e, W, H, info


def _mse(a, b) -> float:
    return float(((a - b) ** 2).mean())


def scale_workload(args, torch, frames: int = 8, warm: int = 32, spp: int = 32) -> dict:
    """Guided and batched together at scale: 16 requests, alternating between
    the field of 256 instanced 40,500-triangle displaced spheres (10.4M
    triangles) and the window-lit room, at 3840x2160 and 1920x1080, max depth
    12.  The same engines render `frames` frames per request through one
    mrtxRenderBatch call, first plain (requests share one set of kernels per
    frame), then each with its own online-trained guiding network (after
    `warm` training frames whose samples are discarded).

    Equal-RMSE speedup of guided over plain batched rendering: per-sample MSE
    ratio (guided / plain at `spp` samples per pixel, after `warm` training
    frames, against a plain render of 4 x spp samples, the unbiased
    estimator) measured on the 1920x1080 field and room requests, times the
    ratio of batched throughputs.  The reference's own error adds the same
    amount to both MSEs: plain's is sigma^2 (1/spp + 1/(4 spp)), so that
    amount is mse_plain / 5, and it is subtracted from both before the ratio
    (the raw ratio is reported too)."""
    from miniraytrace_amd import scenes
    from miniraytrace_amd.batch import render_batch
    cfg = scenes.CONFIG5
    sky = scenes.sky()
    t0 = time.perf_counter()
    engines = [_scale_engine(k, sky, 100 + k) for k in range(cfg["requests"])]
    t_setup = time.perf_counter() - t0
    samples = sum(W * H for _, W, H, _ in engines) * frames
    _log(f"scale: {len(engines)} requests, setup {t_setup:.1f} s, {samples / frames / 1e6:.1f} Mpixels per frame")
    es = [e for e, _, _, _ in engines]
    ray_counts = [_profiled_rays(e) for e in es]
    counts = None
    if not args.no_cpu_baseline:
        counts = []
        for k, (e, W, H, info) in enumerate(engines):
            counts.append(_cpu_fetch_counts(e.export_scene(), sky, info["camera"], W, H,
                                            _centre_crop(W, H, 48, 27), cfg["max_depth"]))
    render_batch(es, [1] * len(es))  # first use: graph capture and caches
    for e in es:
        e.reset_accumulation()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st_plain = render_batch(es, [frames] * len(es))
    t_plain = time.perf_counter() - t0
    _log(f"scale: plain batched {samples / t_plain / 1e6:.1f} Msamples/s (mode {st_plain['mode']})")
    for e in es:
        e.reset_accumulation()
        e.set_guiding(sampling=True, training=True)
    render_batch(es, [warm] * len(es))
    for e in es:
        e.reset_accumulation()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st_guided = render_batch(es, [frames] * len(es))
    t_guided = time.perf_counter() - t0
    _log(f"scale: guided batched {samples / t_guided / 1e6:.1f} Msamples/s (mode {st_guided['mode']})")
    for e in es:
        e.close()
    engines = es = None

    # Per-sample MSE ratio on the two 1920x1080 requests (k = 2: field, 3: room).
    per_request = {}
    for k in (2, 3):
        ref, W, H, _ = _scale_engine(k, sky, 900 + k)
        for _ in range(4 * spp):
            ref.render_frame()
        img_r = ref.read_rgba32f(W, H)[..., :3].astype(np.float64)
        ref.close()
        plain, _, _, _ = _scale_engine(k, sky, 700 + k)
        for _ in range(spp):
            plain.render_frame()
        mse_p = _mse(plain.read_rgba32f(W, H)[..., :3].astype(np.float64), img_r)
        plain.close()
        g, _, _, _ = _scale_engine(k, sky, 800 + k)
        g.set_guiding(sampling=True, training=True)
        for _ in range(warm):
            g.render_frame()
        g.reset_accumulation()
        for _ in range(spp):
            g.render_frame()
        mse_g = _mse(g.read_rgba32f(W, H)[..., :3].astype(np.float64), img_r)
        g.close()
        name = "field" if k % 2 == 0 else "room"
        ref_share = mse_p / 5.0  # the 4 x spp reference's own error
        ratio = (mse_g - ref_share) / (mse_p - ref_share)
        per_request[name] = {"spp": spp, "reference_spp": 4 * spp, "reference": "plain",
                             "mse_plain": mse_p, "mse_guided": mse_g, "mse_ratio_raw": mse_g / mse_p,
                             "reference_mse_share": ref_share, "mse_ratio": ratio}
        _log(f"scale: {name} 1080p per-sample MSE ratio guided / plain {ratio:.3f} "
             f"(raw {mse_g / mse_p:.3f}, before the reference's share is removed)")
    thr_ratio = t_plain / t_guided  # guided samples/s over plain samples/s
    speedups = {n: thr_ratio / r["mse_ratio"] for n, r in per_request.items()}
    out = {
        "workload": "16 requests: 256 x 40,500-tri displaced spheres (10.4M tris) / 495k-tri room, "
                    "3840x2160 and 1920x1080, depth 12, per-request guiding network",
        "requests": cfg["requests"],
        "frames": frames,
        "samples": samples,
        "setup_s": t_setup,
        "plain_batched_s": t_plain,
        "plain_msamples_s": samples / t_plain / 1e6,
        "plain_mode": st_plain["mode"],
        "guided_batched_s": t_guided,
        "guided_msamples_s": samples / t_guided / 1e6,
        "guided_mode": st_guided["mode"],
        "guided_warmup_frames": warm,
        "per_sample_mse": per_request,
        "equal_rmse_speedup": speedups,
    }
    if counts is not None:
        total = frames * sum(_trace_bytes(c, rr, ss)["bytes_per_frame"] for c, (rr, ss) in zip(counts, ray_counts))
        achieved = total / (st_plain["launchMs"] * 1e-3) / 1e9
        out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                           "kernel": "the whole plain batched graph (trace, shade, tail, generate and resolve "
                                     "launches of all 16 requests): a lower bound for the trace kernels",
                           "bytes": total, "launch_ms": st_plain["launchMs"]}
    return out


def bvh_workload(args, torch, reps: int = 5) -> dict:
    """BLAS build of the 1,000,000-triangle terrain (1000 x 500 quads): the
    GPU builder (mrtxBuildBvhDevice, bounds and outputs in HBM) against the
    threaded CPU builder (mrtxBuildBvh, the same nodes), and the engine's
    commit of that mesh (bounds on the host, GPU build, packing, upload)
    against the reference's budget of 300 ms per 1M triangles (PRD.md:172).
    Algorithmic bytes per build: the bounds read once (24 B per triangle),
    the order written (4 B) and the nodes written (32 B each)."""
    from miniraytrace_amd import scenes
    from miniraytrace_amd.bvh import build_bvh, build_bvh_device
    from miniraytrace_amd.engine import Engine
    pos, idx = scenes.terrain(1000, 500, 20.0, 10.0, 0.2, 1)
    tri = pos[idx.reshape(-1, 3)]
    b = np.concatenate([tri.min(1), tri.max(1)], 1).astype(np.float32)
    n = b.shape[0]
    d = torch.from_numpy(b).cuda()
    build_bvh_device(d, 4)  # workspace allocation
    ms = []
    for _ in range(reps):
        t = {}
        gn, go = build_bvh_device(d, 4, timing=t)
        ms.append(t["ms"])
    t0 = time.perf_counter()
    cn, co = build_bvh(b, 4)
    cpu_ms = (time.perf_counter() - t0) * 1e3
    identical = bool(gn.tobytes() == cn.tobytes() and np.array_equal(go, co))
    e = Engine(width=64, height=64)
    m = e.add_mesh(pos, idx)
    e.add_instance(m, None, 0)
    commit_ms = e.commit()["commitMs"]
    e.close()
    gpu_ms = float(np.median(ms))
    traffic = n * (24 + 4) + len(gn) * 32
    gbs = traffic / (gpu_ms * 1e-3) / 1e9
    _log(f"bvh: {n} triangles, GPU {gpu_ms:.2f} ms (median of {reps}), CPU {cpu_ms:.1f} ms, "
         f"identical {identical}, engine commit {commit_ms:.1f} ms")
    return {"workload": "BLAS build, 1,000,000-triangle terrain", "triangles": n, "nodes": int(len(gn)),
            "gpu_ms": round(gpu_ms, 3), "gpu_ms_all": [round(x, 3) for x in ms],
            "identical_to_cpu_builder": identical, "engine_commit_ms": round(commit_ms, 2),
            "commit_budget_ms": 300.0,
            "roofline": {"bound": "hbm", "achieved": round(gbs, 2), "peak": 8000.0, "unit": "GB/s",
                         "frac": round(gbs / 8000.0, 5), "traffic": None,
                         "note": "whole build: dependent levels with a host wait each, then one wavefront per "
                                 "subtree; latency-bound, far from the byte bound"},
            "cpu_baseline": {"value": round(cpu_ms, 2), "unit": "ms", "cores": os.cpu_count(), "kind": "port",
                             "sample": "the whole 1M-triangle build, product CPU builder (threaded)"}}


def edit_workload(args, torch, reps: int = 5) -> dict:
    """Edit-to-frame latency on the 1080p terrain scene (1,256,000
    triangles, depth 8): wall time from the edit call through the commit to
    the end of the next 1-spp frame, against the 100 ms target of the
    reference's interactive integration (docs/rhino-integration-design.md:30).
    Edits: one sphere instance moved (TLAS only); the sphere mesh (1,280
    triangles, 200 instances) deformed in place; the 1M-triangle terrain
    deformed in place (GPU BLAS build); a new mesh and instance appended."""
    from miniraytrace_amd import scenes
    from miniraytrace_amd.engine import Engine
    W, H = 1920, 1080
    e = Engine(width=W, height=H)
    e.set_render_settings(width=W, height=H, max_bounces=8, spp_limit=1 << 20, firefly_clamp=0.0, denoise=False)
    info = scenes.load_terrain_scene(e, scenes.CONFIG2, seed=1, sky_img=scenes.sky())
    e.commit()
    for _ in range(3):
        e.render_frame()
    cfg = scenes.CONFIG2
    tpos, tidx = scenes.terrain(cfg["nx"], cfg["nz"], cfg["sx"], cfg["sz"], cfg["amplitude"], 1)
    spos, sidx = scenes.icosphere(cfg["sphere_level"])

    def timed(edit):
        torch.cuda.synchronize()
        before = e.frame_stats()
        t0 = time.perf_counter()
        edit()
        tc = time.perf_counter()
        e.commit()
        t1 = time.perf_counter()
        e.render_frame()
        t2 = time.perf_counter()
        st = e.frame_stats()
        return {"edit_to_frame_ms": (t2 - t0) * 1e3, "edit_call_ms": (tc - t0) * 1e3, "commit_ms": (t1 - tc) * 1e3,
                "frame_ms": (t2 - t1) * 1e3,
                "upload_bytes": st["commitUploadBytes"],
                "repacks": st["geometryRepacks"] - before["geometryRepacks"],
                "meshes_written": st["meshesWritten"] - before["meshesWritten"]}

    def summary(runs):
        out = {k: round(float(np.median([r[k] for r in runs])), 3)
               for k in ("edit_to_frame_ms", "edit_call_ms", "commit_ms", "frame_ms")}
        out.update(upload_bytes=int(runs[-1]["upload_bytes"]), repacks=int(sum(r["repacks"] for r in runs)),
                   meshes_written=int(sum(r["meshes_written"] for r in runs)), runs=len(runs))
        return out

    res = {"workload": "edit-to-frame latency, 1080p terrain scene (1,256,000 triangles), depth 8",
           "target_ms": 100.0}
    inst = info["instance_ids"][1]
    xf0 = np.asarray(info["sphere_xforms"][0], np.float32).reshape(3, 4).copy()

    def move(k):
        xf = xf0.copy()
        xf[:, 3] += np.float32(0.05 * (k + 1))
        return lambda: e.set_instance_transform(inst, xf.reshape(12))

    res["instance_transform"] = summary([timed(move(k)) for k in range(reps)])

    def deform_sphere(k):
        v = (spos * (1 + 0.05 * (k + 1) * np.sin(4 * spos[:, :1]))).astype(np.float32)
        return lambda: e.update_mesh(info["sphere_mesh"], v, sidx, normals=v)

    res["sphere_mesh_update"] = summary([timed(deform_sphere(k)) for k in range(reps)])

    def deform_terrain(k):
        p = tpos.copy()
        p[:, 1] += np.float32(0.02 * (k + 1)) * np.sin(p[:, 0]).astype(np.float32)
        return lambda: e.update_mesh(info["terrain_mesh"], p, tidx)

    res["terrain_mesh_update"] = summary([timed(deform_terrain(k)) for k in range(3)])

    def add_mesh(k):
        def go():
            m = e.add_mesh(spos * np.float32(0.5), sidx, normals=spos)
            xf = xf0.copy()
            xf[1, 3] += np.float32(1.0 + 0.3 * k)
            e.add_instance(m, xf.reshape(12), info["material_ids"][1])
        return go

    res["mesh_added"] = summary([timed(add_mesh(k)) for k in range(3)])
    e.close()
    _log("edit: " + ", ".join(f"{k} {v['edit_to_frame_ms']:.1f} ms (commit {v['commit_ms']:.1f})"
                              for k, v in res.items() if isinstance(v, dict)))
    return res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="pt",
                    choices=["pt", "traversal", "guided", "batch", "scale", "converge", "copy", "bvh", "edit"])
    ap.add_argument("--guided-spp", type=int, default=1024,
                    help="plain sample count whose RMSE the guided run must reach")
    ap.add_argument("--guided-warmup", type=int, default=0,
                    help="guided frames rendered (and timed) before its samples count")
    ap.add_argument("--no-guided", dest="with_guided", action="store_false", default=True)
    ap.add_argument("--batch-requests", type=int, default=64)
    ap.add_argument("--converge-spp", type=int, default=1024)
    ap.add_argument("--no-converge", dest="with_converge", action="store_false", default=True)
    ap.add_argument("--no-scale", dest="with_scale", action="store_false", default=True)
    ap.add_argument("--no-batch", dest="with_batch", action="store_false", default=True)
    ap.add_argument("--no-bvh", dest="with_bvh", action="store_false", default=True)
    ap.add_argument("--no-edit", dest="with_edit", action="store_false", default=True)
    ap.add_argument("--autotune", action="store_true", default=True)
    ap.add_argument("--no-autotune", dest="autotune", action="store_false")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--with-traversal", action="store_true", default=True)
    ap.add_argument("--no-traversal", dest="with_traversal", action="store_false")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        return 2

    import torch
    if not torch.cuda.is_available():
        print("no HIP device", file=sys.stderr)
        return 1
    torch.cuda.synchronize()

    if args.workload == "bvh":
        res = bvh_workload(args, torch)
        print(json.dumps({"metric": "BLAS build ms, 1M-triangle terrain", "value": res["gpu_ms"], "unit": "ms",
                          "n_gpus": 1, "higher_is_better": False, "dtype": "f32", "data": "synthetic",
                          "config": {"workload": res["workload"]}, "roofline": res["roofline"],
                          "cpu_baseline": res["cpu_baseline"], "bvh": res}))
        return 0
    if args.workload == "edit":
        res = edit_workload(args, torch)
        print(json.dumps({"metric": "edit-to-frame ms, terrain mesh update", "value":
                          res["terrain_mesh_update"]["edit_to_frame_ms"], "unit": "ms", "n_gpus": 1,
                          "higher_is_better": False, "dtype": "f32", "data": "synthetic",
                          "config": {"workload": res["workload"]}, "edit": res}))
        return 0
    if args.workload == "copy":
        res = copy_workload(args, torch)
        print(json.dumps({"workload": "copy", **res}))
        return 0
    if args.workload == "batch":
        res = batch_workload(args, torch)
        print(json.dumps({"metric": "batched Msamples/s", "value": res["msamples_s_batch"], "unit": "Msamples/s",
                          "n_gpus": 1, "higher_is_better": True, "dtype": "f32", "data": "synthetic",
                          "config": {"workload": res["workload"]}, "batch": res}))
        return 0
    if args.workload == "converge":
        from miniraytrace_amd import scenes
        res = convergence_workload(args, torch, scenes.sky())
        print(json.dumps({"metric": "RMSE vs FP64 at 1024 spp, 1080p terrain crop", "value": res["rmse"][-1],
                          "unit": "rmse", "n_gpus": 1, "higher_is_better": False, "dtype": "f32",
                          "data": "synthetic", "config": {"workload": res["workload"]}, "convergence": res}))
        return 0
    if args.workload == "scale":
        res = scale_workload(args, torch)
        print(json.dumps({"metric": "guided + batched Msamples/s, 16 requests at 4K / 1080p", "value":
                          res["guided_msamples_s"], "unit": "Msamples/s", "n_gpus": 1, "higher_is_better": True,
                          "dtype": "f32", "data": "synthetic", "config": {"workload": res["workload"]},
                          "scale": res}))
        return 0
    if args.workload == "guided":
        res = guided_workload(args, torch)
        print(json.dumps({"metric": "equal-RMSE wall time, guided / plain", "value": res["time_ratio"],
                          "unit": "ratio", "n_gpus": 1, "higher_is_better": False, "dtype": "f32",
                          "data": "synthetic", "config": {"workload": res["workload"]}, "guided": res}))
        return 0
    if args.workload == "traversal":
        res = traversal_workload(args, torch)
        line = {"metric": "closest-hit Mrays/s, 10k-triangle soup, 1M rays", "value": res["value_mrays_s"],
                "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                "ms_per_step": res["ms_per_step"], "higher_is_better": True, "scaling": "weak",
                "vs_baseline": None, "dtype": "f32", "data": "synthetic",
                "config": {"workload": res["workload"]}, "traversal": res}
        for k in ("roofline", "cpu_baseline"):
            if k in res:
                line[k] = res[k]
        print(json.dumps(line))
        return 0

    res = pt_workload(args, torch)
    line = {
        "metric": METRIC,
        "value": res["msamples_s"],
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["ms_per_frame"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": res["workload"], "width": 1920, "height": 1080, "spp_per_frame": 1,
                   "max_depth": 8},
        "pt": res,
    }
    for k in ("roofline", "cpu_baseline"):
        if k in res:
            line[k] = res.pop(k)
    if args.with_traversal:
        trav = traversal_workload(args, torch)
        line["traversal"] = trav
    if args.with_batch:
        line["batch"] = batch_workload(args, torch)
    if args.with_guided:
        gd = guided_workload(args, torch)
        line["guided"] = gd
        line["equal_rmse_time_ratio"] = gd["time_ratio"]
    if args.with_converge and not args.no_cpu_baseline:
        from miniraytrace_amd import scenes
        line["convergence"] = convergence_workload(args, torch, scenes.sky())
    if args.with_scale:
        line["scale"] = scale_workload(args, torch, frames=4, warm=32, spp=16)
    if args.with_bvh:
        line["bvh"] = bvh_workload(args, torch)
    if args.with_edit:
        line["edit"] = edit_workload(args, torch)
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
