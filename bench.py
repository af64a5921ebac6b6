#!/usr/bin/env python3
"""Benchmark driver.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Workloads:
  traversal  closest-hit queries: 10,000-triangle seeded soup, 1,000,000
             rays (one step = one mrtxTraceClosest launch over all rays).

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
        blk, stk = qs.autotune(d_rays, hits)
    else:
        blk, stk = qs.config()
    _log(f"traversal: {info['nodeCount']} nodes, depth {info['maxDepth']}, shape {blk}x{stk}")

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
        "shape": [blk, stk],
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
        out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
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


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="traversal", choices=["traversal", "copy"])
    ap.add_argument("--autotune", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
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

    if args.workload == "copy":
        res = copy_workload(args, torch)
        print(json.dumps({"workload": "copy", **res}))
        return 0
    res = traversal_workload(args, torch)
    line = {
        "metric": METRIC,
        "value": None,
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": res["workload"]},
        "traversal": res,
    }
    for k in ("roofline", "cpu_baseline"):
        if k in res:
            line[k] = res[k]
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
