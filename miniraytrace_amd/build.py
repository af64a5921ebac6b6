"""Build recipe for the native libraries of miniraytrace_amd.

Outputs (in-tree, next to this file):
  libmrt.so         the renderer: C ABI (include/mini_raytrace.h + ext) over
                    HIP kernels compiled for gfx950
  libmrt_scenes.so  host-only synthetic workload generators

Every call recompiles everything (objects go to miniraytrace_amd/_build/).
Run as `python -m miniraytrace_amd.build`.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
INCLUDE = ROOT / "include"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
          f"-I{INCLUDE}", f"-I{CSRC}", "-fvisibility=hidden"]
# Device code: no packed-FP32 instructions; correctly rounded fp32 divide and
# sqrt (so reciprocal directions match the CPU oracle bit for bit); no
# type-based alias analysis (lanes hand data to each other through LDS with
# different access widths, e.g. bf16 stores read back as 16-byte fragments).
HIP_FLAGS = ["-x", "hip", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-fhip-fp32-correctly-rounded-divide-sqrt", "-munsafe-fp-atomics", "-fno-strict-aliasing"]
# Host code: no FP contraction, so the BVH builder's float arithmetic is the
# reference's operation sequence.
HOST_FLAGS = ["-x", "c++", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"]


def _sources() -> tuple[list[Path], list[Path]]:
    hip = sorted((CSRC / "kernels").glob("*.hip"))
    host = sorted((CSRC / "host").glob("*.cpp"))
    return hip, host


def _compile(src: Path, flags: list[str]) -> Path:
    out = OBJ / (src.parent.name + "_" + src.stem + ".o")
    cmd = [HIPCC, *COMMON, *flags, "-c", str(src), "-o", str(out)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    # Surface real warnings, not the host pass's note about the GPU-only feature.
    warn = [l for l in res.stderr.splitlines() if "warning" in l and "packed-fp32-ops" not in l]
    if warn:
        print("\n".join(warn), file=sys.stderr)
    return out


def build(verbose: bool = True) -> dict[str, Path]:
    if not Path(HIPCC).exists():
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    if OBJ.exists():
        shutil.rmtree(OBJ)
    OBJ.mkdir(parents=True)
    hip, host = _sources()
    jobs = [(s, HIP_FLAGS) for s in hip] + [(s, HOST_FLAGS) for s in host]
    workers = max(1, min(len(jobs), os.cpu_count() or 4, 16))
    with cf.ThreadPoolExecutor(workers) as ex:
        objs = list(ex.map(lambda j: _compile(*j), jobs))

    lib = PKG / "libmrt.so"
    tmp = PKG / "libmrt.so.tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
           *map(str, objs), "-pthread"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, lib)

    scenes = PKG / "libmrt_scenes.so"
    cmd = ["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-o", str(scenes),
           str(CSRC / "scenes" / "scenes.cpp"), "-pthread"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"scenes build failed: {res.stderr}")
    if verbose:
        print(f"built {lib} ({lib.stat().st_size // 1024} KiB) and {scenes.name}")
    return {"mrt": lib, "scenes": scenes}


if __name__ == "__main__":
    build()
