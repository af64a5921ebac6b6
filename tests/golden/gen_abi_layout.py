"""Writes tests/golden/abi_layout.json: sizeof / offsetof of every struct
field and the value of every enum constant of a mini_raytrace.h.

Run in a container that holds the reference checkout:

    python tests/golden/gen_abi_layout.py /root/reference/capi/include/mini_raytrace.h

The probe is compiled with gcc (C99) into oracle/_ref/ and run; only its
printed numbers are kept (the fixture).  tests/test_abi.py compiles the same
probe against this project's include/mini_raytrace.h and compares.
"""
from __future__ import annotations

import json
import subprocess
import sys
import tempfile
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
OUT = Path(__file__).resolve().parent / "abi_layout.json"

STRUCTS = {
    "mrtEngineDesc": ["structSize", "enableValidation", "width", "height", "gpuIndex1"],
    "mrtMeshDesc": ["structSize", "vertexCount", "indexCount", "positions", "normals", "uvs", "indices"],
    "mrtMaterialDesc": ["structSize", "baseColor", "opacity", "roughness", "metallic", "transmission", "ior",
                        "emission", "baseColorTex", "roughnessTex", "metallicTex", "normalTex", "emissionTex"],
    "mrtTextureDesc": ["structSize", "width", "height", "format", "pixels", "generateMips"],
    "mrtLightDesc": ["structSize", "type", "direction", "position", "radiance", "angularRadius", "corner",
                     "edge0", "edge1", "twoSided"],
    "mrtCameraDesc": ["structSize", "position", "target", "up", "fovYDeg"],
    "mrtCameraDescEx": ["structSize", "projection", "position", "forward", "up", "left", "right", "bottom", "top"],
    "mrtRenderSettings": ["structSize", "width", "height", "maxBounces", "sppLimit", "exposureEV",
                          "fireflyClamp", "tonemap", "denoise"],
    "mrtFrameInfo": ["structSize", "spp", "converged", "denoised", "frameMs"],
    "mrtCommitStats": ["structSize", "blasRebuilt", "tlasRebuilt", "commitMs"],
}
ENUMS = ["MRT_SUCCESS", "MRT_ERROR_UNKNOWN", "MRT_ERROR_VULKAN_INIT", "MRT_ERROR_OUT_OF_MEMORY",
         "MRT_ERROR_INVALID_ARGUMENT", "MRT_ERROR_INVALID_HANDLE", "MRT_ERROR_SHADER_LOAD",
         "MRT_ERROR_DEVICE_LOST", "MRT_INVALID_ID", "MRT_PIXEL_RGBA8_SRGB", "MRT_PIXEL_RGBA32F_LINEAR",
         "MRT_TEX_RGBA8_SRGB", "MRT_TEX_RGBA8_UNORM", "MRT_TEX_RGBA32F", "MRT_LIGHT_SUN", "MRT_LIGHT_POINT",
         "MRT_LIGHT_RECT", "MRT_TONEMAP_LINEAR", "MRT_TONEMAP_ACES", "MRT_CAMERA_PERSPECTIVE",
         "MRT_CAMERA_ORTHOGRAPHIC"]
TYPEDEF_SIZES = ["mrtResult", "mrtPixelFormat", "mrtTextureFormat", "mrtLightType", "mrtTonemap",
                 "mrtMeshId", "mrtLogFn"]


def probe_source(header: Path) -> str:
    lines = ["#include <stdio.h>", "#include <stddef.h>", f'#include "{header}"', "int main(void) {",
             '  printf("{\\n");']
    for s, fields in STRUCTS.items():
        lines.append(f'  printf("\\"{s}\\": {{\\"size\\": %zu", sizeof({s}));')
        for f in fields:
            lines.append(f'  printf(", \\"{f}\\": %zu", offsetof({s}, {f}));')
        lines.append('  printf("},\\n");')
    for t in TYPEDEF_SIZES:
        lines.append(f'  printf("\\"sizeof_{t}\\": %zu,\\n", sizeof({t}));')
    for i, e in enumerate(ENUMS):
        sep = "," if i + 1 < len(ENUMS) else ""
        lines.append(f'  printf("\\"{e}\\": %lld{sep}\\n", (long long)({e}));')
    lines += ['  printf("}\\n");', "  return 0;", "}"]
    return "\n".join(lines)


def layout_of(header: Path, out_dir: Path) -> dict:
    out_dir.mkdir(parents=True, exist_ok=True)
    src = out_dir / "abi_probe.c"
    exe = out_dir / "abi_probe"
    src.write_text(probe_source(header.resolve()))
    subprocess.run(["gcc", "-std=c99", "-Wall", "-o", str(exe), str(src)], check=True)
    res = subprocess.run([str(exe)], check=True, capture_output=True, text=True)
    return json.loads(res.stdout)


if __name__ == "__main__":
    header = Path(sys.argv[1]) if len(sys.argv) > 1 else Path("/root/reference/capi/include/mini_raytrace.h")
    layout = layout_of(header, ROOT / "oracle" / "_ref")
    OUT.write_text(json.dumps({"source": "capi/include/mini_raytrace.h (reference), gcc -std=c99 x86_64",
                               "layout": layout}, indent=1, sort_keys=True) + "\n")
    print(f"wrote {OUT}")
