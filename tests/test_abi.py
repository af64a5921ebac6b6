"""C ABI checks that need no GPU.

* include/mini_raytrace.h has the reference layout: every struct size and
  field offset and every enum value equals tests/golden/abi_layout.json
  (generated from the reference header by tests/golden/gen_abi_layout.py).
* The ctypes mirror in miniraytrace_amd/_lib.py has the same sizes.
* libmrt.so loads and exports every MRT_API function declared in include/*.h.
* Argument validation of the reference capi/src/capi.cpp:93-316 that is
  decided before any device work.
"""
import ctypes as C
import json
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = json.loads((ROOT / "tests" / "golden" / "abi_layout.json").read_text())["layout"]


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        names.update(re.findall(r"MRT_API\s+[^;(]*?\b(mrtx?\w+)\s*\(", text))
    return sorted(names)


def test_header_layout_matches_reference(tmp_path):
    from tests.golden.gen_abi_layout import layout_of
    ours = layout_of(ROOT / "include" / "mini_raytrace.h", tmp_path)
    assert ours == GOLDEN


def test_ctypes_mirror_sizes():
    from miniraytrace_amd import _lib as L
    for name in ["mrtEngineDesc", "mrtMeshDesc", "mrtMaterialDesc", "mrtTextureDesc", "mrtLightDesc",
                 "mrtCameraDesc", "mrtCameraDescEx", "mrtRenderSettings", "mrtFrameInfo", "mrtCommitStats"]:
        cls = getattr(L, name)
        assert C.sizeof(cls) == GOLDEN[name]["size"], name
        for f, _ in cls._fields_:
            assert getattr(cls, f).offset == GOLDEN[name][f], (name, f)
    # the C# smoke test's struct-size asserts (rhino/MiniRaytrace.Smoke/Program.cs:28-30)
    assert C.sizeof(L.mrtCameraDescEx) == 64
    assert C.sizeof(L.F3) == 12


def test_library_exports_every_declared_symbol():
    from miniraytrace_amd import _lib as L
    lib = L.lib()
    names = declared_functions()
    assert len([n for n in names if not n.startswith("mrtx")]) == 27
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", str(L.LIB_PATH)], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (mrtx?\w+)", out))
    assert set(names) <= exported


def test_every_declared_symbol_is_bound_in_python():
    from miniraytrace_amd import _lib as L
    assert set(declared_functions()) <= set(L._SIGS)


def test_argument_validation_without_device():
    from miniraytrace_amd import _lib as L
    lib = L.lib()
    h = C.c_void_p()
    assert lib.mrtCreateEngine(None, C.byref(h)) == L.MRT_ERROR_INVALID_ARGUMENT
    short = L.mrtEngineDesc()
    short.structSize = C.sizeof(short) - 4
    assert lib.mrtCreateEngine(C.byref(short), C.byref(h)) == L.MRT_ERROR_INVALID_ARGUMENT
    assert lib.mrtCreateEngine(C.byref(L.mrtEngineDesc()), None) == L.MRT_ERROR_INVALID_ARGUMENT
    assert lib.mrtCommitScene(None, None) == L.MRT_ERROR_INVALID_HANDLE
    assert lib.mrtRenderFrame(None, None) == L.MRT_ERROR_INVALID_HANDLE
    buf = (C.c_uint8 * 16)()
    assert lib.mrtReadFramebuffer(None, 0, buf, 16) == L.MRT_ERROR_INVALID_ARGUMENT
    assert lib.mrtSetRenderSettings(None, None) == L.MRT_ERROR_INVALID_ARGUMENT
    assert lib.mrtAddMesh(None, None) == L.MRT_INVALID_ID
    assert lib.mrtAddMaterial(None, None) == L.MRT_INVALID_ID
    assert lib.mrtAddInstance(None, 1, None, 0) == L.MRT_INVALID_ID
    lib.mrtDestroyEngine(None)  # safe
    lib.mrtRemoveMesh(None, 3)
    lib.mrtResetAccumulation(None)
    assert lib.mrtGetLastErrorMessage() is not None


def test_result_strings():
    from miniraytrace_amd import _lib as L
    lib = L.lib()
    names = {0: b"Success", 2: b"ErrorVulkanInit", 3: b"ErrorOutOfMemory", 4: b"ErrorInvalidArgument",
             5: b"ErrorInvalidHandle", 6: b"ErrorShaderLoad", 7: b"ErrorDeviceLost", 1: b"ErrorUnknown",
             99: b"ErrorUnknown"}
    for code, s in names.items():
        assert lib.mrtResultToString(code) == s


def test_out_of_range_device_index():
    """gpuIndex1 past the device count fails with MRT_ERROR_VULKAN_INIT and a
    message (tests/gpu_selection_tests.cpp:72-80); with no device at all the
    same code is returned."""
    from miniraytrace_amd import _lib as L
    lib = L.lib()
    d = L.mrtEngineDesc(width=16, height=16, gpuIndex1=9999)
    h = C.c_void_p()
    r = lib.mrtCreateEngine(C.byref(d), C.byref(h))
    assert r == L.MRT_ERROR_VULKAN_INIT
    msg = lib.mrtGetLastErrorMessage().decode()
    assert "out of range" in msg or "no HIP device" in msg
