"""Seeded synthetic workloads (libmrt_scenes.so, host-only C++).

Random numbers come from std::mt19937_64 with floats u = (x >> 40) * 2^-24.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

_PATH = Path(__file__).resolve().parent / "libmrt_scenes.so"
_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not _PATH.exists():
            raise RuntimeError(f"{_PATH} is missing; build with `python -m miniraytrace_amd.build`")
        h = C.CDLL(str(_PATH))
        h.mrts_soup.argtypes = [C.c_uint32, C.c_uint64, C.c_void_p]
        h.mrts_soup_rays.argtypes = [C.c_uint64, C.c_uint64, C.c_void_p]
        _lib = h
    return _lib


def soup(count: int = 10_000, seed: int = 1) -> np.ndarray:
    """Triangle soup in the unit cube: (count, 9) float32 (v0, v1, v2)."""
    out = np.empty((count, 9), dtype=np.float32)
    lib().mrts_soup(count, seed, out.ctypes.data)
    return out


def soup_rays(n: int = 1_000_000, seed: int = 2) -> np.ndarray:
    """Rays from a radius-2 sphere around (0.5,0.5,0.5) towards points in the
    unit cube: (n, 8) float32 (ox oy oz tmin dx dy dz tmax)."""
    out = np.empty((n, 8), dtype=np.float32)
    lib().mrts_soup_rays(n, seed, out.ctypes.data)
    return out
