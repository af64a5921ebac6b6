"""Oracle: CPU restatement of the reference resolve pass (RGBA32F
accumulation -> RGBA8 display image).

TEST INFRASTRUCTURE ONLY: only tests/ may import it, as the checker.

Restates core/shaders/resolve.comp:13-55 (accumulation / sample count,
x exposure, Narkowicz ACES when tonemap == 1, clamp to [0, 1e4], sRGB OETF,
rgba8 UNORM store with alpha 1) and the exposure factor exp2(exposureEV)
of core/src/render/Renderer.cpp:759.  Arithmetic is float64; the UNORM
store rounds to nearest (Vulkan's float -> unorm conversion), so the GPU's
float32 result may differ by one code.
"""
from __future__ import annotations

import numpy as np

TONEMAP_LINEAR, TONEMAP_ACES = 0, 1


def aces(x):
    """Narkowicz 2015 fit: clamp(x (a x + b) / (x (c x + d) + e), 0, 1)."""
    a, b, c, d, e = 2.51, 0.03, 2.43, 0.59, 0.14
    x = np.asarray(x, np.float64)
    return np.clip((x * (a * x + b)) / (x * (c * x + d) + e), 0.0, 1.0)


def linear_to_srgb(c):
    """sRGB OETF (IEC 61966-2-1): 12.92 c up to 0.0031308, else
    1.055 c^(1/2.4) - 0.055."""
    c = np.asarray(c, np.float64)
    return np.where(c <= 0.0031308, c * 12.92, 1.055 * np.power(np.maximum(c, 0.0), 1.0 / 2.4) - 0.055)


def resolve_rgba8(accum_rgb_over_spp: np.ndarray, exposure_ev: float = 0.0, tonemap: int = TONEMAP_ACES):
    """(H, W, 3) average radiance -> (H, W, 4) uint8, alpha 255."""
    x = np.asarray(accum_rgb_over_spp, np.float64) * np.exp2(exposure_ev)
    if tonemap == TONEMAP_ACES:
        x = aces(x)
    s = linear_to_srgb(np.clip(x, 0.0, 1e4))
    out = np.empty(x.shape[:-1] + (4,), np.uint8)
    out[..., :3] = np.clip(np.floor(np.clip(s, 0.0, 1.0) * 255.0 + 0.5), 0, 255).astype(np.uint8)
    out[..., 3] = 255
    return out
