"""Pins the FP64 oracle path tracer (oracle/pathtracer.cpp) on the CPU.

* Emission is reproduced exactly with maxBounces = 0 (the reference's own
  determinism trick, tests/camera_tests.cpp:24-51).
* White furnace: a Lambert+GGX sphere under a constant environment renders
  its directional albedo, computed here by deterministic quadrature of the
  reference BSDF value (bsdf.glsl:100-123) -- this checks the estimator
  (light/BSDF sampling pdfs, MIS weights, Russian roulette) independently of
  how the oracle samples.
* The environment CDF sampler's texel probabilities equal luminance x
  sin(theta) normalised (closed form for a constant map).
"""
import numpy as np
import pytest

from tests.scene_pack import material, pack


def quad_scene(emission):
    s = 1000.0
    pos = np.array([[-s, -s, -20], [s, -s, -20], [s, s, -20], [-s, s, -20]], np.float32)
    return pack(pos, [0, 1, 2, 0, 2, 3], materials=[material(base=(0, 0, 0), emission=emission)],
                instances=[(None, 1)], env=False)


def test_emission_exact(oracle):
    sc = quad_scene((1.0, 0.5, 0.25))
    cam = oracle.camera_lookat((0, 0, 0), (0, 0, -1), (0, 1, 0), 90.0, 16, 16)
    img, _ = oracle.pt_render(sc, None, cam, 16, 16, spp=3, max_bounces=0)
    assert np.all(img[..., 0] == 3.0) and np.all(img[..., 1] == 1.5) and np.all(img[..., 2] == 0.75)


def _albedo_quadrature(base, rough, metal, n=1200):
    """Directional albedo at normal incidence of the reference uber-BSDF
    (reflection half, transmission 0): integral of f(wo, wi) cos(theta_i)."""
    alpha = max(rough * rough, 1e-3)
    th = (np.arange(n) + 0.5) / n * (np.pi / 2)
    ci, si = np.cos(th), np.sin(th)
    wo = np.array([0, 0, 1.0])
    # wi in the xz plane suffices (isotropic, wo along the normal)
    wi = np.stack([si, np.zeros_like(si), ci], 1)
    h = wi + wo
    h /= np.linalg.norm(h, axis=1, keepdims=True)
    a2 = alpha * alpha
    d = (h[:, 2] ** 2) * (a2 - 1) + 1
    D = a2 / np.maximum(np.pi * d * d, 1e-12)

    def lam(c):
        c2 = c * c
        t2 = np.maximum(1 - c2, 0) / np.maximum(c2, 1e-12)
        return 0.5 * (np.sqrt(1 + a2 * t2) - 1)
    G2 = 1 / (1 + lam(np.ones_like(ci)) + lam(ci))
    f0 = 0.04 * (1 - metal) + base * metal
    m = np.clip(1 - (h @ wo), 0, 1)
    F = f0 + (1 - f0) * m ** 5
    spec = D * G2 * F / np.maximum(4 * 1.0 * ci, 1e-7)
    diff = (1 - metal) * base / np.pi
    f = diff + spec
    return float(np.sum(f * ci * si) * (np.pi / 2 / n) * 2 * np.pi)


@pytest.mark.parametrize("rough,metal,base", [(1.0, 0.0, 0.5), (0.5, 0.0, 0.7), (0.6, 1.0, 0.6)])
def test_white_furnace_matches_quadrature(oracle, rough, metal, base):
    from miniraytrace_amd import scenes
    v, f = scenes.icosphere(5)
    sc = pack(v, f, normals=v, materials=[material(base=(base,) * 3, roughness=rough, metallic=metal)],
              instances=[(None, 1)], env=True)
    env = np.ones((64, 128, 4), np.float32)
    W = H = 64
    cam = oracle.camera_lookat((0, 0, 4), (0, 0, 0), (0, 1, 0), 30.0, W, H)
    spp = 256
    img, _ = oracle.pt_render(sc, env, cam, W, H, crop=(30, 30, 4, 4), spp=spp, max_bounces=8, threads=0)
    est = img[..., 0] / spp
    expect = _albedo_quadrature(base, rough, metal)
    # per-pixel standard error from the spread over the 16 pixels
    assert abs(est.mean() - expect) < max(4 * est.std() / np.sqrt(est.size), 2e-3), (est.mean(), expect)


def test_env_cdf_probabilities_closed_form(oracle):
    h, w = 32, 64
    env = np.ones((h, w, 4), np.float32)
    prob, integral = oracle.env_cdf_probs(env)
    s = np.sin((np.arange(h) + 0.5) * np.pi / h)
    expect = np.repeat((s / s.sum() / w)[:, None], w, axis=1)
    # the reference stores its CDFs in float32: each probability is a
    # difference of float32 values in [0, 1], exact to ~2^-23 absolute
    assert np.allclose(prob, expect, rtol=0, atol=2.5e-7)
    assert prob.sum() == pytest.approx(1.0, abs=1e-6)
    assert integral == pytest.approx(4 * np.pi, rel=2e-3)  # midpoint rule of the solid angle


def test_env_cdf_probabilities_random_map(oracle):
    rng = np.random.default_rng(2)
    h, w = 24, 48
    env = rng.exponential(1.0, size=(h, w, 4)).astype(np.float32)
    env[3, 7, :3] = 500.0  # a bright texel
    env[10, :, :3] = 0.0  # a black row
    prob, _ = oracle.env_cdf_probs(env)
    lum = (0.2126 * env[..., 0] + 0.7152 * env[..., 1] + 0.0722 * env[..., 2]).astype(np.float64)
    wgt = lum * np.sin((np.arange(h) + 0.5) * np.pi / h)[:, None]
    expect = wgt / wgt.sum()
    assert np.allclose(prob, expect, rtol=1e-6, atol=2.5e-7)
    assert np.all(prob[10] == 0.0)
