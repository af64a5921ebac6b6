"""Pins the oracle path tracer's transmission, analytic-light, opacity,
texture and normal-map branches (oracle/pathtracer.cpp) on the CPU, each
against what the mathematics fixes rather than against a retyped formula:

* BTDF value: Walter et al. 2007 eq. 21 written with (eta_i, eta_o) and the
  paper's half vector, and its generalised reciprocity f(i,o)/eta_o^2 =
  f(o,i)/eta_i^2 (a misplaced eta or Jacobian breaks it).
* BSDF sampling: chi-square of 400k sampled directions against the density
  the sampler realises, integrated per bin by quadrature, for glass entering
  and exiting, a mixed diffuse/specular/transmissive material, a metal and a
  rough dielectric.  On the refraction side that density is bsdfEval's pdf
  where the half vector faces wo (bsdfEval also gives f and a pdf to pairs
  whose half vector faces away from wo, which sampling never produces:
  bsdf.glsl:130-133 tests only woH * wiH < 0); on the reflection side it is
  bsdfEval's pdf plus the transmission lobe's Fresnel-reflection branch,
  which the reference draws (bsdf.glsl:170-172) but leaves out of
  bsdfEval's pdf (bsdf.glsl:108-123).  Both are reference behaviours the
  oracle and the GPU reproduce.
* Point light (maxBounces 0): every pixel equals f cos I / d^2 at its
  jittered hit point, f from a textbook Lambert + GGX (Walter 2007, Heitz
  2014 height-correlated Smith, Schlick) written here.
* Rect light and sun cone (maxBounces 0): per-sample values are unbiased for
  the quadrature of f cos L over the rect (area form) and over the cone
  (z-test); the rect quadrature itself is checked against Lambert's exact
  polygon form factor.
* Opacity: a quad of opacity p over an emissive background gives exactly
  the front emission (probability p) or the background (1 - p) per sample,
  and at maxBounces 0 the pass-through ends the path (a pass-through uses up
  a bounce, pathtrace.comp:152-156).
* Textures: texel centres, bilinear interpolation of a linear ramp, REPEAT
  wrapping, sRGB decoding (IEC 61966-2-1 at every 8-bit code), UNORM and
  RGBA32F, empty slots white.
* Normal map: a constant tangent-space normal tilts the shading normal by
  the encoded angle (closed form of the point-light value); base colour,
  roughness, metallic and emission textures scale their parameters
  (closed form with an independent bilinear / sRGB lookup).
* The look-at camera's basis and frustum (ray through the centre hits the
  target, edge rays at fov / 2, right vector level, aspect ratio).
"""
import numpy as np
import pytest

from tests.scene_pack import SceneDesc

PI = np.pi


# ---------------------------------------------------------------- textbook BSDF --
def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def ggx_d(cos_h, a):
    a2 = a * a
    d = cos_h * cos_h * (a2 - 1) + 1
    return a2 / (PI * d * d)


def smith_lambda(cos_t, a):
    c2 = cos_t * cos_t
    t2 = (1 - c2) / c2
    return 0.5 * (-1 + np.sqrt(1 + a * a * t2))


def fresnel_dielectric(cos_i, eta):
    """Unpolarised Fresnel reflectance, eta = n_t / n_i."""
    cos_i = np.clip(cos_i, 0, 1)
    s2 = (1 - cos_i ** 2) / eta ** 2
    out = np.ones_like(np.asarray(cos_i, np.float64))
    ok = s2 < 1
    ct = np.sqrt(1 - s2[ok])
    ci = np.asarray(cos_i)[ok]
    rs = (ci - eta * ct) / (ci + eta * ct)
    rp = (eta * ci - ct) / (eta * ci + ct)
    out[ok] = 0.5 * (rs * rs + rp * rp)
    return out


def reflect_f(base, rough, metal, trans, n, wo, wi):
    """Reflection half of the reference uber-BSDF in world space: Lambert
    (1 - m)(1 - t) base / pi plus GGX D G2 F / (4 cos_o cos_i), F = Schlick
    with F0 = mix(0.04, base, metallic)."""
    base = np.asarray(base, np.float32).astype(np.float64)  # as the float32 material record holds it
    rough, metal = float(np.float32(rough)), float(np.float32(metal))
    a = max(rough * rough, 1e-3)
    co, ci = np.dot(n, wo), np.einsum("...k,k->...", wi, n)
    h = _unit(wo + wi)
    f0 = 0.04 * (1 - metal) + base * metal
    F = f0 + (1 - f0) * (1 - np.einsum("...k,k->...", h, wo))[..., None] ** 5
    G2 = 1 / (1 + smith_lambda(co, a) + smith_lambda(ci, a))
    spec = (ggx_d(np.einsum("...k,k->...", h, n), a) * G2 / (4 * co * ci))[..., None] * F
    return (1 - metal) * (1 - trans) * base / PI + spec


def _r32(v):
    """Values as the float32 light and material records hold them."""
    return np.asarray(v, np.float32).astype(np.float64)


def material(**kw):
    m = dict(base=(0.7, 0.7, 0.7), roughness=0.5, metallic=0.0, transmission=0.0, ior=1.5, entering=True)
    m.update(kw)
    return m


# --------------------------------------------------------------------- transmission --
def walter_btdf(m, wo, wi):
    """Walter et al. 2007 eq. 21 with i = wo in medium eta_i and o = wi in
    medium eta_o (local frame, n = +z), times the reference's tint
    base (1 - metallic) transmission; height-correlated G."""
    eta_i, eta_o = (1.0, m["ior"]) if m["entering"] else (m["ior"], 1.0)
    a = max(m["roughness"] ** 2, 1e-3)
    ht = -(eta_i * wo + eta_o * wi)
    ht = ht / np.linalg.norm(ht, axis=-1, keepdims=True)
    ht = np.where(ht[..., 2:3] < 0, -ht, ht)
    ih, oh = np.einsum("...k,...k->...", wo, ht), np.einsum("...k,...k->...", wi, ht)
    F = fresnel_dielectric(np.abs(ih), eta_o / eta_i)
    G = 1 / (1 + smith_lambda(np.abs(wo[..., 2]), a) + smith_lambda(np.abs(wi[..., 2]), a))
    num = np.abs(ih) * np.abs(oh) * eta_o ** 2 * (1 - F) * G * ggx_d(ht[..., 2], a)
    den = np.abs(wo[..., 2]) * np.abs(wi[..., 2]) * (eta_i * ih + eta_o * oh) ** 2
    tint = np.asarray(m["base"]) * (1 - m["metallic"]) * m["transmission"]
    return (num / den)[..., None] * tint, ih * oh < 0


def _dirs(rng, n, lower):
    v = _unit(rng.normal(size=(n, 3)))
    v[:, 2] = -np.abs(v[:, 2]) if lower else np.abs(v[:, 2])
    return v


@pytest.mark.parametrize("entering", [True, False])
@pytest.mark.parametrize("rough", [0.2, 0.6])
def test_btdf_matches_walter(oracle, entering, rough):
    rng = np.random.default_rng(5)
    m = material(base=(0.9, 0.8, 0.7), roughness=rough, metallic=0.2, transmission=0.8, ior=1.45, entering=entering)
    for _ in range(20):
        wo = _dirs(rng, 1, lower=False)[0]
        wi = _dirs(rng, 400, lower=True)
        f, _ = oracle.bsdf_eval(m, wo, wi)
        ref, valid = walter_btdf(m, wo, wi)
        assert valid.sum() > 0
        np.testing.assert_allclose(f[valid], ref[valid], rtol=1e-9, atol=1e-14)
        assert np.all(f[~valid] == 0)


@pytest.mark.parametrize("rough", [0.15, 0.5])
def test_btdf_generalised_reciprocity(oracle, rough):
    """Walter 2007 eq. 25: f_t(i, o) / eta_o^2 = f_t(o, i) / eta_i^2.  Seen
    from the inside, the frame flips (z -> -z) and the relative index
    inverts."""
    rng = np.random.default_rng(7)
    ior = 1.5
    inside = material(roughness=rough, transmission=1.0, ior=ior, entering=True)
    outside = material(roughness=rough, transmission=1.0, ior=ior, entering=False)
    wo = _dirs(rng, 1500, lower=False)
    wi = _dirs(rng, 1500, lower=True)
    flip = np.array([1, 1, -1.0])
    n_ok = 0
    for a, b in zip(wo, wi):
        f_in, _ = oracle.bsdf_eval(inside, a, b[None])
        f_out, _ = oracle.bsdf_eval(outside, b * flip, (a * flip)[None])
        if f_in[0, 0] > 0:
            n_ok += 1
            assert f_in[0, 0] / ior ** 2 == pytest.approx(f_out[0, 0], rel=1e-9)
        else:
            assert f_out[0, 0] == 0
    assert n_ok > 200


# -------------------------------------------------------------- sampling chi-square --
def true_sampling_pdf(oracle, m, wo, wi):
    """Density of the directions bsdfSample returns: bsdfEval's pdf, plus
    on the reflection side the transmission lobe's Fresnel reflection,
    w_t F(wo.h) D_wo(h) / (4 wo.h) with the visible-normal density
    D_wo(h) = G1(wo) (wo.h) D(h) / cos_o."""
    _, pdf = oracle.bsdf_eval(m, wo, wi)
    pdf = pdf.copy()
    eta = m["ior"] if m["entering"] else 1 / m["ior"]
    # refraction side: bsdfEval tests only woH * wiH < 0 (bsdf.glsl:130-133),
    # so it also gives a pdf to pairs whose half vector faces away from wo,
    # which a visible-normal sample never is
    down = wi[:, 2] < 0
    ht = _unit(wo + wi[down] * eta)
    ht = np.where(ht[:, 2:3] < 0, -ht, ht)
    pdf[np.flatnonzero(down)[ht @ wo < 0]] = 0
    base = np.asarray(m["base"], np.float64)
    lum = lambda c: 0.2126 * c[0] + 0.7152 * c[1] + 0.0722 * c[2]
    f0 = 0.04 * (1 - m["metallic"]) + base * m["metallic"]
    ws = lum(f0 + (1 - f0) * (1 - abs(wo[2])) ** 5)
    wd = (1 - m["metallic"]) * (1 - m["transmission"]) * lum(base)
    wt = (1 - m["metallic"]) * m["transmission"]
    tot = wd + ws + wt
    wt /= tot
    up = wi[:, 2] > 0
    if wt > 0 and up.any():
        a = max(m["roughness"] ** 2, 1e-3)
        h = _unit(wo + wi[up])
        F = fresnel_dielectric(np.abs(h @ wo), eta)
        g1 = 1 / (1 + smith_lambda(wo[2], a))
        pdf[up] += wt * F * g1 * ggx_d(h[:, 2], a) / (4 * wo[2])
    return pdf


CHI2_CASES = [
    ("glass in", material(base=(1, 1, 1), roughness=0.4, transmission=1.0, ior=1.5, entering=True)),
    ("glass out", material(base=(1, 1, 1), roughness=0.4, transmission=1.0, ior=1.5, entering=False)),
    ("mixed", material(base=(0.7, 0.6, 0.5), roughness=0.6, transmission=0.5, ior=1.33)),
    ("metal", material(base=(0.9, 0.8, 0.6), roughness=0.3, metallic=1.0)),
    ("dielectric", material(base=(0.8, 0.8, 0.8), roughness=0.9)),
]


@pytest.mark.parametrize("name,m", CHI2_CASES, ids=[c[0] for c in CHI2_CASES])
def test_bsdf_sampling_chi_square(oracle, name, m):
    from scipy import stats
    rng = np.random.default_rng(11)
    th, ph = np.radians(35.0), 0.7
    wo = np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
    n = 400_000
    wi, _, _, _, valid = oracle.bsdf_sample(m, wo, rng.random((n, 4)))
    wi = wi[valid]
    nc, nphi, sub = 24, 24, 16
    ci = np.clip(wi[:, 2], -1, 1)
    phi = np.mod(np.arctan2(wi[:, 1], wi[:, 0]), 2 * PI)
    cnt, _, _ = np.histogram2d(ci, phi, bins=[nc, nphi], range=[[-1, 1], [0, 2 * PI]])
    # expected counts: midpoint rule on a 16 x 16 sub-grid of each (cos, phi) bin
    cc = -1 + (np.arange(nc * sub) + 0.5) * (2 / (nc * sub))
    pp = (np.arange(nphi * sub) + 0.5) * (2 * PI / (nphi * sub))
    C, P = np.meshgrid(cc, pp, indexing="ij")
    s = np.sqrt(1 - C ** 2)
    dirs = np.stack([s * np.cos(P), s * np.sin(P), C], -1).reshape(-1, 3)
    dens = true_sampling_pdf(oracle, m, wo, dirs).reshape(nc * sub, nphi * sub)
    cell = (2 / (nc * sub)) * (2 * PI / (nphi * sub))
    expect = n * (dens * cell).reshape(nc, sub, nphi, sub).sum(axis=(1, 3))
    assert abs(expect.sum() - valid.sum()) < 5 * np.sqrt(valid.sum()) + 0.003 * n, (expect.sum(), valid.sum())
    keep = expect > 20
    chi2 = ((cnt[keep] - expect[keep]) ** 2 / expect[keep]).sum()
    dof = int(keep.sum()) - 1
    p = stats.chi2.sf(chi2, dof)
    print(f"{name}: chi2 {chi2:.1f} dof {dof} p {p:.3g}; outside the kept bins: "
          f"{cnt[~keep].sum():.0f} sampled, {expect[~keep].sum():.1f} expected")
    assert p > 1e-4
    assert abs(cnt[~keep].sum() - expect[~keep].sum()) < 6 * np.sqrt(expect[~keep].sum() + 25)


# --------------------------------------------------------------------------- scenes --
def _plane(desc, half=50.0, uv_scale=None, mat=0):
    pos = np.array([[-half, 0, -half], [half, 0, -half], [half, 0, half], [-half, 0, half]], np.float32)
    idx = [0, 2, 1, 0, 3, 2]  # counter-clockwise seen from +y: normal +y
    uvs = None
    if uv_scale is not None:
        uvs = (pos[:, [0, 2]] * uv_scale).astype(np.float32)
    mesh = desc.add_mesh(pos, idx, uvs=uvs)
    desc.add_instance(mesh, None, mat)
    return mesh


def _pcg4d(v):
    v = v.astype(np.uint64)
    M = np.uint64(0xFFFFFFFF)
    v = (v * np.uint64(1664525) + np.uint64(1013904223)) & M
    x, y, z, w = v[..., 0], v[..., 1], v[..., 2], v[..., 3]
    x = (x + y * w) & M; y = (y + z * x) & M; z = (z + x * y) & M; w = (w + y * z) & M
    x ^= x >> np.uint64(16); y ^= y >> np.uint64(16); z ^= z >> np.uint64(16); w ^= w >> np.uint64(16)
    x = (x + y * w) & M; y = (y + z * x) & M; z = (z + x * y) & M; w = (w + y * z) & M
    return np.stack([x, y, z, w], -1)


def camera_hits(cam, W, H, frame):
    """World ray (o, d) of every pixel's first sample in `frame` (jitter =
    the path's first rng2, random.glsl:16-31) and its hit on the plane y=0."""
    py, px = np.mgrid[0:H, 0:W]
    st = np.stack([px, py, np.full_like(px, frame), np.ones_like(px)], -1)
    h = _pcg4d(st)
    j = h[..., :2].astype(np.float32).astype(np.float64) * 2.0 ** -32
    ux, uy = (px + j[..., 0]) / W, (py + j[..., 1]) / H
    fr = np.array(cam.frustum, np.float64)
    sx = fr[0] * (1 - ux) + fr[1] * ux
    sy = fr[3] * (1 - uy) + fr[2] * uy
    R, U, Fw = (np.array(getattr(cam, k), np.float64) for k in ("right", "up", "forward"))
    d = _unit(Fw + sx[..., None] * R + sy[..., None] * U)
    o = np.array(cam.pos, np.float64)
    t = -o[1] / d[..., 1]
    return o + t[..., None] * d, d


def _cam(oracle, W, H, pos=(0.4, 1.6, 2.2), target=(0.0, 0.0, 0.0)):
    return oracle.camera_lookat(pos, target, (0, 1, 0), 40.0, W, H)


@pytest.mark.parametrize("rough,metal", [(0.5, 0.0), (0.3, 1.0), (1.0, 0.0)])
def test_point_light_closed_form(oracle, rough, metal):
    W = H = 24
    d = SceneDesc()
    base = (0.6, 0.5, 0.4)
    mat = d.add_material(base_color=base, roughness=rough, metallic=metal)
    _plane(d, mat=mat)
    lp, inten = _r32([0.3, 2.0, -0.2]), _r32([5.0, 4.0, 3.0])
    d.add_light(type=1, position=lp, radiance=inten)
    cam = _cam(oracle, W, H)
    img, _ = d.render_oracle(cam, W, H, spp=1, max_bounces=0)
    p, dirs = camera_hits(cam, W, H, 0)
    n = np.array([0, 1.0, 0])
    v = lp - p
    dist2 = (v * v).sum(-1)
    wi = v / np.sqrt(dist2)[..., None]
    ref = np.zeros((H, W, 3))
    for y in range(H):
        for x in range(W):
            fx = reflect_f(base, rough, metal, 0.0, n, -dirs[y, x], wi[y, x][None])[0]
            ref[y, x] = fx * wi[y, x, 1] * inten / dist2[y, x]
    np.testing.assert_allclose(img, ref, rtol=1e-9, atol=1e-13)


def _rect_quadrature(p, n, wo, corner, e0, e1, L, bsdf, k=48):
    """Integral of f(wo, wi) cos_i L over a one-sided rect light (area form:
    cos_i cos_l / d^2 dA), midpoint rule on k x k cells."""
    s = (np.arange(k) + 0.5) / k
    q = corner + s[:, None, None] * e0 + s[None, :, None] * e1
    v = q.reshape(-1, 3) - p
    d2 = (v * v).sum(-1)
    wi = v / np.sqrt(d2)[:, None]
    ln = _unit(np.cross(e0, e1))
    cos_l = -(wi @ ln)
    cos_i = wi @ n
    area = np.linalg.norm(np.cross(e0, e1))
    w = np.where((cos_l > 0) & (cos_i > 0), cos_i * cos_l / d2, 0.0) * area / (k * k)
    if bsdf is None:
        return w.sum()
    return (bsdf(wo, wi) * w[:, None]).sum(0) * L


def _polygon_form_factor(p, n, verts):
    """Lambert's closed form: (1/2pi) sum_i angle_i (n . unit normal of the
    edge plane), i.e. the projected solid angle / pi."""
    r = _unit(verts - p)
    s = 0.0
    for i in range(len(r)):
        a, b = r[i], r[(i + 1) % len(r)]
        g = np.arccos(np.clip(a @ b, -1, 1))
        c = np.cross(a, b)
        s += g * (c / np.linalg.norm(c)) @ n
    return abs(s) / (2 * PI)


def test_rect_quadrature_matches_polygon_form_factor():
    p, n = np.array([0.1, 0.0, -0.2]), np.array([0, 1.0, 0])
    corner, e0, e1 = np.array([-0.3, 1.5, -0.2]), np.array([0.6, 0, 0]), np.array([0, 0, 0.4])
    q = _rect_quadrature(p, n, None, corner, e0, e1, None, None, k=400)
    verts = np.array([corner, corner + e0, corner + e0 + e1, corner + e1])
    assert q / PI == pytest.approx(_polygon_form_factor(p, n, verts), rel=1e-5)


def _per_sample(desc, cam, W, H, frames, **kw):
    return np.stack([desc.render_oracle(cam, W, H, frame0=f, spp=1, **kw)[0] for f in range(frames)])


def test_rect_light_unbiased_vs_quadrature(oracle):
    W = H = 12
    frames = 40
    d = SceneDesc()
    base, rough = (0.6, 0.6, 0.6), 0.5
    mat = d.add_material(base_color=base, roughness=rough)
    _plane(d, mat=mat)
    corner, e0, e1 = _r32([-0.3, 1.5, -0.2]), _r32([0.6, 0, 0]), _r32([0, 0, 0.4])
    # edge0 x edge1 = (0, -0.24, 0): the light faces down
    L = _r32([3.0, 2.0, 1.0])
    d.add_light(type=2, corner=corner, edge0=e0, edge1=e1, radiance=L)
    cam = _cam(oracle, W, H)
    vals = _per_sample(d, cam, W, H, frames, max_bounces=0)
    n = np.array([0, 1.0, 0])
    exp = np.zeros_like(vals)
    for f in range(frames):
        p, dirs = camera_hits(cam, W, H, f)
        for y in range(H):
            for x in range(W):
                bs = lambda wo, wi: reflect_f(base, rough, 0.0, 0.0, n, wo, wi)
                exp[f, y, x] = _rect_quadrature(p[y, x], n, -dirs[y, x], corner, e0, e1, L, bs)
    r = (vals - exp).reshape(-1, 3)
    z = r.mean(0) / (r.std(0) / np.sqrt(r.shape[0]))
    print("rect: z", z, "mean", vals.mean((0, 1, 2)), "expected", exp.mean((0, 1, 2)))
    assert np.all(np.abs(z) < 4.5)
    np.testing.assert_allclose(vals.mean((0, 1, 2)), exp.mean((0, 1, 2)), rtol=0.02)


def _cone_quadrature(n, wo, axis, cos_max, L, bsdf, k=64):
    t = np.cross(axis, [1.0, 0, 0]) if abs(axis[0]) < 0.9 else np.cross(axis, [0, 1.0, 0])
    t = _unit(t)
    b = np.cross(axis, t)
    tm = np.arccos(cos_max)
    psi = (np.arange(k) + 0.5) / k * tm
    phi = (np.arange(2 * k) + 0.5) / (2 * k) * 2 * PI
    Ps, Ph = np.meshgrid(psi, phi, indexing="ij")
    wi = (np.sin(Ps)[..., None] * (np.cos(Ph)[..., None] * t + np.sin(Ph)[..., None] * b) +
          np.cos(Ps)[..., None] * axis).reshape(-1, 3)
    dw = (np.sin(Ps) * (tm / k) * (2 * PI / (2 * k))).reshape(-1)
    cos_i = np.maximum(wi @ n, 0)
    return (bsdf(wo, wi) * (cos_i * dw)[:, None]).sum(0) * L


def test_sun_cone_unbiased_vs_quadrature(oracle):
    W = H = 12
    frames = 40
    d = SceneDesc()
    base, rough = (0.5, 0.6, 0.7), 0.35
    mat = d.add_material(base_color=base, roughness=rough)
    _plane(d, mat=mat)
    axis = _r32(_unit([0.4, 1.0, 0.25]))
    ang = 0.08
    L = _r32([20.0, 18.0, 15.0])
    d.add_light(type=0, direction=axis, radiance=L, angular_radius=ang)
    cam = _cam(oracle, W, H)
    vals = _per_sample(d, cam, W, H, frames, max_bounces=0)
    cos_max = float(np.cos(np.float32(ang)))
    n = np.array([0, 1.0, 0])
    exp = np.zeros_like(vals)
    for f in range(frames):
        p, dirs = camera_hits(cam, W, H, f)
        for y in range(H):
            for x in range(W):
                bs = lambda wo, wi: reflect_f(base, rough, 0.0, 0.0, n, wo, wi)
                exp[f, y, x] = _cone_quadrature(n, -dirs[y, x], axis, cos_max, L, bs)
    r = (vals - exp).reshape(-1, 3)
    z = r.mean(0) / (r.std(0) / np.sqrt(r.shape[0]))
    print("sun: z", z, "mean", vals.mean((0, 1, 2)), "expected", exp.mean((0, 1, 2)))
    assert np.all(np.abs(z) < 4.5)
    np.testing.assert_allclose(vals.mean((0, 1, 2)), exp.mean((0, 1, 2)), rtol=0.01)


def _quad(desc, z, half, mat):
    pos = np.array([[-half, -half, z], [half, -half, z], [half, half, z], [-half, half, z]], np.float32)
    mesh = desc.add_mesh(pos, [0, 1, 2, 0, 2, 3])  # normal +z, towards a camera at the origin
    desc.add_instance(mesh, None, mat)


@pytest.mark.parametrize("bounces", [0, 1])
def test_opacity_pass_through(oracle, bounces):
    W = H = 16
    frames = 48
    p = 0.3
    E1, E2 = np.array([1.0, 0, 0]), np.array([0, 1.0, 0])
    d = SceneDesc()
    front = d.add_material(base_color=(0, 0, 0), opacity=p, emission=E1)
    back = d.add_material(base_color=(0, 0, 0), emission=E2)
    _quad(d, -2.0, 50.0, front)
    _quad(d, -3.0, 80.0, back)
    cam = oracle.camera_lookat((0, 0, 0), (0, 0, -1), (0, 1, 0), 60.0, W, H)
    vals = _per_sample(d, cam, W, H, frames, max_bounces=bounces).reshape(-1, 3)
    is1 = np.all(vals == E1, axis=1)
    other = E2 if bounces == 1 else np.zeros(3)
    is2 = np.all(vals == other, axis=1)
    assert np.all(is1 | is2)
    frac = is1.mean()
    sd = np.sqrt(p * (1 - p) / vals.shape[0])
    assert abs(frac - p) < 4.5 * sd, (frac, p)


# ------------------------------------------------------------------------- textures --
def srgb_eotf(c):
    """IEC 61966-2-1 decoding."""
    c = np.asarray(c, np.float64)
    return np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)


def bilinear_repeat(img, u, v):
    """Independent REPEAT bilinear lookup of a linear (H, W, 4) image,
    texel centres at ((i + 0.5) / W, (j + 0.5) / H)."""
    h, w = img.shape[:2]
    x, y = u * w - 0.5, v * h - 0.5
    x0, y0 = int(np.floor(x)), int(np.floor(y))
    ax, ay = x - x0, y - y0
    g = lambda i, j: img[j % h, i % w]
    return (g(x0, y0) * (1 - ax) + g(x0 + 1, y0) * ax) * (1 - ay) + (g(x0, y0 + 1) * (1 - ax) + g(x0 + 1, y0 + 1) * ax) * ay


def test_texture_lookup(oracle):
    rng = np.random.default_rng(3)
    px8 = rng.integers(0, 256, size=(5, 7, 4), dtype=np.uint8)
    lin = np.concatenate([srgb_eotf(px8[..., :3] / 255.0), px8[..., 3:] / 255.0], -1)
    for fmt, ref in ((0, lin), (1, px8 / 255.0)):
        t = dict(pixels=px8, format=fmt)
        for j in range(5):
            for i in range(7):  # texel centres are exact
                np.testing.assert_allclose(oracle.texture_sample(t, (i + 0.5) / 7, (j + 0.5) / 5), ref[j, i],
                                           rtol=1e-12, atol=1e-15)
        for u, v in rng.random((50, 2)) * 3 - 1:  # interpolation and REPEAT, any (u, v)
            np.testing.assert_allclose(oracle.texture_sample(t, u, v), bilinear_repeat(ref, u, v), rtol=1e-12,
                                       atol=1e-15)
    # a linear ramp is reproduced between texel centres; the seam blends the
    # last and first columns half and half at u = 0
    ramp = np.zeros((1, 8, 4), np.float32)
    ramp[0, :, 0] = np.arange(8)
    t = dict(pixels=ramp, format=2)
    for u in np.linspace(0.5 / 8, 7.5 / 8, 29):
        assert oracle.texture_sample(t, u, 0.5)[0] == pytest.approx(u * 8 - 0.5, abs=1e-12)
    assert oracle.texture_sample(t, 0.0, 0.5)[0] == pytest.approx(3.5)
    # every 8-bit sRGB code decodes by the IEC formula
    codes = np.zeros((1, 256, 4), np.uint8)
    codes[0, :, 0] = np.arange(256)
    t = dict(pixels=codes, format=0)
    got = np.array([oracle.texture_sample(t, (i + 0.5) / 256, 0.5)[0] for i in range(256)])
    np.testing.assert_allclose(got, srgb_eotf(np.arange(256) / 255.0), rtol=1e-13, atol=1e-16)
    assert got[0] == 0 and got[255] == pytest.approx(1.0)
    assert got[128] == pytest.approx(0.2158605001138992, rel=1e-12)  # ((128/255 + 0.055) / 1.055)^2.4
    assert list(oracle.texture_sample(None, 0.3, 0.3)) == [1, 1, 1, 1]  # empty slot


def test_normal_map_tilts_shading_normal(oracle):
    W = H = 16
    alpha = np.radians(25.0)
    tn = np.array([np.sin(alpha), 0.0, np.cos(alpha)])
    nmap = np.zeros((1, 1, 4), np.float32)
    nmap[0, 0, :3] = (tn + 1) / 2
    d = SceneDesc()
    tex = d.add_texture(nmap, 2)
    base, rough = (0.7, 0.7, 0.7), 0.6
    mat = d.add_material(base_color=base, roughness=rough, normal_tex=tex)
    _plane(d, uv_scale=0.25, mat=mat)  # u along +x, v along +z
    lp, inten = _r32([-0.5, 2.0, 0.4]), _r32([6.0, 6.0, 6.0])
    d.add_light(type=1, position=lp, radiance=inten)
    cam = _cam(oracle, W, H)
    img, _ = d.render_oracle(cam, W, H, spp=1, max_bounces=0)
    sc = d.pack()
    t = sc["tangents"][0, :3].astype(np.float64)
    hand = float(sc["tangents"][0, 3])
    n = np.array([0, 1.0, 0])
    # the tangent of an affine uv map is dP/du, the handedness sign (n x t) . dP/dv
    assert np.allclose(t, [1, 0, 0], atol=1e-6) and hand == np.sign(np.cross(n, t) @ [0, 0, 1.0])
    ns = _unit(t * tn[0] + np.cross(n, t) * hand * tn[1] + n * tn[2])
    assert np.degrees(np.arccos(ns @ n)) == pytest.approx(25.0, abs=1e-4)
    p, dirs = camera_hits(cam, W, H, 0)
    ref = np.zeros((H, W, 3))
    for y in range(H):
        for x in range(W):
            v = lp - p[y, x]
            d2 = v @ v
            wi = v / np.sqrt(d2)
            fx = reflect_f(base, rough, 0.0, 0.0, ns, -dirs[y, x], wi[None])[0]
            ref[y, x] = fx * (wi @ ns) * inten / d2
    np.testing.assert_allclose(img, ref, rtol=1e-6, atol=1e-12)


def test_textured_parameters(oracle):
    W = H = 16
    rng = np.random.default_rng(9)
    base_px = rng.integers(0, 256, size=(4, 4, 4), dtype=np.uint8)
    rough_px = rng.integers(60, 256, size=(4, 4, 4), dtype=np.uint8)
    metal_px = rng.integers(0, 256, size=(4, 4, 4), dtype=np.uint8)
    d = SceneDesc()
    tb, tr, tm = d.add_texture(base_px, 0), d.add_texture(rough_px, 1), d.add_texture(metal_px, 1)
    base, rough, metal = np.array([0.9, 0.8, 0.7]), 0.8, 0.6
    mat = d.add_material(base_color=base, roughness=rough, metallic=metal, base_color_tex=tb, roughness_tex=tr,
                         metallic_tex=tm)
    uvs = 0.37
    _plane(d, uv_scale=uvs, mat=mat)
    lp, inten = _r32([0.2, 1.8, 0.3]), _r32([5.0, 5.0, 5.0])
    d.add_light(type=1, position=lp, radiance=inten)
    cam = _cam(oracle, W, H)
    img, _ = d.render_oracle(cam, W, H, spp=1, max_bounces=0)
    lin_b = np.concatenate([srgb_eotf(base_px[..., :3] / 255.0), base_px[..., 3:] / 255.0], -1)
    p, dirs = camera_hits(cam, W, H, 0)
    n = np.array([0, 1.0, 0])
    ref = np.zeros((H, W, 3))
    for y in range(H):
        for x in range(W):
            u, v = p[y, x, 0] * uvs, p[y, x, 2] * uvs
            b = base * bilinear_repeat(lin_b, u, v)[:3]
            r = rough * bilinear_repeat(rough_px / 255.0, u, v)[0]
            m = metal * bilinear_repeat(metal_px / 255.0, u, v)[0]
            vv = lp - p[y, x]
            d2 = vv @ vv
            wi = vv / np.sqrt(d2)
            ref[y, x] = reflect_f(b, r, m, 0.0, n, -dirs[y, x], wi[None])[0] * wi[1] * inten / d2
    # uv comes from float32 vertex uvs: 1e-6-level differences in the lookup
    np.testing.assert_allclose(img, ref, rtol=2e-5, atol=1e-9)


def test_emission_texture(oracle):
    W = H = 16
    rng = np.random.default_rng(4)
    em_px = rng.integers(0, 256, size=(3, 5, 4), dtype=np.uint8)
    d = SceneDesc()
    te = d.add_texture(em_px, 0)
    E = np.array([2.0, 1.5, 1.0])
    mat = d.add_material(base_color=(0, 0, 0), emission=E, emission_tex=te)
    _plane(d, uv_scale=0.5, mat=mat)
    cam = _cam(oracle, W, H)
    img, _ = d.render_oracle(cam, W, H, spp=1, max_bounces=0)
    lin = np.concatenate([srgb_eotf(em_px[..., :3] / 255.0), em_px[..., 3:] / 255.0], -1)
    p, _ = camera_hits(cam, W, H, 0)
    ref = np.array([[E * bilinear_repeat(lin, p[y, x, 0] * 0.5, p[y, x, 2] * 0.5)[:3] for x in range(W)]
                    for y in range(H)])
    np.testing.assert_allclose(img, ref, rtol=2e-5, atol=1e-9)


def test_camera_lookat_basis(oracle):
    W, H, fov = 64, 36, 50.0
    pos, target = np.array([1.0, 2.0, 3.0]), np.array([-0.5, 0.2, -1.0])
    cam = oracle.camera_lookat(pos, target, (0, 1, 0), fov, W, H)
    R, U, Fw = (np.array(getattr(cam, k), np.float64) for k in ("right", "up", "forward"))
    fr = np.array(cam.frustum, np.float64)
    np.testing.assert_allclose(Fw, _unit(target - pos), atol=1e-6)
    for a, b in ((R, U), (R, Fw), (U, Fw)):
        assert abs(a @ b) < 1e-6
    for a in (R, U, Fw):
        assert np.linalg.norm(a) == pytest.approx(1.0, abs=1e-6)
    assert abs(R[1]) < 1e-6 and U[1] > 0               # right is level, up points up
    assert np.cross(R, U) @ Fw < 0                       # right-handed: right x up = -forward
    assert fr[3] == pytest.approx(np.tan(np.radians(fov / 2)), rel=1e-6) and fr[2] == -fr[3]
    assert fr[1] / fr[3] == pytest.approx(W / H, rel=1e-6) and fr[0] == -fr[1]


# -------------------------------------------------------------------------- resolve --
def test_resolve_closed_forms():
    from oracle import resolve as R
    # Narkowicz ACES fit: 0 -> 0, 1 -> (a + b) / (c + d + e), saturates at 1, increasing
    assert R.aces(0.0) == 0.0
    assert R.aces(1.0) == pytest.approx((2.51 + 0.03) / (2.43 + 0.59 + 0.14), rel=1e-15)
    assert R.aces(1e3) == 1.0
    x = np.linspace(0, 8, 4001)
    assert np.all(np.diff(R.aces(x)) >= 0)
    # sRGB OETF is the inverse of the IEC decoding, continuous at the knee
    c = np.linspace(0, 1, 10001)
    np.testing.assert_allclose(R.linear_to_srgb(srgb_eotf(c)), c, atol=1e-12)
    knee = 0.0031308
    assert R.linear_to_srgb(knee) == pytest.approx(1.055 * knee ** (1 / 2.4) - 0.055, abs=2e-6)
    assert R.linear_to_srgb(0.5) == pytest.approx(1.055 * 0.5 ** (1 / 2.4) - 0.055, rel=1e-15)
    # one EV doubles the linear value; above 1 (after the curve) the store saturates
    v = np.array([[[0.01, 0.1, 0.3]]])
    assert np.array_equal(R.resolve_rgba8(v, 1.0, R.TONEMAP_LINEAR), R.resolve_rgba8(2 * v, 0.0, R.TONEMAP_LINEAR))
    out = R.resolve_rgba8(np.array([[[1.5, 0.0, 2e4]]]), 0.0, R.TONEMAP_LINEAR)
    assert list(out[0, 0]) == [255, 0, 255, 255]
    assert list(R.resolve_rgba8(np.array([[[1.0, 0.5, 0.2]]]), 0.0, R.TONEMAP_LINEAR)[0, 0, :3]) == [
        255, round(255 * (1.055 * 0.5 ** (1 / 2.4) - 0.055)), round(255 * (1.055 * 0.2 ** (1 / 2.4) - 0.055))]
