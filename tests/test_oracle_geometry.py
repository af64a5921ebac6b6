"""Pins the oracle intersectors against closed forms and an independent
float64 numpy intersector.

Reference semantics: Moeller-Trumbore as tests/bvh_tests.cpp:19-34 (accept
1e-4 < t < tMax; u in [0,1], v >= 0, u+v <= 1); watertight test per Woop,
Benthin & Wald 2013 (tmin < t <= tLimit).
"""
import numpy as np
import pytest

TRI = (np.array([0, 0, 0], np.float32), np.array([1, 0, 0], np.float32), np.array([0, 1, 0], np.float32))


def isect_f64(v0, v1, v2, o, d):
    """Plane intersection + barycentric inside test in float64 (independent)."""
    v0, v1, v2, o, d = (np.asarray(x, np.float64) for x in (v0, v1, v2, o, d))
    n = np.cross(v1 - v0, v2 - v0)
    den = n @ d
    if abs(den) < 1e-300:
        return None
    t = (n @ (v0 - o)) / den
    p = o + t * d
    # barycentrics by areas
    a = np.cross(v1 - p, v2 - p) @ n
    b = np.cross(v2 - p, v0 - p) @ n
    c = np.cross(v0 - p, v1 - p) @ n
    if min(a, b, c) < 0:
        return None
    s = a + b + c
    return t, b / s, c / s


@pytest.mark.parametrize("px,py", [(0.25, 0.25), (0.1, 0.7), (0.6, 0.3), (0.33, 0.33)])
def test_closed_form_hits(oracle, px, py):
    o = np.array([px, py, 5.0], np.float32)
    d = np.array([0, 0, -1], np.float32)
    t_mt = oracle.isect_mt(*TRI, o, d)
    wt = oracle.isect_wt(*TRI, o, d)
    assert t_mt == pytest.approx(5.0, rel=1e-6)
    assert wt is not None
    t, b1, b2 = wt
    assert t == pytest.approx(5.0, rel=1e-6)
    assert b1 == pytest.approx(px, abs=1e-6) and b2 == pytest.approx(py, abs=1e-6)


def test_miss_outside_and_behind(oracle):
    d = np.array([0, 0, -1], np.float32)
    assert oracle.isect_mt(*TRI, np.array([0.8, 0.8, 5], np.float32), d) < 0
    assert oracle.isect_wt(*TRI, np.array([0.8, 0.8, 5], np.float32), d) is None
    # behind the origin
    assert oracle.isect_mt(*TRI, np.array([0.2, 0.2, -5], np.float32), d) < 0
    assert oracle.isect_wt(*TRI, np.array([0.2, 0.2, -5], np.float32), d) is None
    # t below tmin (1e-4) is rejected by both
    o = np.array([0.2, 0.2, 5e-5], np.float32)
    assert oracle.isect_mt(*TRI, o, d) < 0
    assert oracle.isect_wt(*TRI, o, d) is None


def test_tmax_strict_and_limit(oracle):
    o = np.array([0.2, 0.2, 2.0], np.float32)
    d = np.array([0, 0, -1], np.float32)
    assert oracle.isect_mt(*TRI, o, d, tmax=2.0) < 0  # t < tMax strictly
    assert oracle.isect_wt(*TRI, o, d, tlimit=2.0) is not None  # t <= limit (tie)
    assert oracle.isect_wt(*TRI, o, d, tlimit=1.999) is None


def test_parallel_ray_misses(oracle):
    o = np.array([0.2, 0.2, 0.0], np.float32)
    d = np.array([1, 0, 0], np.float32)
    assert oracle.isect_mt(*TRI, o, d) < 0
    assert oracle.isect_wt(*TRI, o, d) is None


def test_watertight_shared_edge(oracle):
    """Rays through the shared diagonal of a quad: the watertight test never
    lets a ray through both triangles (no crack)."""
    a, b, c, d4 = (np.array(p, np.float32) for p in ([0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]))
    rng = np.random.default_rng(5)
    misses = 0
    for s in np.linspace(0.0, 1.0, 257)[1:-1]:
        for _ in range(4):
            # aim (after float32 rounding of the origin) at a point of the
            # diagonal x == y from an oblique direction
            dirv = np.array([rng.uniform(-0.3, 0.3), rng.uniform(-0.3, 0.3), -1.0], np.float32)
            target = np.array([s, s, 0.0], np.float32)
            o = (target - dirv * np.float32(rng.uniform(0.5, 3.0))).astype(np.float32)
            h1 = oracle.isect_wt(a, b, c, o, dirv)
            h2 = oracle.isect_wt(a, c, d4, o, dirv)
            if h1 is None and h2 is None:
                misses += 1
    assert misses == 0


def test_matches_float64_away_from_edges(oracle):
    rng = np.random.default_rng(11)
    checked = 0
    for _ in range(800):
        v = rng.uniform(-1, 1, size=(3, 3)).astype(np.float32)
        o = rng.uniform(-3, 3, size=3).astype(np.float32)
        target = (v[0] * 0.3 + v[1] * 0.3 + v[2] * 0.4).astype(np.float32)
        d = (target - o + rng.normal(scale=0.3, size=3)).astype(np.float32)
        ref = isect_f64(v[0], v[1], v[2], o, d)
        wt = oracle.isect_wt(v[0], v[1], v[2], o, d, tmin=1e-4)
        mt = oracle.isect_mt(v[0], v[1], v[2], o, d)
        if ref is None or ref[0] <= 1e-3:
            continue
        t, b1, b2 = ref
        if min(b1, b2, 1 - b1 - b2) < 1e-4:
            continue  # too close to an edge for a float32 decision to be pinned
        checked += 1
        assert wt is not None
        assert wt[0] == pytest.approx(t, rel=2e-5)
        assert wt[1] == pytest.approx(b1, abs=2e-4) and wt[2] == pytest.approx(b2, abs=2e-4)
        assert mt == pytest.approx(t, rel=2e-5)
    assert checked > 200
