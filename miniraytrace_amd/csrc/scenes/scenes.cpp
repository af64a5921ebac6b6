// Synthetic workload generators (host C++).  They produce the seeded inputs
// the benchmark and the tests feed to the renderer; they are not part of the
// rendering path.
//
// Random numbers: std::mt19937_64 (its output sequence is fixed by the C++
// standard), converted to floats as u = (x >> 40) * 2^-24, which is exact in
// binary32 and lies in [0, 1).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>

namespace {

struct Rng {
    std::mt19937_64 g;
    explicit Rng(uint64_t seed) : g(seed) {}
    float u() { return float(g() >> 40) * (1.0f / 16777216.0f); }
    double ud() { return double(g() >> 40) * (1.0 / 16777216.0); }
};

} // namespace

extern "C" {

// Triangle soup: for each triangle a centroid c ~ U[0,1]^3 (3 draws), then per
// vertex c + (2u - 1) * 0.03 per component (9 draws), all in float.
// verts9: count x 9 floats (v0 xyz, v1 xyz, v2 xyz).
void mrts_soup(uint32_t count, uint64_t seed, float* verts9) {
    Rng r(seed);
    for (uint32_t t = 0; t < count; ++t) {
        const float cx = r.u(), cy = r.u(), cz = r.u();
        float* v = verts9 + size_t(t) * 9;
        for (int k = 0; k < 3; ++k) {
            v[k * 3 + 0] = cx + (2.0f * r.u() - 1.0f) * 0.03f;
            v[k * 3 + 1] = cy + (2.0f * r.u() - 1.0f) * 0.03f;
            v[k * 3 + 2] = cz + (2.0f * r.u() - 1.0f) * 0.03f;
        }
    }
}

// Rays into the unit cube: origin (0.5,0.5,0.5) + 2 s with s uniform on the
// unit sphere (z = 2u - 1, phi = 2 pi u'), direction normalize(p - o) with
// p ~ U[0,1]^3; computed in double, stored as float.  tmin = 1e-4, tmax = +inf.
// rays8: n x 8 floats (ox oy oz tmin dx dy dz tmax).
void mrts_soup_rays(uint64_t n, uint64_t seed, float* rays8) {
    Rng r(seed);
    const double kTwoPi = 6.283185307179586476925;
    for (uint64_t i = 0; i < n; ++i) {
        const double z = 2.0 * r.ud() - 1.0;
        const double phi = kTwoPi * r.ud();
        const double rad = std::sqrt(std::fmax(0.0, 1.0 - z * z));
        const double ox = 0.5 + 2.0 * rad * std::cos(phi);
        const double oy = 0.5 + 2.0 * rad * std::sin(phi);
        const double oz = 0.5 + 2.0 * z;
        const double px = r.ud(), py = r.ud(), pz = r.ud();
        double dx = px - ox, dy = py - oy, dz = pz - oz;
        const double len = std::sqrt(dx * dx + dy * dy + dz * dz);
        dx /= len; dy /= len; dz /= len;
        float* o = rays8 + i * 8;
        o[0] = float(ox); o[1] = float(oy); o[2] = float(oz); o[3] = 1e-4f;
        o[4] = float(dx); o[5] = float(dy); o[6] = float(dz); o[7] = INFINITY;
    }
}

} // extern "C"
