// Synthetic workload generators (host C++).  They produce the seeded inputs
// the benchmark and the tests feed to the renderer; they are not part of the
// rendering path.
//
// Random numbers: std::mt19937_64 (its output sequence is fixed by the C++
// standard), converted to floats as u = (x >> 40) * 2^-24, which is exact in
// binary32 and lies in [0, 1).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
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

// ---- terrain ----------------------------------------------------------------
// Improved Perlin gradient noise (Perlin 2002) in 2D with a permutation
// shuffled by the seeded generator; fBm of 4 octaves, normalised to [-1, 1].
struct Perlin {
    int perm[512];
    explicit Perlin(uint64_t seed) {
        Rng r(seed);
        int p[256];
        for (int i = 0; i < 256; ++i) p[i] = i;
        for (int i = 255; i > 0; --i) {
            const int j = int(r.g() % uint64_t(i + 1));
            const int t = p[i]; p[i] = p[j]; p[j] = t;
        }
        for (int i = 0; i < 512; ++i) perm[i] = p[i & 255];
    }
    static double fade(double t) { return t * t * t * (t * (t * 6 - 15) + 10); }
    static double grad(int h, double x, double y) {
        switch (h & 7) {
            case 0: return x + y;  case 1: return -x + y; case 2: return x - y;  case 3: return -x - y;
            case 4: return x;      case 5: return -x;     case 6: return y;      default: return -y;
        }
    }
    double noise(double x, double y) const {
        const int xi = int(std::floor(x)) & 255, yi = int(std::floor(y)) & 255;
        const double xf = x - std::floor(x), yf = y - std::floor(y);
        const double u = fade(xf), v = fade(yf);
        const int aa = perm[perm[xi] + yi], ab = perm[perm[xi] + yi + 1];
        const int ba = perm[perm[xi + 1] + yi], bb = perm[perm[xi + 1] + yi + 1];
        const double x1 = grad(aa, xf, yf) + u * (grad(ba, xf - 1, yf) - grad(aa, xf, yf));
        const double x2 = grad(ab, xf, yf - 1) + u * (grad(bb, xf - 1, yf - 1) - grad(ab, xf, yf - 1));
        return x1 + v * (x2 - x1);
    }
    double fbm(double x, double y) const {
        double s = 0, a = 1, f = 1, norm = 0;
        for (int o = 0; o < 4; ++o) {
            s += a * noise(x * f, y * f);
            norm += a;
            a *= 0.5;
            f *= 2.0;
        }
        return s / norm;
    }
};

// Height of the terrain at (x, z): amplitude * fbm(x * 0.35, z * 0.35).
static double terrainHeight(const Perlin& p, double x, double z, double amplitude) {
    return amplitude * p.fbm(x * 0.35 + 17.0, z * 0.35 + 31.0);
}

// Heightfield of nx x nz quads over [-sx/2, sx/2] x [-sz/2, sz/2]
// ((nx+1) x (nz+1) vertices, 2 triangles per quad, normals facing +y).
// positions: (nx+1)(nz+1) x 3; indices: nx*nz*6.
void mrts_terrain(uint32_t nx, uint32_t nz, float sx, float sz, float amplitude, uint64_t seed,
                  float* positions, uint32_t* indices) {
    const Perlin p(seed);
    for (uint32_t j = 0; j <= nz; ++j)
        for (uint32_t i = 0; i <= nx; ++i) {
            const double x = -0.5 * sx + double(sx) * i / nx, z = -0.5 * sz + double(sz) * j / nz;
            float* v = positions + (size_t(j) * (nx + 1) + i) * 3;
            v[0] = float(x);
            v[1] = float(terrainHeight(p, x, z, amplitude));
            v[2] = float(z);
        }
    for (uint32_t j = 0; j < nz; ++j)
        for (uint32_t i = 0; i < nx; ++i) {
            const uint32_t v00 = j * (nx + 1) + i, v10 = v00 + 1, v01 = v00 + (nx + 1), v11 = v01 + 1;
            uint32_t* t = indices + (size_t(j) * nx + i) * 6;
            t[0] = v00; t[1] = v01; t[2] = v11;
            t[3] = v00; t[4] = v11; t[5] = v10;
        }
}

float mrts_terrain_height(float x, float z, float amplitude, uint64_t seed) {
    const Perlin p(seed);
    return float(terrainHeight(p, x, z, amplitude));
}

// ---- icosphere -----------------------------------------------------------------
// Unit icosphere by recursive 4:1 subdivision of an icosahedron.  Level n:
// 10*4^n + 2 vertices, 20*4^n triangles.  Returns the vertex count written.
uint32_t mrts_icosphere(uint32_t level, float* positions, uint32_t* indices) {
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    std::vector<double> v = { -1, t, 0, 1, t, 0, -1, -t, 0, 1, -t, 0, 0, -1, t, 0, 1, t,
                              0, -1, -t, 0, 1, -t, t, 0, -1, t, 0, 1, -t, 0, -1, -t, 0, 1 };
    for (size_t i = 0; i < v.size(); i += 3) {
        const double l = std::sqrt(v[i] * v[i] + v[i + 1] * v[i + 1] + v[i + 2] * v[i + 2]);
        v[i] /= l; v[i + 1] /= l; v[i + 2] /= l;
    }
    std::vector<uint32_t> f = { 0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4,
                                11, 10, 2, 10, 7, 6, 7, 1, 8, 3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8,
                                3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1 };
    for (uint32_t lv = 0; lv < level; ++lv) {
        std::vector<uint32_t> nf;
        std::map<uint64_t, uint32_t> cache;
        auto mid = [&](uint32_t a, uint32_t b) {
            const uint64_t key = (uint64_t(std::min(a, b)) << 32) | std::max(a, b);
            auto it = cache.find(key);
            if (it != cache.end()) return it->second;
            double x = v[a * 3] + v[b * 3], y = v[a * 3 + 1] + v[b * 3 + 1], z = v[a * 3 + 2] + v[b * 3 + 2];
            const double l = std::sqrt(x * x + y * y + z * z);
            v.push_back(x / l); v.push_back(y / l); v.push_back(z / l);
            const uint32_t id = uint32_t(v.size() / 3 - 1);
            cache.emplace(key, id);
            return id;
        };
        for (size_t i = 0; i < f.size(); i += 3) {
            const uint32_t a = f[i], b = f[i + 1], c = f[i + 2];
            const uint32_t ab = mid(a, b), bc = mid(b, c), ca = mid(c, a);
            const uint32_t tri[12] = { a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca };
            nf.insert(nf.end(), tri, tri + 12);
        }
        f.swap(nf);
    }
    for (size_t i = 0; i < v.size(); ++i) positions[i] = float(v[i]);
    std::memcpy(indices, f.data(), f.size() * 4);
    return uint32_t(v.size() / 3);
}

// ---- sky ---------------------------------------------------------------------
// Procedural equirectangular sky (row 0 = zenith, u = atan2(z, x) / 2pi):
// horizon-to-zenith gradient and dark ground after the reference's
// tools/sky_gen.py, plus a sun disc of angular radius sunRadius whose
// luminance is sunScale x the zenith luminance, towards (0.4, 0.6, 0.3).
// rgba: h x w x 4 floats.
void mrts_sky(uint32_t w, uint32_t h, float sunRadius, float sunScale, float* rgba) {
    const double kPi = 3.14159265358979323846;
    const double sl = std::sqrt(0.4 * 0.4 + 0.6 * 0.6 + 0.3 * 0.3);
    const double sx = 0.4 / sl, sy = 0.6 / sl, sz = 0.3 / sl;
    const double zr = 0.22, zg = 0.42, zb = 0.85;
    const double zenithLum = 0.2126 * zr + 0.7152 * zg + 0.0722 * zb;
    const double sunTint[3] = { 1.0, 0.93, 0.83 };
    const double tintLum = 0.2126 * sunTint[0] + 0.7152 * sunTint[1] + 0.0722 * sunTint[2];
    const double sunK = sunScale * zenithLum / tintLum;
    const double cosSun = std::cos(double(sunRadius));
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x) {
            const double theta = (y + 0.5) / h * kPi, phi = (x + 0.5) / w * 2.0 * kPi;
            const double dx = std::sin(theta) * std::cos(phi), dy = std::cos(theta), dz = std::sin(theta) * std::sin(phi);
            double r, g, b;
            if (dy < 0.0) {
                const double f = 1.0 + dy;
                r = 0.28 * f + 0.05; g = 0.26 * f + 0.05; b = 0.24 * f + 0.05;
            } else {
                const double t = std::pow(dy, 0.55);
                r = 0.95 * (1 - t) + zr * t; g = 0.97 * (1 - t) + zg * t; b = 1.00 * (1 - t) + zb * t;
                const double c = dx * sx + dy * sy + dz * sz;
                if (c >= cosSun) {
                    r = sunK * sunTint[0]; g = sunK * sunTint[1]; b = sunK * sunTint[2];
                } else if (c > 0.0) {
                    const double glow = std::exp(-std::acos(std::min(1.0, c)) * 14.0) * 3.0;
                    r += glow; g += glow * 0.9; b += glow * 0.7;
                }
            }
            float* o = rgba + (size_t(y) * w + x) * 4;
            o[0] = float(r); o[1] = float(g); o[2] = float(b); o[3] = 1.0f;
        }
}

// ---- scattered instances --------------------------------------------------------
// count uniformly scaled instances on the terrain: radius U[rmin, rmax],
// x in [-sx/2, sx/2] * 0.9, z likewise, resting on the surface; material slot
// uniform in [0, nmat).  xforms: count x 12 (row-major 3x4); mats: count.
void mrts_scatter_on_terrain(uint32_t count, float sx, float sz, float amplitude, uint64_t terrainSeed,
                             float rmin, float rmax, uint32_t nmat, uint64_t seed, float* xforms,
                             uint32_t* mats) {
    Rng r(seed);
    const Perlin p(terrainSeed);
    for (uint32_t i = 0; i < count; ++i) {
        const double rad = rmin + (rmax - rmin) * r.ud();
        const double x = (r.ud() - 0.5) * sx * 0.9, z = (r.ud() - 0.5) * sz * 0.9;
        const double y = terrainHeight(p, x, z, amplitude) + rad;
        float* m = xforms + size_t(i) * 12;
        for (int k = 0; k < 12; ++k) m[k] = 0.0f;
        m[0] = m[5] = m[10] = float(rad);
        m[3] = float(x); m[7] = float(y); m[11] = float(z);
        mats[i] = uint32_t(r.g() % nmat);
    }
}

// ---- room ------------------------------------------------------------------------
// Closed box room [-hx, hx] x [0, hy] x [-hz, hz], each of the 6 walls an
// n x n grid of quads (2 triangles each), facing inwards.  Vertices strictly
// inside a wall are jittered by U[-jitter, jitter] per axis (boundary
// vertices are shared exactly, so the room stays closed).  A rectangular
// window covering windowFrac of the +z wall is cut out (its quads are
// dropped): the only opening.  Returns the triangle count written;
// positions: 6 (n+1)^2 x 3, indices: up to 6 n^2 x 2 x 3.
uint32_t mrts_room(uint32_t n, float hx, float hy, float hz, float windowFrac, float jitter, uint64_t seed,
                   float* positions, uint32_t* indices, uint32_t* vertexCount) {
    Rng r(seed);
    uint32_t nv = 0, nt = 0;
    // wall frames: origin corner, u axis, v axis (u x v points inwards)
    struct Wall { double o[3], u[3], v[3]; };
    const Wall walls[6] = {
        { { -hx, 0, -hz }, { 2 * hx, 0, 0 }, { 0, 0, 2 * hz } },     // floor (y = 0), normal +y
        { { -hx, hy, -hz }, { 0, 0, 2 * hz }, { 2 * hx, 0, 0 } },    // ceiling, normal -y
        { { -hx, 0, -hz }, { 0, 0, 2 * hz }, { 0, hy, 0 } },         // x = -hx, normal +x
        { { hx, 0, -hz }, { 0, hy, 0 }, { 0, 0, 2 * hz } },          // x = +hx, normal -x
        { { -hx, 0, -hz }, { 0, hy, 0 }, { 2 * hx, 0, 0 } },         // z = -hz, normal +z
        { { -hx, 0, hz }, { 2 * hx, 0, 0 }, { 0, hy, 0 } },          // z = +hz (window wall), normal -z
    };
    // window: centred on the +z wall, aspect 2:1, area windowFrac of the wall
    const double wallArea = 2.0 * hx * hy;
    const double ww = std::sqrt(windowFrac * wallArea * 2.0), wh = ww / 2.0;
    for (int w = 0; w < 6; ++w) {
        const Wall& f = walls[w];
        const uint32_t base = nv;
        // winding so that the face normal (u x v, or v x u) points into the room
        const double cr[3] = { f.u[1] * f.v[2] - f.u[2] * f.v[1], f.u[2] * f.v[0] - f.u[0] * f.v[2],
                               f.u[0] * f.v[1] - f.u[1] * f.v[0] };
        const double toC[3] = { 0.0 - f.o[0], 0.5 * hy - f.o[1], 0.0 - f.o[2] };
        const bool flip = cr[0] * toC[0] + cr[1] * toC[1] + cr[2] * toC[2] < 0.0;
        for (uint32_t j = 0; j <= n; ++j)
            for (uint32_t i = 0; i <= n; ++i) {
                const double a = double(i) / n, b = double(j) / n;
                double p[3];
                for (int k = 0; k < 3; ++k) p[k] = f.o[k] + a * f.u[k] + b * f.v[k];
                const bool interior = i > 0 && i < n && j > 0 && j < n;
                for (int k = 0; k < 3; ++k) {
                    const double jit = (2.0 * r.ud() - 1.0) * jitter;  // drawn for every vertex
                    if (interior) p[k] += jit;
                }
                float* o = positions + size_t(nv) * 3;
                o[0] = float(p[0]); o[1] = float(p[1]); o[2] = float(p[2]);
                ++nv;
            }
        for (uint32_t j = 0; j < n; ++j)
            for (uint32_t i = 0; i < n; ++i) {
                if (w == 5) {
                    // quad centre in wall coordinates: x in [-hx, hx], y in [0, hy]
                    const double cx = -hx + (i + 0.5) / n * 2.0 * hx, cy = (j + 0.5) / n * hy;
                    if (std::fabs(cx) < ww / 2 && std::fabs(cy - hy * 0.55) < wh / 2) continue;
                }
                const uint32_t v00 = base + j * (n + 1) + i, v10 = v00 + 1, v01 = v00 + (n + 1), v11 = v01 + 1;
                uint32_t* t = indices + size_t(nt) * 3;
                if (!flip) {
                    t[0] = v00; t[1] = v10; t[2] = v11;
                    t[3] = v00; t[4] = v11; t[5] = v01;
                } else {
                    t[0] = v00; t[1] = v11; t[2] = v10;
                    t[3] = v00; t[4] = v01; t[5] = v11;
                }
                nt += 2;
            }
    }
    *vertexCount = nv;
    return nt;
}

// GGX material parameters: roughness U[0.05, 0.8], metallic 0 or 1,
// albedo U[0.2, 0.9]^3.  out: count x 5 (r, g, b, roughness, metallic).
void mrts_materials(uint32_t count, uint64_t seed, float* out) {
    Rng r(seed);
    for (uint32_t i = 0; i < count; ++i) {
        float* o = out + size_t(i) * 5;
        o[3] = float(0.05 + 0.75 * r.ud());
        o[4] = r.ud() < 0.5 ? 0.0f : 1.0f;
        for (int c = 0; c < 3; ++c) o[c] = float(0.2 + 0.7 * r.ud());
    }
}

// ---- displaced geodesic sphere ----------------------------------------------
// Each icosahedron face split into n x n triangles (20 n^2 triangles; n = 45
// gives 40,500), vertices shared along face edges, projected to the unit
// sphere and moved radially: r = 1 + amplitude * sum_k w_k cos(f_k (d . a_k) +
// phi_k) / sum_k w_k over 8 seeded waves (a_k uniform directions, f_k in
// [3, 11], phi_k uniform, w_k = 1 / f_k).  Returns the vertex count;
// positions: at most 10 n^2 + 2 vertices x 3, indices: 20 n^2 x 3.
uint32_t mrts_displaced_sphere(uint32_t n, float amplitude, uint64_t seed, float* positions, uint32_t* indices) {
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    const double ico[12][3] = { { -1, t, 0 }, { 1, t, 0 }, { -1, -t, 0 }, { 1, -t, 0 }, { 0, -1, t }, { 0, 1, t },
                                { 0, -1, -t }, { 0, 1, -t }, { t, 0, -1 }, { t, 0, 1 }, { -t, 0, -1 }, { -t, 0, 1 } };
    const uint32_t faces[20][3] = { { 0, 11, 5 }, { 0, 5, 1 }, { 0, 1, 7 }, { 0, 7, 10 }, { 0, 10, 11 },
                                    { 1, 5, 9 }, { 5, 11, 4 }, { 11, 10, 2 }, { 10, 7, 6 }, { 7, 1, 8 },
                                    { 3, 9, 4 }, { 3, 4, 2 }, { 3, 2, 6 }, { 3, 6, 8 }, { 3, 8, 9 },
                                    { 4, 9, 5 }, { 2, 4, 11 }, { 6, 2, 10 }, { 8, 6, 7 }, { 9, 8, 1 } };
    Rng r(seed);
    double wa[8][3], wf[8], wp[8], ww[8], wsum = 0.0;
    for (int k = 0; k < 8; ++k) {
        const double z = 2.0 * r.ud() - 1.0, ph = 2.0 * M_PI * r.ud(), s = std::sqrt(std::max(0.0, 1.0 - z * z));
        wa[k][0] = s * std::cos(ph); wa[k][1] = s * std::sin(ph); wa[k][2] = z;
        wf[k] = 3.0 + 8.0 * r.ud();
        wp[k] = 2.0 * M_PI * r.ud();
        ww[k] = 1.0 / wf[k];
        wsum += ww[k];
    }
    // Vertex key: the grid point's barycentric coordinates (a, b, c; a + b +
    // c = n) on the global icosahedron corners, canonical per shared point.
    std::map<std::array<uint64_t, 3>, uint32_t> ids;
    uint32_t nv = 0, nt = 0;
    auto vert = [&](uint32_t f, uint32_t i, uint32_t j) {
        // point = (i * A + j * B + (n - i - j) * C) / n on face f = (A, B, C)
        const uint32_t w[3] = { i, j, n - i - j };
        std::array<std::pair<uint32_t, uint32_t>, 3> parts;
        for (int c = 0; c < 3; ++c) parts[c] = { faces[f][c], w[c] };
        std::sort(parts.begin(), parts.end());
        std::array<uint64_t, 3> key;
        for (int c = 0; c < 3; ++c) key[c] = (uint64_t(parts[c].second == 0 ? 0xFFFF : parts[c].first) << 32) | parts[c].second;
        std::sort(key.begin(), key.end());
        auto it = ids.find(key);
        if (it != ids.end()) return it->second;
        double p[3] = { 0, 0, 0 };
        for (int c = 0; c < 3; ++c)
            for (int a = 0; a < 3; ++a) p[a] += double(w[c]) * ico[faces[f][c]][a];
        const double l = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
        for (double& x : p) x /= l;
        double h = 0.0;
        for (int k = 0; k < 8; ++k)
            h += ww[k] * std::cos(wf[k] * (p[0] * wa[k][0] + p[1] * wa[k][1] + p[2] * wa[k][2]) + wp[k]);
        const double rad = 1.0 + amplitude * h / wsum;
        for (int a = 0; a < 3; ++a) positions[size_t(nv) * 3 + a] = float(p[a] * rad);
        ids.emplace(key, nv);
        return nv++;
    };
    for (uint32_t f = 0; f < 20; ++f)
        for (uint32_t i = 0; i < n; ++i)
            for (uint32_t j = 0; i + j < n; ++j) {
                const uint32_t a = vert(f, i, j), b = vert(f, i + 1, j), c = vert(f, i, j + 1);
                uint32_t* o = indices + size_t(nt++) * 3;
                o[0] = a; o[1] = b; o[2] = c;
                if (i + j + 1 < n) {
                    const uint32_t d = vert(f, i + 1, j + 1);
                    o = indices + size_t(nt++) * 3;
                    o[0] = b; o[1] = d; o[2] = c;
                }
            }
    return nv;
}

// Instance transforms (3 x 4 row-major) uniformly random in a box of edge
// `box` centred at the origin: uniform rotation (random unit quaternion),
// uniform scale in [smin, smax], uniform translation; material index
// uniform in [0, nmat).
void mrts_random_instances(uint32_t count, float box, float smin, float smax, uint32_t nmat, uint64_t seed,
                           float* xforms, uint32_t* mats) {
    Rng r(seed);
    for (uint32_t i = 0; i < count; ++i) {
        // Shoemake's uniform quaternion from three uniforms
        const double u1 = r.ud(), u2 = r.ud(), u3 = r.ud();
        const double a = std::sqrt(1.0 - u1), b = std::sqrt(u1);
        const double qx = a * std::sin(2 * M_PI * u2), qy = a * std::cos(2 * M_PI * u2);
        const double qz = b * std::sin(2 * M_PI * u3), qw = b * std::cos(2 * M_PI * u3);
        const double s = smin + (smax - smin) * r.ud();
        const double R[3][3] = { { 1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw) },
                                 { 2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw) },
                                 { 2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy) } };
        float* m = xforms + size_t(i) * 12;
        for (int row = 0; row < 3; ++row) {
            for (int c = 0; c < 3; ++c) m[row * 4 + c] = float(R[row][c] * s);
            m[row * 4 + 3] = float((r.ud() - 0.5) * box);
        }
        mats[i] = uint32_t(r.g() % nmat);
    }
}

} // extern "C"
