// Seeded workload generators (include/bonsai_scenes.h).
#include <cmath>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "../../include/bonsai_scenes.h"

namespace {

inline uint64_t mix64(uint64_t z) {
    // splitmix64 finaliser (Steele, Lea, Flood 2014)
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline uint64_t hash2(uint64_t seed, uint64_t i) { return mix64(mix64(seed) ^ i); }
inline double u01(uint64_t seed, uint64_t i) { return (double)(hash2(seed, i) >> 11) * (1.0 / 9007199254740992.0); }

struct P3 {
    double x, y, z;
};
inline P3 unit(P3 p) {
    const double l = std::sqrt(p.x * p.x + p.y * p.y + p.z * p.z);
    return P3{p.x / l, p.y / l, p.z / l};
}

// Lattice value in [-1, 1].
inline double lattice(uint64_t seed, int64_t i, int64_t j, int64_t k) {
    uint64_t h = mix64(seed ^ mix64((uint64_t)i * 0x8CB92BA72F3D8DD7ull ^ mix64((uint64_t)j * 0xD6E8FEB86659FD93ull ^
                                                                               (uint64_t)k * 0x9E3779B97F4A7C15ull)));
    return (double)(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}
inline double smooth(double t) { return t * t * (3.0 - 2.0 * t); }

double value_noise(uint64_t seed, double x, double y, double z) {
    const double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    const int64_t ix = (int64_t)fx, iy = (int64_t)fy, iz = (int64_t)fz;
    const double tx = smooth(x - fx), ty = smooth(y - fy), tz = smooth(z - fz);
    double c[2][2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int d = 0; d < 2; ++d) c[a][b][d] = lattice(seed, ix + a, iy + b, iz + d);
    const double x00 = c[0][0][0] + (c[1][0][0] - c[0][0][0]) * tx, x10 = c[0][1][0] + (c[1][1][0] - c[0][1][0]) * tx;
    const double x01 = c[0][0][1] + (c[1][0][1] - c[0][0][1]) * tx, x11 = c[0][1][1] + (c[1][1][1] - c[0][1][1]) * tx;
    const double y0 = x00 + (x10 - x00) * ty, y1 = x01 + (x11 - x01) * ty;
    return y0 + (y1 - y0) * tz;
}

}  // namespace

extern "C" {

uint64_t bgx_icosphere_tri_count(int level) { return level < 0 || level > 12 ? 0 : 20ull << (2 * level); }

int bgx_icosphere(int level, const float center[3], float radius, float amplitude, float frequency, uint64_t seed,
                  float* out) {
    if (level < 0 || level > 12 || !out || !center) return BG_ERR_INVALID;
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    std::vector<P3> v = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                         {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    for (P3& p : v) p = unit(p);
    std::vector<uint32_t> f = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4,
                               11, 10, 2, 10, 7, 6, 7, 1, 8, 3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8,
                               3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
    for (int l = 0; l < level; ++l) {
        std::unordered_map<uint64_t, uint32_t> mid;
        mid.reserve(f.size());
        auto midpoint = [&](uint32_t a, uint32_t b) -> uint32_t {
            const uint64_t key = a < b ? ((uint64_t)a << 32 | b) : ((uint64_t)b << 32 | a);
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            const P3 m = unit(P3{(v[a].x + v[b].x) * 0.5, (v[a].y + v[b].y) * 0.5, (v[a].z + v[b].z) * 0.5});
            v.push_back(m);
            const uint32_t id = (uint32_t)(v.size() - 1);
            mid.emplace(key, id);
            return id;
        };
        std::vector<uint32_t> g;
        g.reserve(f.size() * 4);
        for (size_t i = 0; i < f.size(); i += 3) {
            const uint32_t a = f[i], b = f[i + 1], c = f[i + 2];
            const uint32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
            const uint32_t nf[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
            g.insert(g.end(), nf, nf + 12);
        }
        f.swap(g);
    }
    std::vector<float> pos(3 * v.size());
    for (size_t i = 0; i < v.size(); ++i) {
        const P3& p = v[i];
        const double n = value_noise(seed, p.x * frequency, p.y * frequency, p.z * frequency);
        const double r = (double)radius * (1.0 + (double)amplitude * n);
        pos[3 * i + 0] = (float)((double)center[0] + r * p.x);
        pos[3 * i + 1] = (float)((double)center[1] + r * p.y);
        pos[3 * i + 2] = (float)((double)center[2] + r * p.z);
    }
    for (size_t i = 0; i < f.size(); ++i) std::memcpy(out + 3 * i, &pos[3 * f[i]], 3 * sizeof(float));
    return BG_OK;
}

int bgx_rays_incoherent(uint64_t n, const float bmin[3], const float bmax[3], uint64_t seed, bg_ray* out) {
    if (!out || !bmin || !bmax) return BG_ERR_INVALID;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const uint64_t k = (uint64_t)i * 64;
        bg_ray r;
        for (int c = 0; c < 3; ++c) {
            const double u = u01(seed, k + c);
            r.o[c] = (float)((double)bmin[c] + u * ((double)bmax[c] - (double)bmin[c]));
        }
        double x1 = 0, x2 = 0, s = 2;
        for (uint64_t j = 0; s >= 1.0 || s == 0.0; ++j) {
            x1 = 2.0 * u01(seed, k + 3 + 2 * j) - 1.0;
            x2 = 2.0 * u01(seed, k + 4 + 2 * j) - 1.0;
            s = x1 * x1 + x2 * x2;
            if (j > 28) { x1 = 0.6; x2 = 0.0; s = 0.36; }
        }
        const double q = 2.0 * std::sqrt(1.0 - s);
        r.d[0] = (float)(x1 * q);
        r.d[1] = (float)(x2 * q);
        r.d[2] = (float)(1.0 - 2.0 * s);
        r.t_max = INFINITY;
        r.time = 0.0f;
        out[i] = r;
    }
    return BG_OK;
}

float bgx_uniform(uint64_t seed, uint64_t index) { return (float)(hash2(seed, index) >> 40) * (1.0f / 16777216.0f); }

}  // extern "C"
