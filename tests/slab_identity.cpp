// Checks the traversal's slab test (bg::dev::slab_entry, compiled here for the
// host from the library's own dev_geom.h) against the oracle's line-by-line
// restatement of bounds3.rs:246-288 (orc::intersect_range_precomputed) on
// boxes and rays built from special values: +-0, +-inf, denormals, huge
// values, origins on box planes, flat boxes and zero direction components
// (whose plane distances are 0 * inf = NaN or +-inf). For every case the
// product's "entered before t_max" (entry < t_max) must equal the oracle's
// result for several t_max, including the entry distance itself and the next
// float above it. Prints "cases N entered E mismatches M"; exits 1 on a
// mismatch. Built and run by tests/test_slab_identity.py.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>

#include "../bonsai_amd/csrc/dev_geom.h"
#include "../oracle/rt.h"

namespace {

uint64_t g_state = 0x9E3779B97F4A7C15ull;
uint64_t next_u64() {
    g_state ^= g_state << 13;
    g_state ^= g_state >> 7;
    g_state ^= g_state << 17;
    return g_state;
}
float uniform(float a, float b) { return a + (b - a) * (float)((next_u64() >> 40) * (1.0 / 16777216.0)); }

const float kSpecial[] = {0.0f, -0.0f, 1.0f, -1.0f, 0.5f, -0.5f, 2.0f, -2.0f, INFINITY, -INFINITY,
                          1e-38f, -1e-38f, 1e-45f, -1e-45f, 3e38f, -3e38f, 0.25f, 1.5f};
constexpr int kNumSpecial = sizeof(kSpecial) / sizeof(kSpecial[0]);

float pick_special() { return kSpecial[next_u64() % kNumSpecial]; }
float grid_value() { return 0.5f * (float)((int)(next_u64() % 9) - 4); }

}  // namespace

int main(int argc, char** argv) {
    const long n_cases = argc > 1 ? std::atol(argv[1]) : 2000000;
    long entered = 0, mismatches = 0;
    for (long it = 0; it < n_cases; ++it) {
        float o[3], d[3], lo[3], hi[3];
        for (int k = 0; k < 3; ++k) {
            switch (next_u64() % 4) {
                case 0: d[k] = pick_special(); break;
                case 1: d[k] = grid_value(); break;
                default: d[k] = uniform(-1.0f, 1.0f);
            }
            const int mode = (int)(next_u64() % 5);
            o[k] = mode == 0 ? pick_special() : grid_value();
            lo[k] = grid_value();
            hi[k] = lo[k] + 0.5f * (float)(next_u64() % 4);  // lo <= hi, sometimes flat
            if (mode == 1) lo[k] = o[k];                     // origin on the lower plane
            if (mode == 2) {                                 // origin on the upper plane
                hi[k] = o[k];
                if (lo[k] > hi[k]) lo[k] = hi[k];
            }
            if (mode == 3) {
                lo[k] = uniform(-1.0f, 1.0f);
                hi[k] = lo[k] + uniform(0.0f, 1.0f);
                o[k] = uniform(-2.0f, 2.0f);
            }
        }
        // the product's per-ray state, as ray_setup fills the fields the slab test reads
        bg::dev::RayState rs{};
        rs.ox = o[0]; rs.oy = o[1]; rs.oz = o[2];
        rs.dx = d[0]; rs.dy = d[1]; rs.dz = d[2];
        rs.ix = 1.0f / d[0]; rs.iy = 1.0f / d[1]; rs.iz = 1.0f / d[2];
        const float entry = bg::dev::slab_entry(rs, lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]);
        // the oracle: the reference's ray and box, inv_dir and dir_is_neg (bvh.rs:281-290)
        orc::Bounds3 b;
        b.mn = orc::V3(lo[0], lo[1], lo[2]);
        b.mx = orc::V3(hi[0], hi[1], hi[2]);
        orc::Ray ray;
        ray.o = orc::V3(o[0], o[1], o[2]);
        ray.d = orc::V3(d[0], d[1], d[2]);
        ray.time = 0.0f;
        const orc::V3 inv(1.0f / d[0], 1.0f / d[1], 1.0f / d[2]);
        const int neg[3] = {inv.x < 0.0f, inv.y < 0.0f, inv.z < 0.0f};
        const float t_maxes[] = {INFINITY, 1.0f, 0.5f, 0.0f, -1.0f, uniform(0.0f, 4.0f), entry,
                                 std::nextafter(entry, INFINITY)};
        bool any = false;
        for (float tm : t_maxes) {
            if (std::isnan(tm)) continue;
            ray.t_max = tm;
            const bool want = orc::intersect_range_precomputed(b, ray, inv, neg);
            const bool got = entry < tm;
            any |= want;
            if (want != got && ++mismatches <= 10)
                std::printf("MISMATCH o (%g %g %g) d (%g %g %g) lo (%g %g %g) hi (%g %g %g) t_max %g: oracle %d, "
                            "entry %g\n", o[0], o[1], o[2], d[0], d[1], d[2], lo[0], lo[1], lo[2], hi[0], hi[1],
                            hi[2], tm, (int)want, entry);
        }
        entered += any;
    }
    std::printf("cases %ld entered %ld mismatches %ld\n", n_cases, entered, mismatches);
    return mismatches != 0;
}
