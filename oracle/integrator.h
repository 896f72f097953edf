// Oracle (test infrastructure only): the path integrator of bonsai restated on
// the CPU: materials (materials/matte.rs, a conductor assembled from
// bxdfs/microfacet.rs + base/microfacet.rs + base/fresnel.rs), BSDF
// (base/bsdf.rs), diffuse area lights (lights/diffuse.rs, base/light.rs),
// next-event estimation (base/integrator.rs:309-437), PathIntegrator::radiance
// (integrators/path.rs:51-146) and the box-filtered film (base/film.rs).
//
// Two semantics are provided. REF reproduces the reference files as written,
// including the defects catalogued beside each use (Q1-Q7). CORRECTED follows
// the evident pbrt-v3 intent (continuation rays, area-measure light pdf,
// front-face emission, BSDF-sampled MIS branch, G1 = 1/(1+Lambda), slope_x_2).
#pragma once

#include <cstdint>
#include <vector>

#include "bvh.h"
#include "camera.h"

namespace orc {

enum Mode { MODE_REF = 0, MODE_CORRECTED = 1 };

struct RGB {
    Float c[3];
    RGB() : c{0, 0, 0} {}
    explicit RGB(Float v) : c{v, v, v} {}
    RGB(Float a, Float b, Float d) : c{a, b, d} {}
    Float operator[](int i) const { return c[i]; }
};
inline RGB operator+(const RGB& a, const RGB& b) { return RGB(a[0] + b[0], a[1] + b[1], a[2] + b[2]); }
inline RGB operator-(const RGB& a, const RGB& b) { return RGB(a[0] - b[0], a[1] - b[1], a[2] - b[2]); }
inline RGB operator*(const RGB& a, const RGB& b) { return RGB(a[0] * b[0], a[1] * b[1], a[2] * b[2]); }
inline RGB operator*(const RGB& a, Float s) { return RGB(a[0] * s, a[1] * s, a[2] * s); }
inline RGB operator/(const RGB& a, const RGB& b) { return RGB(a[0] / b[0], a[1] / b[1], a[2] / b[2]); }
// rgb.rs:407-417: division by a scalar multiplies by its reciprocal
inline RGB operator/(const RGB& a, Float s) { Float inv = 1.0f / s; return RGB(a[0] * inv, a[1] * inv, a[2] * inv); }
inline bool is_black(const RGB& a) { return a[0] == 0.0f && a[1] == 0.0f && a[2] == 0.0f; }
// rgb.rs:97-104
inline Float max_component_value(const RGB& a) { return fmax_r(fmax_r(a[0], a[1]), a[2]); }
inline Float luminance(const RGB& a) { return 0.212671f * a[0] + 0.715160f * a[1] + 0.072169f * a[2]; }
inline RGB sqrt_rgb(const RGB& a) { return RGB(std::sqrt(a[0]), std::sqrt(a[1]), std::sqrt(a[2])); }

// MAT_PLASTIC (plastic.rs:24-52): r = Kd, eta = Ks, alpha = roughness, k[0] != 0: remap roughness.
enum MatType { MAT_MATTE = 0, MAT_CONDUCTOR = 1, MAT_PLASTIC = 2 };
struct Material {
    uint32_t type = MAT_MATTE;
    RGB r;        // matte Kd / conductor R
    Float alpha = 0.0f;
    RGB eta, k;
};

struct AreaLight {
    uint32_t prim;
    RGB radiance;
    bool double_sided;
};

// Scene with shading data. prim_material/prim_light are indexed by primitive id.
struct ShadeScene {
    Scene* geom = nullptr;
    const RefBVH* bvh = nullptr;
    std::vector<Material> materials;
    std::vector<uint32_t> prim_material;
    std::vector<int32_t> prim_light;
    std::vector<AreaLight> lights;
};

// Counter-based sampler stream. Each draw is the splitmix64 finaliser of
// (stream key + (dim + 1) * golden); the stream key mixes (seed, pixel, sample).
// The reference draws f32 uniforms in [0,1) from a per-tile StdRng
// (samplers/stratified.rs:129-151, base/rng.rs); this keeps the draw order
// and the independent-uniform distribution, with a stream per camera sample.
struct SampleStream {
    uint64_t key;
    uint32_t dim;
};
inline uint64_t splitmix_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline SampleStream make_stream(uint64_t seed, uint64_t pixel, uint64_t sample) {
    uint64_t k = splitmix_mix(seed + 0x9E3779B97F4A7C15ull);
    k = splitmix_mix(k ^ ((pixel << 24) + sample));
    return SampleStream{k, 0};
}
inline Float get_1d(SampleStream& s) {
    uint64_t h = splitmix_mix(s.key + (uint64_t)(s.dim + 1) * 0x9E3779B97F4A7C15ull);
    s.dim++;
    return (Float)(uint32_t)(h >> 40) * (1.0f / 16777216.0f);
}

// The integrator a render runs: PathIntegrator (path.rs), DirectLightingIntegrator
// (direct.rs) with LightStrategy::UniformSampleAll / UniformSampleOne, WhittedIntegrator (whitted.rs).
enum Integrator : uint32_t { INTEGRATOR_PATH = 0, INTEGRATOR_DIRECT = 1, INTEGRATOR_WHITTED = 2 };
enum LightStrategy : uint32_t { SAMPLE_ALL_LIGHTS = 0, SAMPLE_ONE_LIGHT = 1 };

struct RenderParams {
    uint32_t width, height, spp, max_depth;
    Float rr_threshold;
    uint64_t seed;
    Mode mode;
    uint32_t x0, y0, x1, y1;  // pixel window to render (film is still width x height)
    uint32_t integrator = INTEGRATOR_PATH;
    uint32_t strategy = SAMPLE_ALL_LIGHTS;
};

// Radiance of one camera sample (path.rs:51-146 / base/integrator.rs:309-437).
RGB radiance(const ShadeScene& sc, const Ray& camera_ray, SampleStream& st, const RenderParams& p);
// DirectLightingIntegrator::radiance (direct.rs:72-113, integrator.rs:274-328).
RGB radiance_direct(const ShadeScene& sc, const Ray& camera_ray, SampleStream& st, const RenderParams& p);
// WhittedIntegrator::radiance (whitted.rs:41-93).
RGB radiance_whitted(const ShadeScene& sc, const Ray& camera_ray, SampleStream& st, const RenderParams& p);

// SamplerIntegrator::render (integrator.rs:35-133): returns per-pixel double
// RGB sums and filter-weight sums over the window (film.rs:232-284 box filter).
void render(const ShadeScene& sc, const PerspCamera& cam, const RenderParams& p, int threads, double* rgb_sum,
            double* weight_sum);

// The BSDF of material m in the frame n = ns = z, dpdu = x: f and pdf of (wo, wi) for every
// lobe (bsdf.rs f, pdf), and one BSDF::sample with (u0, u1). For tests of the BSDF alone.
struct BsdfProbe {
    RGB f;
    Float pdf;
    V3 swi;
    RGB sf;
    Float spdf;
    uint32_t stype;
    int ncomp;
};
BsdfProbe bsdf_probe(const Material& m, Mode mode, const V3& wo, const V3& wi, Float u0, Float u1);

// Single functions of the restatement, for the tests that pin them.
// fresnel_conductor (bxdf.rs:190-217) with eta_i = 1.
RGB probe_fresnel_conductor(Float cos_theta_i, const RGB& eta, const RGB& k);
// Isotropic Trowbridge-Reitz (microfacet.rs) of alpha: out = D(wh), Lambda(wo), G1(wo), pdf(wo, wh).
void probe_tr(Mode mode, Float alpha, const V3& wo, const V3& wh, Float out[4]);
// TrowbridgeReitzDistribution::sample11 (microfacet.rs:55-97): (slope_x, slope_y).
void probe_sample11(Mode mode, Float cos_theta, Float u1, Float u2, Float out[2]);
// MicrofacetDistribution::sample (microfacet.rs:144-153): the sampled microfacet normal.
V3 probe_tr_sample(Mode mode, Float alpha, const V3& wo, Float u1, Float u2);
// Interaction::spawn_ray (interaction.rs:115-121) from (p, p_error, n, time) in direction d.
Ray probe_spawn_ray(const V3& p, const V3& perr, const V3& n, Float time, const V3& d, Mode mode);
// DiffuseAreaLight::sample_point (diffuse.rs:134-167) of light `light` seen from (p, perr, n):
// out = point[3], normal[3], pdf (solid angle), radiance toward the point[3]; pdf 0: no sample.
void probe_light_sample(const ShadeScene& sc, uint32_t light, Mode mode, const V3& p, const V3& perr, const V3& n,
                        Float u0, Float u1, Float out[10]);
// The light's point_pdf (shape.rs:40-57, sphere.rs:356-373) of direction wi from (p, perr, n).
Float probe_light_pdf(const ShadeScene& sc, uint32_t light, Mode mode, const V3& p, const V3& perr, const V3& n,
                      const V3& wi);
// Interaction::spawn_ray_to_it (interaction.rs:132-143); *target receives the offset target point.
Ray probe_spawn_ray_to(const V3& p0, const V3& e0, const V3& n0, Float time, const V3& p1, const V3& e1,
                       const V3& n1, V3* target);

}  // namespace orc
