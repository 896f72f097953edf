// Device pieces shared by the integrators: the surface at a hit, the BSDF of a
// material, area-light point sampling and its pdf, and ray writes. Every
// function keeps the reference's order of operations (the sources are built
// with -ffp-contract=off).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "bg_internal.h"
#include "dev_geom.h"
#include "dev_shade.h"

namespace bg {
namespace dev {

__device__ __forceinline__ uint64_t key_of(uint2 k) { return (uint64_t)k.x | ((uint64_t)k.y << 32); }

__device__ __forceinline__ void write_ray(float4* rays, uint64_t i, V o, float t_max, V d, float time) {
    rays[2 * i] = make_float4(o.x, o.y, o.z, t_max);
    rays[2 * i + 1] = make_float4(d.x, d.y, d.z, time);
}
// a ray no traversal accepts anything on (t_max = -inf)
__device__ __forceinline__ void kill_ray(float4* rays, uint64_t i) {
    rays[2 * i] = make_float4(0.0f, 0.0f, 0.0f, -__builtin_inff());
    rays[2 * i + 1] = make_float4(0.0f, 0.0f, 1.0f, 0.0f);
}

// DiffuseAreaLight::emission (diffuse.rs:163-169)
__device__ __forceinline__ C3 emission(const LightDev& l, V n, V w) {
    return (l.double_sided || dot(n, w) > 0.0f) ? C3{l.L[0], l.L[1], l.L[2]} : mkc(0.0f);
}

__device__ __forceinline__ void tri_verts(const float* __restrict__ v, V* p0, V* p1, V* p2) {
    *p0 = mkv(v[0], v[1], v[2]);
    *p1 = mkv(v[3], v[4], v[5]);
    *p2 = mkv(v[6], v[7], v[8]);
}

// Interaction::spawn_ray (interaction.rs:115-121); t_max 1 - 1e-4 in the reference, +inf corrected.
__device__ __forceinline__ void spawn(const Surf& s, V d, bool ref, V* o, float* t_max) {
    *o = offset_ray_origin(s.p, s.perr, s.n, d);
    *t_max = ref ? 1.0f - 0.0001f : __builtin_inff();
}

// The surface interaction of the closest hit `prim` of ray (r0, r1), as the
// primitive's own intersect computes it (triangle.rs / sphere.rs); false only
// for a primitive the ray does not hit, which a traversal never reports.
__device__ __forceinline__ bool hit_surface(const ShadeSceneDev& sc, float4 r0, float4 r1, uint32_t prim, Surf* s,
                                            float* t) {
    RayState rs;
    rs.ox = r0.x; rs.oy = r0.y; rs.oz = r0.z; rs.t_max = r0.w;
    rs.dx = r1.x; rs.dy = r1.y; rs.dz = r1.z;
    ray_setup(rs);
    bool ok;
    if (prim < sc.n_tris) ok = tri_interaction(rs, sc.tri_orig + 9 * (uint64_t)prim, sc.tri_flip[prim] != 0, s, t);
    else ok = sphere_interaction(rs, sc.spheres[prim - sc.n_tris], s, t);
    s->time = r1.w;
    return ok;
}

// MatteMaterial (matte.rs:20-41, sigma 0), the conductor, PlasticMaterial
// (plastic.rs:24-52); BSDF::new (bsdf.rs:33-52). NL = 1 leaves plastic out.
// Returned by value: filling a caller's BSDF through a reference made the
// compiler keep more values in scratch across the shading kernels' calls.
template <int NL>
__device__ __forceinline__ BsdfN<NL> make_bsdf(const MaterialDev& m, const Surf& s, bool ref) {
    BsdfN<NL> bs;
    bs.ns = s.ns;
    bs.ng = s.n;
    bs.ss = unit(s.dpdu);
    bs.ts = cross(bs.ns, bs.ss);
    bs.n = 0;
    if (m.type == 0 || m.type == 2) {  // Lambertian Kd (clamped at 0), if not black
        const C3 kd = C3{clampf(m.r[0], 0.0f, __builtin_inff()), clampf(m.r[1], 0.0f, __builtin_inff()),
                         clampf(m.r[2], 0.0f, __builtin_inff())};
        if (!black(kd)) {
            Lobe l{};
            l.type = kRefl | kDiffuse;
            l.r = kd;
            l.ref = ref;
            bs.add(l);
        }
    }
    if (NL > 1 && m.type == 2) {  // plastic: Trowbridge-Reitz Ks with FresnelDielectric(1.5, 1.0), if not black
        const C3 ks = C3{clampf(m.eta[0], 0.0f, __builtin_inff()), clampf(m.eta[1], 0.0f, __builtin_inff()),
                         clampf(m.eta[2], 0.0f, __builtin_inff())};
        if (!black(ks)) {
            const float rough = m.k[0] != 0.0f ? roughness_to_alpha(m.alpha) : m.alpha;
            Lobe l{};
            l.type = kRefl | kGlossy;
            l.r = ks;
            l.ax = rmax(rough, 0.001f);
            l.ay = rmax(rough, 0.001f);
            l.ref = ref;
            l.dielectric = true;
            bs.add(l);
        }
    } else if (m.type == 1) {  // conductor
        Lobe l{};
        l.type = kRefl | kGlossy;
        l.r = C3{m.r[0], m.r[1], m.r[2]};
        l.ax = rmax(m.alpha, 0.001f);
        l.ay = rmax(m.alpha, 0.001f);
        l.eta = C3{m.eta[0], m.eta[1], m.eta[2]};
        l.k = C3{m.k[0], m.k[1], m.k[2]};
        l.ref = ref;
        l.dielectric = false;
        bs.add(l);
    }
    return bs;
}

// A point sampled on an area light as seen from a surface point; pdf 0 means
// no sample (the caller skips the light).
struct LightPoint {
    V p, n, err;  // the point, its normal and error bounds (the visibility target)
    V wi;         // unit direction from the surface point
    C3 rad;       // emitted radiance along wi
    float pdf;    // solid-angle pdf (the reference drops the 1/area factor)
};

// DiffuseAreaLight::sample_point (diffuse.rs:134-167) over Shape::sample_point
// (shape.rs:17-34) and Triangle::sample (triangle.rs:522-559).
__device__ __forceinline__ LightPoint sample_light_point(const ShadeSceneDev& sc, const LightDev& light, const Surf& s,
                                                         float ul0, float ul1, bool ref) {
    LightPoint o;
    V p0, p1, p2;
    tri_verts(sc.tri_orig + 9 * (uint64_t)light.prim, &p0, &p1, &p2);
    const float sq = __builtin_sqrtf(ul0);
    const float c0 = 1.0f - sq, c1 = ul1 * sq, c2 = 1.0f - c0 - c1;
    o.p = c0 * p0 + c1 * p1 + c2 * p2;
    V ln = unit(cross(p1 - p0, p2 - p0));
    if (sc.tri_flip[light.prim]) ln = ln * -1.0f;
    o.n = ln;
    const V labs = vabs(c0 * p0) + vabs(c1 * p1) + vabs(c2 * p2);
    o.err = gamma_n(6.0f) * labs;
    const float pdf_area = 1.0f / light.area;
    V wi = o.p - s.p;
    float lpdf;
    if (len2(wi) == 0.0f) {
        lpdf = 0.0f;
    } else {
        wi = unit(wi);
        const float conv = len2(s.p - o.p) / __builtin_fabsf(dot(ln, -wi));
        lpdf = ref ? conv : pdf_area * conv;  // shape.rs:26-27 drops 1/area
        if (__builtin_isinf(lpdf)) lpdf = 0.0f;
    }
    o.rad = mkc(0.0f);
    o.wi = mkv(0.0f, 0.0f, 0.0f);
    if (!(lpdf == 0.0f || len2(o.p - s.p) == 0.0f)) {
        o.wi = unit(o.p - s.p);
        o.rad = emission(light, ln, ref ? o.wi : -o.wi);  // diffuse.rs:67 evaluates toward +wi
    } else {
        lpdf = 0.0f;
    }
    o.pdf = lpdf;
    return o;
}

// VisibilityTester's ray (spawn_ray_to_it, interaction.rs:132-143), t_max 1 - 1e-4.
__device__ __forceinline__ void write_shadow_ray(float4* rays, uint64_t i, const Surf& s, const LightPoint& l) {
    const V o = offset_ray_origin(s.p, s.perr, s.n, l.p - s.p);
    const V tgt = offset_ray_origin(l.p, l.err, l.n, o - l.p);
    write_ray(rays, i, o, 1.0f - 0.0001f, tgt - o, s.time);
}

// Shape::pdf (shape.rs:40-57) on a light's triangle: solid-angle pdf of direction wi.
__device__ __noinline__ float light_point_pdf(const ShadeSceneDev& sc, const LightDev& l, const Surf& s, V wi, bool ref) {
    V o;
    float tm;
    spawn(s, wi, ref, &o, &tm);
    RayState rs;
    rs.ox = o.x; rs.oy = o.y; rs.oz = o.z; rs.t_max = tm;
    rs.dx = wi.x; rs.dy = wi.y; rs.dz = wi.z;
    ray_setup(rs);
    const float* v = sc.tri_orig + 9 * (uint64_t)l.prim;
    float t, b0, b1, b2;
    if (!tri_hit_full(rs, v, &t, &b0, &b1, &b2)) return 0.0f;
    V p0, p1, p2;
    tri_verts(v, &p0, &p1, &p2);
    const V hp = b0 * p0 + b1 * p1 + b2 * p2;
    V hn = unit(cross(p0 - p2, p1 - p2));
    if (sc.tri_flip[l.prim]) hn = -hn;
    float pdf = len2(s.p - hp) / (__builtin_fabsf(dot(hn, -wi)) * l.area);
    if (__builtin_isinf(pdf)) pdf = 0.0f;
    return pdf;
}

// Emission of light li seen along the ray that hit triangle `prim` with direction d
// (the triangle's geometric normal, flipped with its orientation).
__device__ __forceinline__ C3 emission_at(const ShadeSceneDev& sc, int32_t li, uint32_t prim, float4 d) {
    V p0, p1, p2;
    tri_verts(sc.tri_orig + 9 * (uint64_t)prim, &p0, &p1, &p2);
    V n = unit(cross(p0 - p2, p1 - p2));
    if (sc.tri_flip[prim]) n = -n;
    return emission(sc.lights[li], n, mkv(-d.x, -d.y, -d.z));
}

}  // namespace dev
}  // namespace bg
