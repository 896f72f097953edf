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
    // (each branch has its own ray state and the sphere branch its own surface record: what the
    // non-inlined sphere call takes by address then lives in scratch only on the sphere path,
    // instead of sending every triangle hit's surface through scratch)
    bool ok;
    if (prim < sc.n_tris) {
        RayState rs;
        rs.ox = r0.x; rs.oy = r0.y; rs.oz = r0.z; rs.t_max = r0.w;
        rs.dx = r1.x; rs.dy = r1.y; rs.dz = r1.z;
        ray_setup(rs);
        ok = tri_interaction(rs, sc.tri_orig + 9 * (uint64_t)prim, (sc.prim_info[prim].y & kPrimFlip) != 0, s, t);
    } else {
        RayState rs;
        rs.ox = r0.x; rs.oy = r0.y; rs.oz = r0.z; rs.t_max = r0.w;
        rs.dx = r1.x; rs.dy = r1.y; rs.dz = r1.z;
        ray_setup(rs);
        Surf h;
        float th;
        ok = sphere_interaction(rs, sc.spheres[prim - sc.n_tris], &h, &th);
        *s = h;
        *t = th;
    }
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

// sampling.rs:211-216
__device__ __forceinline__ V uniform_sample_sphere(float u0, float u1) {
    const float z = 1.0f - 2.0f * u0;
    const float r = __builtin_sqrtf(rmax(1.0f - z * z, 0.0f));
    const float phi = 2.0f * kPi * u1;
    return mkv(r * cosf(phi), r * sinf(phi), z);
}

// The sphere's centre in world space (Point3::transform of the origin, point3.rs:51-69; affine).
__device__ __forceinline__ V sphere_center(const SphereShadeDev& sp) { return mkv(sp.o2w[3], sp.o2w[7], sp.o2w[11]); }

// The centre test of sphere.rs:356-363: the origin of a ray from the surface point toward the
// centre lies inside the sphere.
__device__ __forceinline__ bool sphere_contains_origin(const SphereShadeDev& sp, const Surf& s, V c) {
    const V o = offset_ray_origin(s.p, s.perr, s.n, c - s.p);
    return len2(o - c) <= sp.radius * sp.radius;
}

// Sphere::sample (sphere.rs:328-354): a uniform point on the whole sphere. The reference keeps the
// normal of the unreprojected point unnormalised (length r without scale); corrected normalises it.
__device__ __noinline__ void sphere_sample_area(const SphereShadeDev& sp, float u0, float u1, bool ref, V* p, V* perr,
                                                V* n) {
    const V dir = uniform_sample_sphere(u0, u1);
    V obj = mkv(0.0f + sp.radius * dir.x, 0.0f + sp.radius * dir.y, 0.0f + sp.radius * dir.z);
    const float* w = sp.w2o;
    // normal.rs:21-30
    V nn = mkv(w[0] * obj.x + w[4] * obj.y + w[8] * obj.z, w[1] * obj.x + w[5] * obj.y + w[9] * obj.z,
               w[2] * obj.x + w[6] * obj.y + w[10] * obj.z);
    if (!ref) nn = unit(nn);
    if (sp.reverse) nn = nn * -1.0f;
    obj = obj * (sp.radius / __builtin_sqrtf(len2(obj)));
    const V oerr = gamma_n(5.0f) * vabs(obj);
    // Point3::transform_with_point_error (point3.rs:106-157)
    const float* m = sp.o2w;
    const float g3 = kGamma3;
    const float x = obj.x, y = obj.y, z = obj.z;
    *p = mkv((m[0] * x + m[1] * y) + (m[2] * z + m[3]), (m[4] * x + m[5] * y) + (m[6] * z + m[7]),
             (m[8] * x + m[9] * y) + (m[10] * z + m[11]));
    float e[3];
    for (int k = 0; k < 3; ++k) {
        const float* q = m + 4 * k;
        e[k] = (g3 + 1.0f) * (__builtin_fabsf(q[0]) * oerr.x + __builtin_fabsf(q[1]) * oerr.y + __builtin_fabsf(q[2]) * oerr.z) +
               g3 * (__builtin_fabsf(q[0] * x) + __builtin_fabsf(q[1] * y) + __builtin_fabsf(q[2] * z) + __builtin_fabsf(q[3]));
    }
    *perr = mkv(e[0], e[1], e[2]);
    *n = nn;
}

// sampling.rs:229-231
__device__ __forceinline__ float uniform_cone_pdf(float cos_max) { return 1.0f / (2.0f * kPi * (1.0f - cos_max)); }

// Corrected semantics from outside a sphere (pbrt-v3 Sphere::Sample(ref, u)): a direction uniform
// in the cone the sphere subtends and the point of the sphere it meets first (spheres unscaled).
__device__ __noinline__ void sphere_sample_cone(const SphereShadeDev& sp, V c, V ref_p, float u0, float u1, V* p,
                                                V* perr, V* n, float* pdf) {
    const float dc = __builtin_sqrtf(len2(ref_p - c));
    const float inv_dc = 1.0f / dc;
    const V wc = (c - ref_p) * inv_dc;
    // vec3.rs:25-33
    V wcx;
    if (__builtin_fabsf(wc.x) > __builtin_fabsf(wc.y)) wcx = mkv(-wc.z, 0.0f, wc.x) * (1.0f / __builtin_sqrtf(wc.x * wc.x + wc.z * wc.z));
    else wcx = mkv(0.0f, wc.z, -wc.y) * (1.0f / __builtin_sqrtf(wc.y * wc.y + wc.z * wc.z));
    const V wcy = cross(wc, wcx);
    const float sin_max = sp.radius * inv_dc;
    const float sin_max2 = sin_max * sin_max;
    const float inv_sin_max = 1.0f / sin_max;
    const float cos_max = __builtin_sqrtf(rmax(0.0f, 1.0f - sin_max2));
    float cos_t = (cos_max - 1.0f) * u0 + 1.0f;
    float sin_t2 = 1.0f - cos_t * cos_t;
    if (sin_max2 < 0.00068523f) {  // sin^2(1.5 deg): small cones in the sin^2 parameterisation
        sin_t2 = sin_max2 * u0;
        cos_t = __builtin_sqrtf(1.0f - sin_t2);
    }
    const float cos_a = sin_t2 * inv_sin_max + cos_t * __builtin_sqrtf(rmax(0.0f, 1.0f - sin_t2 * inv_sin_max * inv_sin_max));
    const float sin_a = __builtin_sqrtf(rmax(0.0f, 1.0f - cos_a * cos_a));
    const float phi = u1 * 2.0f * kPi;
    const float sx = sin_a * cosf(phi), sy = sin_a * sinf(phi);
    const V nw = sx * (-wcx) + sy * (-wcy) + cos_a * (-wc);
    *p = c + sp.radius * nw;
    *perr = gamma_n(5.0f) * vabs(*p);
    *n = sp.reverse ? nw * -1.0f : nw;
    *pdf = uniform_cone_pdf(cos_max);
}

// DiffuseAreaLight::sample_point (diffuse.rs:134-167) over Shape::sample_point
// (shape.rs:17-34) and Triangle::sample (triangle.rs:522-559) or Sphere::sample
// (sphere.rs:328-354; corrected semantics sample a sphere seen from outside by its cone).
// SPL: the scene has lights on spheres (compiled out otherwise: that code costs the shading
// kernels registers and call stack even where it never runs).
template <bool SPL>
__device__ __forceinline__ LightPoint sample_light_point(const ShadeSceneDev& sc, const LightDev& light, const Surf& s,
                                                         float ul0, float ul1, bool ref) {
    LightPoint o;
    float lpdf;
    bool solid_angle = false;
    if (SPL && light.prim >= sc.n_tris) {
        const SphereShadeDev& sp = sc.spheres[light.prim - sc.n_tris];
        const V c = sphere_center(sp);
        if (!ref && !sphere_contains_origin(sp, s, c)) {
            sphere_sample_cone(sp, c, s.p, ul0, ul1, &o.p, &o.err, &o.n, &lpdf);
            solid_angle = true;
        } else {
            sphere_sample_area(sp, ul0, ul1, ref, &o.p, &o.err, &o.n);
        }
    } else {
        V p0, p1, p2;
        tri_verts(sc.tri_orig + 9 * (uint64_t)light.prim, &p0, &p1, &p2);
        const float sq = __builtin_sqrtf(ul0);
        const float c0 = 1.0f - sq, c1 = ul1 * sq, c2 = 1.0f - c0 - c1;
        o.p = c0 * p0 + c1 * p1 + c2 * p2;
        V ln = unit(cross(p1 - p0, p2 - p0));
        if (sc.prim_info[light.prim].y & kPrimFlip) ln = ln * -1.0f;
        o.n = ln;
        const V labs = vabs(c0 * p0) + vabs(c1 * p1) + vabs(c2 * p2);
        o.err = gamma_n(6.0f) * labs;
    }
    if (!solid_angle) {
        const float pdf_area = 1.0f / light.area;
        V wi = o.p - s.p;
        if (len2(wi) == 0.0f) {
            lpdf = 0.0f;
        } else {
            wi = unit(wi);
            const float conv = len2(s.p - o.p) / __builtin_fabsf(dot(o.n, -wi));
            lpdf = ref ? conv : pdf_area * conv;  // shape.rs:26-27 drops 1/area
            if (__builtin_isinf(lpdf)) lpdf = 0.0f;
        }
    }
    o.rad = mkc(0.0f);
    o.wi = mkv(0.0f, 0.0f, 0.0f);
    if (!(lpdf == 0.0f || len2(o.p - s.p) == 0.0f)) {
        o.wi = unit(o.p - s.p);
        o.rad = emission(light, o.n, ref ? o.wi : -o.wi);  // diffuse.rs:67 evaluates toward +wi
    } else {
        lpdf = 0.0f;
    }
    o.pdf = lpdf;
    return o;
}

// VisibilityTester's ray (spawn_ray_to_it, interaction.rs:132-143), t_max 1 - 1e-4; returns
// the t_max it wrote.
__device__ __forceinline__ float write_shadow_ray(float4* rays, uint64_t i, const Surf& s, const LightPoint& l) {
    const V o = offset_ray_origin(s.p, s.perr, s.n, l.p - s.p);
    const V tgt = offset_ray_origin(l.p, l.err, l.n, o - l.p);
    const float t_max = 1.0f - 0.0001f;
    write_ray(rays, i, o, t_max, tgt - o, s.time);
    return t_max;
}

// Shape::pdf (shape.rs:40-57) on a light's triangle, or sphere.rs:356-373 on a sphere (the cone
// pdf from outside, whatever wi): solid-angle pdf of direction wi.
template <bool SPL>
__device__ __forceinline__ float light_point_pdf(const ShadeSceneDev& sc, const LightDev& l, const Surf& s, V wi, bool ref) {
    if (SPL && l.prim >= sc.n_tris) {
        const SphereShadeDev& sp = sc.spheres[l.prim - sc.n_tris];
        const V c = sphere_center(sp);
        if (!sphere_contains_origin(sp, s, c)) {
            const float sin2 = sp.radius * sp.radius / len2(s.p - c);
            return uniform_cone_pdf(__builtin_sqrtf(rmax(0.0f, 1.0f - sin2)));
        }
    }
    V o;
    float tm;
    spawn(s, wi, ref, &o, &tm);
    RayState rs;
    rs.ox = o.x; rs.oy = o.y; rs.oz = o.z; rs.t_max = tm;
    rs.dx = wi.x; rs.dy = wi.y; rs.dz = wi.z;
    ray_setup(rs);
    V hp, hn;
    float area;
    if (SPL && l.prim >= sc.n_tris) {
        const SphereShadeDev& sp = sc.spheres[l.prim - sc.n_tris];
        Surf h;
        float t;
        if (!sphere_interaction(rs, sp, &h, &t)) return 0.0f;
        hp = h.p;
        hn = h.n;
        area = l.area;
    } else {
        const float* v = sc.tri_orig + 9 * (uint64_t)l.prim;
        float t, b0, b1, b2;
        if (!tri_hit_full(rs, v, &t, &b0, &b1, &b2)) return 0.0f;
        V p0, p1, p2;
        tri_verts(v, &p0, &p1, &p2);
        hp = b0 * p0 + b1 * p1 + b2 * p2;
        hn = unit(cross(p0 - p2, p1 - p2));
        if (sc.prim_info[l.prim].y & kPrimFlip) hn = -hn;
        area = l.area;
    }
    float pdf = len2(s.p - hp) / (__builtin_fabsf(dot(hn, -wi)) * area);
    if (__builtin_isinf(pdf)) pdf = 0.0f;
    return pdf;
}

// Emission of light li seen along the ray (r0, r1) that hit primitive `prim` (the hit's
// geometric normal: the triangle's, flipped with its orientation, or the sphere's at the hit).
template <bool SPL>
__device__ __forceinline__ C3 emission_at(const ShadeSceneDev& sc, int32_t li, uint32_t prim, float4 r0, float4 r1) {
    V n;
    if (SPL && prim >= sc.n_tris) {
        RayState rs;
        rs.ox = r0.x; rs.oy = r0.y; rs.oz = r0.z; rs.t_max = r0.w;
        rs.dx = r1.x; rs.dy = r1.y; rs.dz = r1.z;
        ray_setup(rs);
        Surf h;
        float t;
        if (!sphere_interaction(rs, sc.spheres[prim - sc.n_tris], &h, &t)) return mkc(0.0f);
        n = h.n;
    } else {
        V p0, p1, p2;
        tri_verts(sc.tri_orig + 9 * (uint64_t)prim, &p0, &p1, &p2);
        n = unit(cross(p0 - p2, p1 - p2));
        if (sc.prim_info[prim].y & kPrimFlip) n = -n;
    }
    return emission(sc.lights[li], n, mkv(-r1.x, -r1.y, -r1.z));
}

// Solid-angle density with which light sampling from the ray origin (a path vertex, offset) picks
// the point the ray hit at distance t, |cos| cosl there (the guided estimator's MIS weight of
// emission found by its continuation): area sampling converted by t^2 / (cosl area), or the cone
// density of a sphere seen from outside.
template <bool SPL>
__device__ __forceinline__ float light_pdf_of_hit(const ShadeSceneDev& sc, const LightDev& l, float4 r0, float t,
                                                  float cosl) {
    if (SPL && l.prim >= sc.n_tris) {
        const SphereShadeDev& sp = sc.spheres[l.prim - sc.n_tris];
        const V c = sphere_center(sp);
        const float d2 = len2(mkv(r0.x, r0.y, r0.z) - c);
        if (d2 > sp.radius * sp.radius) return uniform_cone_pdf(__builtin_sqrtf(rmax(0.0f, 1.0f - sp.radius * sp.radius / d2)));
    }
    return t * t / (cosl * l.area);
}

}  // namespace dev
}  // namespace bg
