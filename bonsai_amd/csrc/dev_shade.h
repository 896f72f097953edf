// Device-side shading for the path integrator: surface interactions,
// Lambertian and Trowbridge-Reitz BxDFs, the BSDF frame, diffuse area light
// sampling and the sample stream. f32 operation order follows the reference
// files cited beside each function so that every step is reproducible;
// transcendental functions (sin, cos, atan2, acos) come from the device math
// library and may differ from the host's by an ulp.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "bg_internal.h"
#include "dev_geom.h"

namespace bg {
namespace dev {

constexpr float kOneMinusEps = 0.99999994f;  // 1 - f32::EPSILON
constexpr float kInvPi = 1.0f / kPi;

// ------------------------------------------------------------ vectors
struct V {
    float x, y, z;
};
__device__ __forceinline__ V mkv(float x, float y, float z) { return V{x, y, z}; }
__device__ __forceinline__ V operator+(V a, V b) { return V{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V operator-(V a, V b) { return V{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V operator-(V a) { return V{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ V operator*(V a, float s) { return V{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V operator*(float s, V a) { return V{s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ float dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float len2(V a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
// vec3.rs:135-137 and 330-342: multiply by the reciprocal length
__device__ __forceinline__ V unit(V a) {
    const float inv = 1.0f / __builtin_sqrtf(len2(a));
    return V{a.x * inv, a.y * inv, a.z * inv};
}
__device__ __forceinline__ V vabs(V a) { return V{__builtin_fabsf(a.x), __builtin_fabsf(a.y), __builtin_fabsf(a.z)}; }
// vec3.rs:108-122
__device__ __forceinline__ V cross(V a, V b) {
    const double x = a.x, y = a.y, z = a.z, vx = b.x, vy = b.y, vz = b.z;
    return V{(float)((y * vz) - (z * vy)), (float)((z * vx) - (x * vz)), (float)((x * vy) - (y * vx))};
}
__device__ __forceinline__ V face_forward(V n, V v) { return dot(n, v) < 0.0f ? -n : n; }
__device__ __forceinline__ float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

// point3.rs:31-49
__device__ __forceinline__ V offset_ray_origin(V p, V perr, V n, V w) {
    V off = dot(vabs(n), perr) * n;
    if (dot(w, n) < 0.0f) off = -off;
    V po = p + off;
    if (off.x > 0.0f) po.x = next_up(po.x); else if (off.x < 0.0f) po.x = next_down(po.x);
    if (off.y > 0.0f) po.y = next_up(po.y); else if (off.y < 0.0f) po.y = next_down(po.y);
    if (off.z > 0.0f) po.z = next_up(po.z); else if (off.z < 0.0f) po.z = next_down(po.z);
    return po;
}

// ------------------------------------------------------------ RGB
struct C3 {
    float r, g, b;
};
__device__ __forceinline__ C3 mkc(float v) { return C3{v, v, v}; }
__device__ __forceinline__ C3 operator+(C3 a, C3 b) { return C3{a.r + b.r, a.g + b.g, a.b + b.b}; }
__device__ __forceinline__ C3 operator-(C3 a, C3 b) { return C3{a.r - b.r, a.g - b.g, a.b - b.b}; }
__device__ __forceinline__ C3 operator*(C3 a, C3 b) { return C3{a.r * b.r, a.g * b.g, a.b * b.b}; }
__device__ __forceinline__ C3 operator*(C3 a, float s) { return C3{a.r * s, a.g * s, a.b * s}; }
__device__ __forceinline__ C3 operator/(C3 a, C3 b) { return C3{a.r / b.r, a.g / b.g, a.b / b.b}; }
// rgb.rs:407-417
__device__ __forceinline__ C3 operator/(C3 a, float s) {
    const float inv = 1.0f / s;
    return C3{a.r * inv, a.g * inv, a.b * inv};
}
__device__ __forceinline__ bool black(C3 a) { return a.r == 0.0f && a.g == 0.0f && a.b == 0.0f; }
__device__ __forceinline__ float cmax(C3 a) { return rmax(rmax(a.r, a.g), a.b); }
__device__ __forceinline__ float luminance(C3 a) { return 0.212671f * a.r + 0.715160f * a.g + 0.072169f * a.b; }
__device__ __forceinline__ C3 csqrt(C3 a) { return C3{__builtin_sqrtf(a.r), __builtin_sqrtf(a.g), __builtin_sqrtf(a.b)}; }

// ------------------------------------------------------------ sample stream
__host__ __device__ __forceinline__ uint64_t splitmix_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
// One independent stream per (seed, pixel, sample); draw k is the splitmix64
// finaliser of key + (k + 1) * golden, as 24 bits in [0, 1).
__host__ __device__ __forceinline__ uint64_t stream_key(uint64_t seed, uint64_t pixel, uint64_t sample) {
    uint64_t k = splitmix_mix(seed + 0x9E3779B97F4A7C15ull);
    return splitmix_mix(k ^ ((pixel << 24) + sample));
}
__device__ __forceinline__ float draw(uint64_t key, uint32_t& dim) {
    const uint64_t h = splitmix_mix(key + (uint64_t)(dim + 1) * 0x9E3779B97F4A7C15ull);
    ++dim;
    return (float)(uint32_t)(h >> 40) * (1.0f / 16777216.0f);
}

// ------------------------------------------------------------ interactions
struct Surf {
    V p, perr, n, ns, dpdu, wo;
    float time;
};

// triangle.rs:135-388 for an accepted hit: barycentrics, point, error, normal.
// Returns false if the triangle no longer reports the hit (never, for a
// primitive the traversal accepted on the same ray).
__device__ __forceinline__ bool tri_interaction(const RayState& r, const float* __restrict__ v, bool flip, Surf* s,
                                                float* t_out) {
    float t, b0, b1, b2;
    if (!tri_hit_full(r, v, &t, &b0, &b1, &b2)) return false;
    const V p0 = mkv(v[0], v[1], v[2]), p1 = mkv(v[3], v[4], v[5]), p2 = mkv(v[6], v[7], v[8]);
    const float xs = __builtin_fabsf(b0 * p0.x) + __builtin_fabsf(b1 * p1.x) + __builtin_fabsf(b2 * p2.x);
    const float ys = __builtin_fabsf(b0 * p0.y) + __builtin_fabsf(b1 * p1.y) + __builtin_fabsf(b2 * p2.y);
    const float zs = __builtin_fabsf(b0 * p0.z) + __builtin_fabsf(b1 * p1.z) + __builtin_fabsf(b2 * p2.z);
    const float g7 = gamma_n(7.0f);
    s->perr = mkv(g7 * xs, g7 * ys, g7 * zs);
    s->p = b0 * p0 + b1 * p1 + b2 * p2;
    // dpdu from the default uvs (triangle.rs:241-268): (-1*dp02 - (-1)*dp12) * 1, with the degenerate
    // fallback to a coordinate frame around the face normal
    const V dp02 = p0 - p2, dp12 = p1 - p2;
    const float m1 = -1.0f, nz = -0.0f;
    V dpdu = mkv((m1 * dp02.x - m1 * dp12.x) * 1.0f, (m1 * dp02.y - m1 * dp12.y) * 1.0f, (m1 * dp02.z - m1 * dp12.z) * 1.0f);
    const V dpdv = mkv((nz * dp02.x + m1 * dp12.x) * 1.0f, (nz * dp02.y + m1 * dp12.y) * 1.0f, (nz * dp02.z + m1 * dp12.z) * 1.0f);
    if (len2(cross(dpdu, dpdv)) == 0.0f) {
        // vec3.rs:25-33 coordinate_system of the normalised face normal
        const V ng = unit(cross(p2 - p0, p1 - p0));
        if (__builtin_fabsf(ng.x) > __builtin_fabsf(ng.y)) dpdu = mkv(-ng.z, 0.0f, ng.x) * (1.0f / __builtin_sqrtf(ng.x * ng.x + ng.z * ng.z));
        else dpdu = mkv(0.0f, ng.z, -ng.y) * (1.0f / __builtin_sqrtf(ng.y * ng.y + ng.z * ng.z));
    }
    V n = unit(cross(dp02, dp12));
    if (flip) n = -n;
    s->n = n;
    s->ns = n;
    s->dpdu = dpdu;
    s->wo = mkv(-r.dx, -r.dy, -r.dz);
    *t_out = t;
    return true;
}

// sphere.rs:80-237: interaction of an accepted sphere hit, in world space.
__device__ __noinline__ bool sphere_interaction(const RayState& r, const SphereShadeDev& sp, Surf* s, float* t_out) {
    SphereHit h;
    if (!sphere_hit(r, sp.w2o, sp.radius, sp.z_min, sp.z_max, sp.phi_max, &h)) return false;
    const float* w = sp.w2o;
    // object-space direction (vec3.rs:35-44)
    const V dobj = mkv(w[0] * r.dx + w[1] * r.dy + w[2] * r.dz, w[4] * r.dx + w[5] * r.dy + w[6] * r.dz,
                       w[8] * r.dx + w[9] * r.dy + w[10] * r.dz);
    const float theta = acosf(clampf(h.pz / sp.radius, -1.0f, 1.0f));
    const float z_radius = __builtin_sqrtf(h.px * h.px + h.py * h.py);
    const float inv_z = 1.0f / z_radius;
    const float cos_phi = h.px * inv_z, sin_phi = h.py * inv_z;
    const V dpdu = mkv(-sp.phi_max * h.py, sp.phi_max * h.px, 0.0f);
    const V dpdv = (sp.theta_max - sp.theta_min) * mkv(h.pz * cos_phi, h.pz * sin_phi, -sp.radius * sinf(theta));
    const float g5 = gamma_n(5.0f);
    const V perr = g5 * vabs(mkv(h.px, h.py, h.pz));
    V n = unit(cross(dpdu, dpdv));
    if (sp.flip) n = n * -1.0f;
    // Interaction::transform (interaction.rs:268-300) with point3.rs:106-157
    const float* m = sp.o2w;
    const float g3 = kGamma3;
    const float x = h.px, y = h.py, z = h.pz;
    const float xp = (m[0] * x + m[1] * y) + (m[2] * z + m[3]);
    const float yp = (m[4] * x + m[5] * y) + (m[6] * z + m[7]);
    const float zp = (m[8] * x + m[9] * y) + (m[10] * z + m[11]);
    float e[3];
    for (int k = 0; k < 3; ++k) {
        const float* q = m + 4 * k;
        e[k] = (g3 + 1.0f) * (__builtin_fabsf(q[0]) * perr.x + __builtin_fabsf(q[1]) * perr.y + __builtin_fabsf(q[2]) * perr.z) +
               g3 * (__builtin_fabsf(q[0] * x) + __builtin_fabsf(q[1] * y) + __builtin_fabsf(q[2] * z) + __builtin_fabsf(q[3]));
    }
    s->p = mkv(xp, yp, zp);  // w = 1 for affine transforms
    s->perr = mkv(e[0], e[1], e[2]);
    const V mwo = -dobj;
    s->wo = mkv(m[0] * mwo.x + m[1] * mwo.y + m[2] * mwo.z, m[4] * mwo.x + m[5] * mwo.y + m[6] * mwo.z,
                m[8] * mwo.x + m[9] * mwo.y + m[10] * mwo.z);
    // normal.rs:21-30: normals transform by the inverse transposed
    const V nt = mkv(w[0] * n.x + w[4] * n.y + w[8] * n.z, w[1] * n.x + w[5] * n.y + w[9] * n.z,
                     w[2] * n.x + w[6] * n.y + w[10] * n.z);
    s->n = unit(nt);
    s->ns = face_forward(unit(nt), s->n);
    s->dpdu = mkv(m[0] * dpdu.x + m[1] * dpdu.y + m[2] * dpdu.z, m[4] * dpdu.x + m[5] * dpdu.y + m[6] * dpdu.z,
                  m[8] * dpdu.x + m[9] * dpdu.y + m[10] * dpdu.z);
    *t_out = h.t;
    return true;
}

// ------------------------------------------------------------ BxDFs
constexpr uint32_t kRefl = 1, kTrans = 2, kDiffuse = 4, kGlossy = 8, kSpecular = 16, kAll = 31;

__device__ __forceinline__ float cos2t(V w) { return w.z * w.z; }
__device__ __forceinline__ float abs_cos(V w) { return __builtin_fabsf(w.z); }
__device__ __forceinline__ float sin2t(V w) { return rmax(0.0f, 1.0f - cos2t(w)); }
__device__ __forceinline__ float sint(V w) { return __builtin_sqrtf(sin2t(w)); }
__device__ __forceinline__ float tant(V w) { return sint(w) / w.z; }
__device__ __forceinline__ float tan2t(V w) { return sin2t(w) / cos2t(w); }
__device__ __forceinline__ float cosp(V w) { const float s = sint(w); return s == 0.0f ? 1.0f : clampf(w.x / s, -1.0f, 1.0f); }
__device__ __forceinline__ float sinp(V w) { const float s = sint(w); return s == 0.0f ? 0.0f : clampf(w.y / s, -1.0f, 1.0f); }
__device__ __forceinline__ float cos2p(V w) { return cosp(w) * cosp(w); }
__device__ __forceinline__ float sin2p(V w) { return sinp(w) * sinp(w); }
__device__ __forceinline__ bool same_hemi(V a, V b) { return a.z * b.z > 0.0f; }

// sampling.rs:181-198
__device__ __forceinline__ void concentric_disk(float u0, float u1, float* dx, float* dy) {
    const float ox = 2.0f * u0 - 1.0f, oy = 2.0f * u1 - 1.0f;
    if (ox == 0.0f && oy == 0.0f) { *dx = 0.0f; *dy = 0.0f; return; }
    float theta, rad;
    if (__builtin_fabsf(ox) > __builtin_fabsf(oy)) { theta = (kPi / 4.0f) * (oy / ox); rad = ox; }
    else { theta = (kPi / 2.0f) - (kPi / 4.0f) * (ox / oy); rad = oy; }
    *dx = rad * cosf(theta);
    *dy = rad * sinf(theta);
}

// bxdf.rs:190-217
__device__ __forceinline__ C3 fresnel_conductor(float cos_i, C3 eta_t, C3 k) {
    const C3 c = mkc(clampf(cos_i, -1.0f, 1.0f));
    const C3 one = mkc(1.0f);
    const C3 eta = eta_t / one, eta_k = k / one;
    const C3 c2 = c * c, s2 = one - c2, eta2 = eta * eta, etak2 = eta_k * eta_k;
    const C3 t0 = eta2 - etak2 - s2;
    const C3 a2b2 = csqrt(t0 * t0 + eta2 * etak2 * 4.0f);
    const C3 t1 = a2b2 + c2;
    const C3 a = csqrt((a2b2 + t0) * 0.5f);
    const C3 t2 = c * a * 2.0f;
    const C3 rs = (t1 - t2) / (t1 + t2);
    const C3 t3 = c2 * a2b2 + s2 * s2;
    const C3 t4 = t2 * s2;
    const C3 rp = rs * (t3 - t4) / (t3 + t4);
    return (rp + rs) * 0.5f;
}

// Trowbridge-Reitz (base/microfacet.rs); `ref` selects the reference's G1 and sample11.
__device__ __forceinline__ float tr_d(float ax, float ay, V wh) {
    const float t2 = tan2t(wh);
    if (__builtin_isinf(t2)) return 0.0f;
    const float cos4 = cos2t(wh) * cos2t(wh);
    const float e = (cos2p(wh) / (ax * ax) + sin2p(wh) / (ay * ay)) * t2;
    return 1.0f / (kPi * ax * ay * cos4 * (1.0f + e) * (1.0f + e));
}
__device__ __forceinline__ float tr_lambda(float ax, float ay, V w) {
    const float att = __builtin_fabsf(tant(w));
    if (__builtin_isinf(att)) return 0.0f;
    const float alpha = __builtin_sqrtf(cos2p(w) * ax * ax + sin2p(w) * ay * ay);
    const float a2t2 = (alpha * att) * (alpha * att);
    return (-1.0f + __builtin_sqrtf(1.0f + a2t2)) / 2.0f;
}
__device__ __forceinline__ float tr_g1(bool ref, float ax, float ay, V w) {
    return ref ? 1.0f / (1.0f / tr_lambda(ax, ay, w)) : 1.0f / (1.0f + tr_lambda(ax, ay, w));
}
__device__ __forceinline__ float tr_g(float ax, float ay, V wo, V wi) {
    return 1.0f / (1.0f + tr_lambda(ax, ay, wo) + tr_lambda(ax, ay, wi));
}
__device__ __forceinline__ float tr_pdf(bool ref, float ax, float ay, V wo, V wh) {
    return tr_d(ax, ay, wh) * tr_g1(ref, ax, ay, wo) * __builtin_fabsf(dot(wo, wh)) / abs_cos(wo);
}
// microfacet.rs:55-97
__device__ __forceinline__ void sample11(bool ref, float cos_theta, float u1, float u2, float* sx, float* sy) {
    if (cos_theta > 0.9999f) {
        const float r = __builtin_sqrtf(u1 / (1.0f - u1));
        const float phi = 6.28318530718f * u2;
        *sx = r * cosf(phi);
        *sy = r * sinf(phi);
        return;
    }
    const float sin_t = __builtin_sqrtf(rmax(1.0f - cos_theta * cos_theta, 0.0f));
    const float tan_t = sin_t / cos_theta;
    float a = 1.0f / tan_t;
    const float g1 = 2.0f / (1.0f + __builtin_sqrtf(1.0f + 1.0f / (a * a)));
    a = 2.0f * u1 / g1 - 1.0f;
    float temp = 1.0f / (a * a - 1.0f);
    if (temp > 1e10f) temp = 1e10f;
    const float b = tan_t;
    const float d = __builtin_sqrtf(rmax(b * b * temp * temp - (a * a - b * b) * temp, 0.0f));
    const float x1 = b * temp - d, x2 = b * temp + d;
    const float slope_x = (a < 0.0f || x2 > 1.0f / tan_t) ? x1 : (ref ? x1 : x2);
    float s;
    if (u2 > 0.5f) { s = 1.0f; u2 = 2.0f * (u2 - 0.5f); }
    else { s = -1.0f; u2 = 2.0f * (0.5f - u2); }
    const float z = (u2 * (u2 * (u2 * 0.27385f - 0.73369f) + 0.46341f)) /
                    (u2 * (u2 * (u2 * 0.093073f + 0.309420f) - 1.000000f) + 0.597999f);
    *sx = slope_x;
    *sy = s * z * __builtin_sqrtf(1.0f + slope_x * slope_x);
}
// microfacet.rs:99-112, 144-153
__device__ __forceinline__ V tr_sample(bool ref, float ax, float ay, V wo, float u0, float u1) {
    const bool flip = wo.z < 0.0f;
    const V w = flip ? -wo : wo;
    const V ws = unit(mkv(ax * w.x, ay * w.y, w.z));
    float sx, sy;
    sample11(ref, ws.z, u0, u1, &sx, &sy);
    const float tmp = cosp(ws) * sx - sinp(ws) * sy;
    sy = sinp(ws) * sx + cosp(ws) * sy;
    sx = tmp;
    sx = ax * sx;
    sy = ay * sy;
    const V wh = unit(mkv(-sx, -sy, 1.0f));
    return flip ? -wh : wh;
}

// bxdf.rs:164-187. The reference keeps the sign of cos_i after swapping the
// indices, so for light arriving from behind the normal the parallel term's
// denominator passes through zero (cos_i = -0.5547 for 1.5 / 1.0); `fold`
// (the corrected mode) takes |cos_i| there, as the Fresnel equations do.
__device__ __forceinline__ float fresnel_dielectric(float cos_i, float eta_i, float eta_t, bool fold) {
    cos_i = clampf(cos_i, -1.0f, 1.0f);
    if (!(cos_i > 0.0f)) {
        const float tmp = eta_i;
        eta_i = eta_t;
        eta_t = tmp;
        if (fold) cos_i = __builtin_fabsf(cos_i);
    }
    const float sin_i = __builtin_sqrtf(rmax(0.0f, 1.0f - cos_i * cos_i));
    const float sin_t = eta_i / eta_t * sin_i;
    if (sin_t >= 1.0f) return 1.0f;
    const float cos_t = __builtin_sqrtf(rmax(0.0f, 1.0f - sin_t * sin_t));
    const float par = ((eta_t * cos_i) - (eta_i * cos_t)) / ((eta_t * cos_i) + (eta_i * cos_t));
    const float perp = ((eta_i * cos_i) - (eta_t * cos_t)) / ((eta_i * cos_i) + (eta_t * cos_t));
    return (par * par + perp * perp) / 2.0f;
}

// microfacet.rs:45-53
__device__ __forceinline__ float roughness_to_alpha(float roughness) {
    roughness = rmax(roughness, 1e-3f);
    const float x = logf(roughness);
    return 1.62142f + 0.819955f * x + 0.1734f * x * x + 0.0171201f * x * x * x + 0.000640711f * x * x * x * x;
}

// One lobe: Lambertian reflection (lambertian.rs) or microfacet reflection
// (bxdfs/microfacet.rs:69-147) with a conductor Fresnel term or, for the
// plastic material, FresnelDielectric(1.5, 1.0) (plastic.rs:40).
struct Lobe {
    uint32_t type;  // 0: none
    C3 r;
    float ax, ay;
    C3 eta, k;
    bool ref;
    bool dielectric;
};

__device__ __forceinline__ C3 lobe_f(const Lobe& l, V wo, V wi) {
    if (l.type & kDiffuse) return l.r * (1.0f / kPi);
    const float co = abs_cos(wo), ci = abs_cos(wi);
    V wh = wi + wo;
    if (ci == 0.0f || co == 0.0f) return mkc(0.0f);
    if (wh.x == 0.0f && wh.y == 0.0f && wh.z == 0.0f) return mkc(0.0f);
    wh = unit(wh);
    const float cwh = dot(wi, face_forward(wh, mkv(0.0f, 0.0f, 1.0f)));
    const C3 F = l.dielectric ? mkc(fresnel_dielectric(cwh, 1.5f, 1.0f, !l.ref)) : fresnel_conductor(__builtin_fabsf(cwh), l.eta, l.k);
    return l.r * tr_d(l.ax, l.ay, wh) * tr_g(l.ax, l.ay, wo, wi) * F / (4.0f * ci * co);
}
__device__ __forceinline__ float lobe_pdf(const Lobe& l, V wo, V wi) {
    if (l.type & kDiffuse) return same_hemi(wo, wi) ? abs_cos(wi) * (1.0f / kPi) : 0.0f;
    if (!same_hemi(wo, wi)) return 0.0f;
    const V wh = unit(wo + wi);
    return tr_pdf(l.ref, l.ax, l.ay, wo, wh) / (4.0f * dot(wo, wh));
}
struct BxS {
    V wi;
    C3 f;
    float pdf;
    uint32_t type;
};
__device__ __forceinline__ BxS lobe_sample(const Lobe& l, V wo, float u0, float u1) {
    BxS s{mkv(0, 0, 0), mkc(0.0f), 0.0f, 0};
    if (l.type & kDiffuse) {
        float dx, dy;
        concentric_disk(u0, u1, &dx, &dy);
        V wi = mkv(dx, dy, __builtin_sqrtf(rmax(0.0f, 1.0f - dx * dx - dy * dy)));
        if (wo.z < 0.0f) wi.z *= -1.0f;
        s.wi = wi;
        s.f = lobe_f(l, wo, wi);
        s.pdf = lobe_pdf(l, wo, wi);
        return s;
    }
    if (wo.z == 0.0f) return s;
    const V wh = tr_sample(l.ref, l.ax, l.ay, wo, u0, u1);
    if (dot(wo, wh) < 0.0f) return s;
    const V wi = -wo + 2.0f * dot(wo, wh) * wh;  // bxdf.rs:144-146
    if (!same_hemi(wo, wi)) return s;
    s.wi = wi;
    s.f = lobe_f(l, wo, wi);
    s.pdf = tr_pdf(l.ref, l.ax, l.ay, wo, wh) / (4.0f * dot(wo, wh));
    return s;
}
__device__ __forceinline__ bool matches(uint32_t type, uint32_t flags) { return (type & flags) == type; }

// bsdf.rs:16-259 with up to NL <= 2 lobes (plastic: Lambertian, then
// microfacet), held as two named members so that every index is known at
// compile time; NL = 1 compiles the single-lobe materials' code alone.
template <int NL>
struct BsdfN {
    V ns, ng, ss, ts;
    Lobe l0, l1;
    int n;
    __device__ const Lobe& lobe(int i) const { return i == 0 ? l0 : l1; }
    __device__ void add(const Lobe& l) {
        if (n == 0) l0 = l;
        else l1 = l;
        ++n;
    }
    __device__ bool any_specular() const {
        return (n > 0 && (l0.type & kSpecular)) || (NL > 1 && n > 1 && (l1.type & kSpecular));
    }
    __device__ V local(V v) const { return mkv(dot(v, ss), dot(v, ts), dot(v, ns)); }
    __device__ V world(V v) const {
        return mkv(ss.x * v.x + ts.x * v.y + ns.x * v.z, ss.y * v.x + ts.y * v.y + ns.y * v.z,
                   ss.z * v.x + ts.z * v.y + ns.z * v.z);
    }
    __device__ int ncomp(uint32_t flags) const {
        int c = 0;
#pragma unroll
        for (int i = 0; i < NL; ++i)
            if (i < n && matches(lobe(i).type, flags)) ++c;
        return c;
    }
    // bsdf.rs:84-104
    __device__ C3 f(V wo_w, V wi_w, uint32_t flags) const {
        const V wo = local(wo_w), wi = local(wi_w);
        if (wo.z == 0.0f) return mkc(0.0f);
        const bool refl = dot(wi_w, ng) * dot(wo_w, ng) > 0.0f;
        C3 r = mkc(0.0f);
#pragma unroll
        for (int i = 0; i < NL; ++i) {
            const Lobe& l = lobe(i);
            if (i < n && matches(l.type, flags) && ((refl && (l.type & kRefl)) || (!refl && (l.type & kTrans))))
                r = r + lobe_f(l, wo, wi);
        }
        return r;
    }
    // bsdf.rs:234-258
    __device__ float pdf(V wo_w, V wi_w, uint32_t flags) const {
        if (n == 0) return 0.0f;
        const V wo = local(wo_w), wi = local(wi_w);
        if (wo.z == 0.0f) return 0.0f;
        float ret = 0.0f, comps = 0.0f;
#pragma unroll
        for (int i = 0; i < NL; ++i)
            if (i < n && matches(lobe(i).type, flags)) {
                comps += 1.0f;
                ret += lobe_pdf(lobe(i), wo, wi);
            }
        return comps > 0.0f ? ret / comps : 0.0f;
    }
    // bsdf.rs:140-232: the component-th matching lobe (u0 remapped) is sampled; for a non-specular
    // pick the pdf averages every matching lobe and f sums them
    __device__ BxS sample(V wo_w, float u0, float u1, uint32_t flags) const {
        BxS out{mkv(0, 0, 0), mkc(0.0f), 0.0f, 0};
        const int m = ncomp(flags);
        if (m == 0) return out;
        uint32_t comp = (uint32_t)floorf(u0 * (float)m);
        if (comp > (uint32_t)m - 1) comp = (uint32_t)m - 1;
        // the comp-th matching lobe
        const bool first_matches = n > 0 && matches(l0.type, flags);
        const int chosen = (NL == 1 || (comp == 0 && first_matches)) ? 0 : 1;
        const Lobe& b = lobe(chosen);
        const float ur = rmin(u0 * (float)m - (float)comp, kOneMinusEps);
        const V wo = local(wo_w);
        if (wo.z == 0.0f) return out;
        BxS s = lobe_sample(b, wo, ur, u1);
        if (s.pdf == 0.0f) return out;
        const V wi_w = world(s.wi);
        if (NL > 1 && !(b.type & kSpecular) && m > 1) {
            const Lobe& o = lobe(1 - chosen);
            if (matches(o.type, flags)) s.pdf += lobe_pdf(o, wo, s.wi);
        }
        if (m > 1) s.pdf /= (float)m;
        C3 f = s.f;
        if (!(b.type & kSpecular)) {
            const bool refl = dot(wi_w, ng) * dot(wo_w, ng) > 0.0f;
            f = mkc(0.0f);
#pragma unroll
            for (int i = 0; i < NL; ++i) {
                const Lobe& l = lobe(i);
                if (i < n && matches(l.type, flags) && ((refl && (l.type & kRefl)) || (!refl && (l.type & kTrans))))
                    f = f + lobe_f(l, wo, s.wi);
            }
        }
        out.wi = wi_w;
        out.f = f;
        out.pdf = s.pdf;
        out.type = b.type;
        return out;
    }
};

__device__ __forceinline__ float power_heuristic(float fp, float gp) {
    const float f = 1.0f * fp, g = 1.0f * gp;
    return (f * f) / (f * f + g * g);
}

}  // namespace dev
}  // namespace bg
