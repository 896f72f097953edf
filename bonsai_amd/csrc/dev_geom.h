// Device-side intersection arithmetic for gfx950.
//
// Every expression keeps the reference's f32 operation order (the library is
// compiled with -ffp-contract=off, so no multiply-add is fused and each
// operation rounds once, as in the Rust code) so that hit decisions, t and
// barycentrics match the CPU reference bit for bit. Division and sqrt are the
// correctly rounded IEEE operations. Per-ray quantities the reference
// recomputes for every primitive (the permutation and shear of
// triangle.rs:148-165) are computed once per ray; they are pure functions of
// the ray, so the values are identical.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "tri_degenerate.h"

namespace bg {
namespace dev {

constexpr float kMachineEps = 5.9604644775390625e-08f;  // f32::EPSILON / 2
__host__ __device__ constexpr float gamma_n(float n) { return (n * kMachineEps) / (1.0f - n * kMachineEps); }
constexpr float kGamma2 = gamma_n(2.0f);
constexpr float kGamma3 = gamma_n(3.0f);
constexpr float kGamma5 = gamma_n(5.0f);
constexpr float kSlabPad = 1.0f + 2.0f * gamma_n(3.0f);  // bounds3.rs:259
constexpr float kPi = 3.14159265358979323846f;

__device__ __forceinline__ float next_up(float v) {
    if (__builtin_isinf(v) && v > 0.0f) return v;
    if (v == -0.0f) v = 0.0f;
    uint32_t u = __float_as_uint(v);
    u = (v >= 0.0f) ? u + 1u : u - 1u;
    return __uint_as_float(u);
}
__device__ __forceinline__ float next_down(float v) {
    if (__builtin_isinf(v) && v < 0.0f) return v;
    if (v == 0.0f) v = -0.0f;
    uint32_t u = __float_as_uint(v);
    u = (v > 0.0f) ? u - 1u : u + 1u;
    return __uint_as_float(u);
}
// Rust f32::min / max semantics (NaN yields the other operand).
__device__ __forceinline__ float rmin(float a, float b) { return a != a ? b : (b != b ? a : (a < b ? a : b)); }
__device__ __forceinline__ float rmax(float a, float b) { return a != a ? b : (b != b ? a : (a > b ? a : b)); }
// Vec3::max_component of non-negative values (vec3.rs:147-149)
__device__ __forceinline__ float max3(float x, float y, float z) { return rmax(x, rmax(y, z)); }

__device__ __forceinline__ float sel3(float x, float y, float z, int k) { return k == 0 ? x : (k == 1 ? y : z); }

// Per-ray state of a query. The traversal loops hold one per lane for the whole
// query, so it carries no value another field already determines: the
// direction signs of bvh.rs:281-290 are the signs of ix, iy, iz, and the
// permutation (kx, ky) of triangle.rs:148-156 follows from kz.
struct RayState {
    float ox, oy, oz;       // origin
    float dx, dy, dz;       // direction
    float t_max;
    float ix, iy, iz;       // 1 / d
    uint32_t negbits;       // (ix < 0) | (iy < 0) << 1 | (iz < 0) << 2 (bit 3, a record's "no axis", is 0)
    int kz;                 // triangle.rs:148-156: kz = max_dimension(|d|), kx = kz + 1, ky = kx + 1 (mod 3)
    float okx, oky, okz;    // origin components in permuted order
    float sx, sy, sz;       // triangle.rs:163-165 shear
    __host__ __device__ bool neg0() const { return ix < 0.0f; }
    __host__ __device__ bool neg1() const { return iy < 0.0f; }
    __host__ __device__ bool neg2() const { return iz < 0.0f; }
};

// Components of (x, y, z) in the ray's permuted order, the x, y and z of triangle.rs:157-160.
__device__ __forceinline__ float sel_kx(const RayState& r, float x, float y, float z) {
    return r.kz == 2 ? x : (r.kz == 0 ? y : z);
}
__device__ __forceinline__ float sel_ky(const RayState& r, float x, float y, float z) {
    return r.kz == 1 ? x : (r.kz == 2 ? y : z);
}
__device__ __forceinline__ float sel_kz(const RayState& r, float x, float y, float z) {
    return r.kz == 0 ? x : (r.kz == 1 ? y : z);
}

__device__ __forceinline__ void ray_setup(RayState& r) {
    r.ix = 1.0f / r.dx;
    r.iy = 1.0f / r.dy;
    r.iz = 1.0f / r.dz;
    r.negbits = (uint32_t)r.neg0() | (uint32_t)r.neg1() << 1 | (uint32_t)r.neg2() << 2;
    const float ax = __builtin_fabsf(r.dx), ay = __builtin_fabsf(r.dy), az = __builtin_fabsf(r.dz);
    // vec3.rs:159-171 max_dimension
    r.kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    const float dpx = sel_kx(r, r.dx, r.dy, r.dz);
    const float dpy = sel_ky(r, r.dx, r.dy, r.dz);
    const float dpz = sel_kz(r, r.dx, r.dy, r.dz);
    r.okx = sel_kx(r, r.ox, r.oy, r.oz);
    r.oky = sel_ky(r, r.ox, r.oy, r.oz);
    r.okz = sel_kz(r, r.ox, r.oy, r.oz);
    r.sx = -dpx / dpz;
    r.sy = -dpy / dpz;
    r.sz = 1.0f / dpz;
}

// bounds3.rs:246-288 with node bounds (lo, hi).
__device__ __forceinline__ bool slab(const RayState& r, float lx, float ly, float lz, float hx, float hy, float hz) {
    float t_min = ((r.neg0() ? hx : lx) - r.ox) * r.ix;
    float t_max = ((r.neg0() ? lx : hx) - r.ox) * r.ix;
    const float ty_min = ((r.neg1() ? hy : ly) - r.oy) * r.iy;
    float ty_max = ((r.neg1() ? ly : hy) - r.oy) * r.iy;
    t_max *= kSlabPad;
    ty_max *= kSlabPad;
    if (t_min > ty_max || ty_min > t_max) return false;
    if (ty_min > t_min) t_min = ty_min;
    if (ty_max < t_max) t_max = ty_max;
    const float tz_min = ((r.neg2() ? hz : lz) - r.oz) * r.iz;
    float tz_max = ((r.neg2() ? lz : hz) - r.oz) * r.iz;
    tz_max *= kSlabPad;
    if (t_min > tz_max || tz_min > t_max) return false;
    if (tz_min > t_min) t_min = tz_min;
    if (tz_max < t_max) t_max = tz_max;
    return (t_min < r.t_max) && (t_max > 0.0f);
}

// The same test split at its last step: the entry distance t_min of the box
// when the box is entered at all (every step but the t_max comparison), NaN
// otherwise; slab() is then (t_entry < ray t_max), NaN comparing false. Entry
// distances stored on the stack give, at pop time and for any smaller t_max,
// exactly the result a fresh slab() would give.
//
// The six plane distances are the reference's (same selection by direction
// sign, same operations, same padding); their combination is an identity of
// the reference's sequence, in three max/min instructions and three compares:
// - a NaN x distance always ends in "not entered" there (a NaN t_min or t_max
//   survives every update and fails the last comparison);
// - with x ordered, every NaN y or z distance takes part in no comparison and
//   no update there (each test involving it is false), as IEEE maxNum/minNum
//   ignore a NaN operand;
// - the early exits there compare each entry with each other axis's exit;
//   entry <= exit also holds on one axis (the sign selection orders the two
//   planes and the padding only raises a non-negative exit), except where that
//   exit is negative, and then the final exit > 0 fails in both forms.
// So entered <=> max(entries) <= min(exits) && min(exits) > 0, with the entry
// distance max(entries) (equal to the reference's up to the sign of a zero,
// which no comparison sees).
//
// The three exits are padded once, after the minimum: x -> round(pad * x) is
// non-decreasing for the positive pad, so it commutes with the minimum, and a
// NaN exit stays NaN, which minNum ignores in both orders; the minimum can
// differ only in the sign of a zero, which neither comparison sees.
//
// slab_enter returns "entered" and stores the entry distance in *t_in (also
// for a box not entered, where the caller must not use it).
__host__ __device__ __forceinline__ bool slab_enter(const RayState& r, float lx, float ly, float lz, float hx,
                                                    float hy, float hz, float* t_in) {
    const float tx_min = ((r.neg0() ? hx : lx) - r.ox) * r.ix;
    const float tx_max = ((r.neg0() ? lx : hx) - r.ox) * r.ix;
    const float ty_min = ((r.neg1() ? hy : ly) - r.oy) * r.iy;
    const float ty_max = ((r.neg1() ? ly : hy) - r.oy) * r.iy;
    const float tz_min = ((r.neg2() ? hz : lz) - r.oz) * r.iz;
    const float tz_max = ((r.neg2() ? lz : hz) - r.oz) * r.iz;
    *t_in = __builtin_fmaxf(__builtin_fmaxf(tx_min, ty_min), tz_min);
    const float t_out = kSlabPad * __builtin_fminf(__builtin_fminf(tx_max, ty_max), tz_max);
    return !__builtin_isunordered(tx_min, tx_max) && *t_in <= t_out && t_out > 0.0f;
}

// The entry distance of a box the ray enters, NaN for one it does not.
__host__ __device__ __forceinline__ float slab_entry(const RayState& r, float lx, float ly, float lz, float hx,
                                                     float hy, float hz) {
    float t_in;
    return slab_enter(r, lx, ly, lz, hx, hy, hz, &t_in) ? t_in : __builtin_nanf("");
}

using bg::cross64_len2;
using bg::tri_degenerate;

// triangle.rs:179-190 (operand order as written there).
__device__ __forceinline__ void edges_f64(float p0x, float p0y, float p1x, float p1y, float p2x, float p2y, float* e0,
                                          float* e1, float* e2) {
    const double p2txp1ty = (double)p2x * (double)p1y;
    const double p2typ1tx = (double)p2y * (double)p1x;
    *e0 = (float)(p2typ1tx - p2txp1ty);
    const double p0txp2ty = (double)p0x * (double)p2y;
    const double p0typ2tx = (double)p0y * (double)p2x;
    *e1 = (float)(p0typ2tx - p0txp2ty);
    const double p1txp0ty = (double)p1x * (double)p0y;
    const double p1typ0tx = (double)p1y * (double)p0x;
    *e2 = (float)(p1typ0tx - p1txp0ty);
}

// triangle.rs:135-240 (+ the degenerate exit): on success returns t and the
// barycentrics b0, b1, b2. DEGEN = false leaves out the degenerate exit, for
// vertex buffers in which every degenerate triangle has been collapsed to one
// point (its edge functions are then all zero and det == 0 rejects it first).
template <bool DEGEN = true>
__device__ __forceinline__ bool tri_hit_full(const RayState& r, const float* __restrict__ v, float* t_out,
                                             float* b0_out, float* b1_out, float* b2_out) {
    const float p0x = v[0], p0y = v[1], p0z = v[2];
    const float p1x = v[3], p1y = v[4], p1z = v[5];
    const float p2x = v[6], p2y = v[7], p2z = v[8];
    // translate, permute (triangle.rs:142-160)
    float a_x = sel_kx(r, p0x, p0y, p0z) - r.okx, a_y = sel_ky(r, p0x, p0y, p0z) - r.oky,
          a_z = sel_kz(r, p0x, p0y, p0z) - r.okz;
    float b_x = sel_kx(r, p1x, p1y, p1z) - r.okx, b_y = sel_ky(r, p1x, p1y, p1z) - r.oky,
          b_z = sel_kz(r, p1x, p1y, p1z) - r.okz;
    float c_x = sel_kx(r, p2x, p2y, p2z) - r.okx, c_y = sel_ky(r, p2x, p2y, p2z) - r.oky,
          c_z = sel_kz(r, p2x, p2y, p2z) - r.okz;
    // shear (triangle.rs:166-171)
    a_x += r.sx * a_z; a_y += r.sy * a_z;
    b_x += r.sx * b_z; b_y += r.sy * b_z;
    c_x += r.sx * c_z; c_y += r.sy * c_z;
    // edge functions (triangle.rs:174-176)
    float e0 = b_x * c_y - b_y * c_x;
    float e1 = c_x * a_y - c_y * a_x;
    float e2 = a_x * b_y - a_y * b_x;
    if (e0 == 0.0f || e1 == 0.0f || e2 == 0.0f) edges_f64(a_x, a_y, b_x, b_y, c_x, c_y, &e0, &e1, &e2);
    if ((e0 < 0.0f || e1 < 0.0f || e2 < 0.0f) && (e0 > 0.0f || e1 > 0.0f || e2 > 0.0f)) return false;
    const float det = e0 + e1 + e2;
    if (det == 0.0f) return false;
    a_z *= r.sz; b_z *= r.sz; c_z *= r.sz;
    const float t_scaled = e0 * a_z + e1 * b_z + e2 * c_z;
    if (det < 0.0f && (t_scaled >= 0.0f || t_scaled < r.t_max * det)) return false;
    else if (det > 0.0f && (t_scaled <= 0.0f || t_scaled > r.t_max * det)) return false;
    const float inv_det = 1.0f / det;
    const float t = t_scaled * inv_det;
    // conservative t bound (triangle.rs:219-239)
    const float max_zt = max3(__builtin_fabsf(a_z), __builtin_fabsf(b_z), __builtin_fabsf(c_z));
    const float delta_z = kGamma3 * max_zt;
    const float max_xt = max3(__builtin_fabsf(a_x), __builtin_fabsf(b_x), __builtin_fabsf(c_x));
    const float max_yt = max3(__builtin_fabsf(a_y), __builtin_fabsf(b_y), __builtin_fabsf(c_y));
    const float delta_x = kGamma5 * (max_xt + max_zt);
    const float delta_y = kGamma5 * (max_yt + max_zt);
    const float delta_e = 2.0f * (kGamma2 * max_xt * max_yt + delta_y * max_xt + delta_x * max_yt);
    const float max_e = max3(__builtin_fabsf(e0), __builtin_fabsf(e1), __builtin_fabsf(e2));
    const float delta_t =
        3.0f * (kGamma3 * max_e * max_zt + delta_e * max_zt + delta_z * max_e) * __builtin_fabsf(inv_det);
    if (t <= delta_t) return false;
    if (DEGEN && tri_degenerate(p0x, p0y, p0z, p1x, p1y, p1z, p2x, p2y, p2z)) return false;
    *t_out = t;
    *b0_out = e0 * inv_det;
    *b1_out = e1 * inv_det;
    *b2_out = e2 * inv_det;
    return true;
}

// The traversal's test, on the slot-ordered vertices (degenerate triangles collapsed).
__device__ __forceinline__ bool tri_hit(const RayState& r, const float* __restrict__ v, float* t_out, float* b1_out,
                                        float* b2_out) {
    float b0;
    return tri_hit_full<false>(r, v, t_out, &b0, b1_out, b2_out);
}

// ---------------------------------------------------------------- spheres
// EFloat (efloat.rs) as (v, lo, hi).
struct EF {
    float v, lo, hi;
};
__device__ __forceinline__ EF ef(float v, float err) {
    if (err == 0.0f) return EF{v, v, v};
    return EF{v, next_down(v - err), next_up(v + err)};
}
__device__ __forceinline__ EF ef_add(EF a, EF b) { return EF{a.v + b.v, next_down(a.lo + b.lo), next_up(a.hi + b.hi)}; }
__device__ __forceinline__ EF ef_sub(EF a, EF b) { return EF{a.v - b.v, next_down(a.lo - b.hi), next_up(a.hi - b.lo)}; }
__device__ __forceinline__ EF ef_mul(EF a, EF b) {
    const float p0 = a.lo * b.lo, p1 = a.hi * b.lo, p2 = a.lo * b.hi, p3 = a.hi * b.hi;
    return EF{a.v * b.v, next_down(rmin(rmin(rmin(p0, p1), p2), p3)), next_up(rmax(rmax(rmax(p0, p1), p2), p3))};
}
__device__ __forceinline__ EF ef_div(EF a, EF b) {
    const float q0 = a.lo / b.lo, q1 = a.hi / b.lo, q2 = a.lo / b.hi, q3 = a.hi / b.hi;
    const bool straddle = b.lo < 0.0f && b.hi > 0.0f;
    const float lo = straddle ? -__builtin_inff() : next_down(rmin(rmin(rmin(q0, q1), q2), q3));
    const float hi = straddle ? __builtin_inff() : next_up(rmax(rmax(rmax(q0, q1), q2), q3));
    return EF{a.v / b.v, lo, hi};
}
__device__ __forceinline__ bool ef_eq(EF a, EF b) { return a.v == b.v && a.lo == b.lo && a.hi == b.hi; }

struct SphereHit {
    float t;
    float px, py, pz;  // refined object-space point
    float phi;
};

// Refined point and phi for parameter t (sphere.rs:119-130).
__device__ __forceinline__ void sphere_point(float ox, float oy, float oz, float dx, float dy, float dz, float t,
                                             float radius, SphereHit* h) {
    float x = ox + dx * t, y = oy + dy * t, z = oz + dz * t;
    const float s = radius / __builtin_sqrtf(x * x + y * y + z * z);
    x = x * s; y = y * s; z = z * s;
    if (x == 0.0f && y == 0.0f) x = 1e-5f * radius;
    float phi = atan2f(y, x);
    if (phi < 0.0f) phi += 2.0f * kPi;
    h->px = x; h->py = y; h->pz = z; h->phi = phi;
}

// sphere.rs:80-163 / 239-326: the hit decision for one sphere. `w` is the
// world_to_object affine rows; prm = (radius, z_min, z_max, phi_max).
__device__ __noinline__ bool sphere_hit(const RayState& r, const float* __restrict__ w, float radius, float z_min,
                                        float z_max, float phi_max, SphereHit* out) {
    // Point3::transform_with_error (point3.rs:71-104); row 3 is (0,0,0,1).
    const float x = r.ox, y = r.oy, z = r.oz;
    const float xp = (w[0] * x + w[1] * y) + (w[2] * z + w[3]);
    const float yp = (w[4] * x + w[5] * y) + (w[6] * z + w[7]);
    const float zp = (w[8] * x + w[9] * y) + (w[10] * z + w[11]);
    const float wp = (0.0f * x + 0.0f * y) + (0.0f * z + 1.0f);
    const float g3 = kGamma3;
    const float ex = g3 * (__builtin_fabsf(w[0] * x) + __builtin_fabsf(w[1] * y) + __builtin_fabsf(w[2] * z) + __builtin_fabsf(w[3]));
    const float ey = g3 * (__builtin_fabsf(w[4] * x) + __builtin_fabsf(w[5] * y) + __builtin_fabsf(w[6] * z) + __builtin_fabsf(w[7]));
    const float ez = g3 * (__builtin_fabsf(w[8] * x) + __builtin_fabsf(w[9] * y) + __builtin_fabsf(w[10] * z) + __builtin_fabsf(w[11]));
    float ox = xp, oy = yp, oz = zp;
    if (wp != 1.0f) {
        const float iw = 1.0f / wp;
        ox = xp * iw; oy = yp * iw; oz = zp * iw;
    }
    // Vec3::transform_with_error (vec3.rs:46-64)
    const float dx = w[0] * r.dx + w[1] * r.dy + w[2] * r.dz;
    const float dy = w[4] * r.dx + w[5] * r.dy + w[6] * r.dz;
    const float dz = w[8] * r.dx + w[9] * r.dy + w[10] * r.dz;
    const float fx = g3 * (__builtin_fabsf(w[0] * r.dx) + __builtin_fabsf(w[1] * r.dy) + __builtin_fabsf(w[2] * r.dz));
    const float fy = g3 * (__builtin_fabsf(w[4] * r.dx) + __builtin_fabsf(w[5] * r.dy) + __builtin_fabsf(w[6] * r.dz));
    const float fz = g3 * (__builtin_fabsf(w[8] * r.dx) + __builtin_fabsf(w[9] * r.dy) + __builtin_fabsf(w[10] * r.dz));
    // ray.rs:87-93: origin pushed along d by the error, t_max kept
    const float l2 = dx * dx + dy * dy + dz * dz;
    if (l2 > 0.0f) {
        const float dt = (__builtin_fabsf(dx) * ex + __builtin_fabsf(dy) * ey + __builtin_fabsf(dz) * ez) / l2;
        ox = ox + dx * dt; oy = oy + dy * dt; oz = oz + dz * dt;
    }
    const EF eox = ef(ox, ex), eoy = ef(oy, ey), eoz = ef(oz, ez);
    const EF edx = ef(dx, fx), edy = ef(dy, fy), edz = ef(dz, fz);
    const EF a = ef_add(ef_add(ef_mul(edx, edx), ef_mul(edy, edy)), ef_mul(edz, edz));
    const EF b = ef_mul(ef(2.0f, 0.0f), ef_add(ef_add(ef_mul(edx, eox), ef_mul(edy, eoy)), ef_mul(edz, eoz)));
    const EF rr = ef(radius, 0.0f);
    const EF c = ef_sub(ef_add(ef_add(ef_mul(eox, eox), ef_mul(eoy, eoy)), ef_mul(eoz, eoz)), ef_mul(rr, rr));
    // EFloat::quadratic (efloat.rs:35-59)
    const double disc = (double)b.v * (double)b.v - 4.0 * (double)a.v * (double)c.v;
    if (disc < 0.0) return false;
    const float root = (float)__builtin_sqrt(disc);
    const EF rootf = ef(root, kMachineEps * root);
    const EF q = (b.v < 0.0f) ? ef_mul(ef(-0.5f, 0.0f), ef_sub(b, rootf)) : ef_mul(ef(-0.5f, 0.0f), ef_add(b, rootf));
    EF t0 = ef_div(q, a), t1 = ef_div(c, q);
    if (t0.v > t1.v) { EF tmp = t0; t0 = t1; t1 = tmp; }
    const float tmax = r.t_max;
    if (t0.hi > tmax || t1.lo <= 0.0f) return false;
    EF ts = t0;
    if (ts.lo <= 0.0f) {
        ts = t1;
        if (ts.hi > tmax) return false;
    }
    SphereHit h;
    sphere_point(ox, oy, oz, dx, dy, dz, ts.v, radius, &h);
    const bool clip = (z_min > -radius && h.pz < z_min) || (z_max < radius && h.pz > z_max) || h.phi > phi_max;
    if (clip) {
        if (ef_eq(ts, t1)) return false;
        if (t1.hi > tmax) return false;
        ts = t1;
        sphere_point(ox, oy, oz, dx, dy, dz, ts.v, radius, &h);
        if ((z_min > -radius && h.pz < z_min) || (z_max < radius && h.pz > z_max) || h.phi > phi_max) return false;
    }
    h.t = ts.v;
    *out = h;
    return true;
}

}  // namespace dev
}  // namespace bg
