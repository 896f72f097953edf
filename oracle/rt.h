// Oracle (test infrastructure only): a CPU restatement of bonsai's geometry and
// intersection numerics. Nothing in the shipped library includes or links this
// file; it exists so tests can check the HIP path against an independent,
// line-by-line restatement of the Rust reference.
//
// Every function follows the reference file:line cited beside it, keeping the
// reference's operation order and parenthesisation. Build with
// -ffp-contract=off -fno-fast-math so each f32 operation rounds once, as Rust's
// f32 arithmetic does.
#pragma once

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>

namespace orc {

typedef float Float;

// constants.rs:7 (std::f32::consts::PI), constants.rs:15-16
static const Float PI = 3.14159265358979323846f;
static const Float MACHINE_EPSILON = FLT_EPSILON * 0.5f;
static const Float ONE_MINUS_EPSILON = 1.0f - FLT_EPSILON;
static const Float INV_PI = 1.0f / PI;
static const Float PI_OVER_TWO = PI / 2.0f;

inline uint32_t f2u(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float u2f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

// math.rs:3-5
inline Float gamma(Float n) { return (n * MACHINE_EPSILON) / (1.0f - n * MACHINE_EPSILON); }

// math.rs:27-41
inline Float next_float_down(Float v) {
    if (std::isinf(v) && v < 0.0f) return v;
    if (v == 0.0f) v = -0.0f;
    uint32_t ui = f2u(v);
    if (v > 0.0f) ui -= 1; else ui += 1;
    return u2f(ui);
}

// math.rs:43-57
inline Float next_float_up(Float v) {
    if (std::isinf(v) && v > 0.0f) return v;
    if (v == -0.0f) v = 0.0f;
    uint32_t ui = f2u(v);
    if (v >= 0.0f) ui += 1; else ui -= 1;
    return u2f(ui);
}

// Rust f32::min / f32::max: a NaN operand yields the other operand. For equal
// operands (including +0/-0) the second operand is returned.
inline Float fmin_r(Float a, Float b) {
    if (a != a) return b;
    if (b != b) return a;
    return a < b ? a : b;
}
inline Float fmax_r(Float a, Float b) {
    if (a != a) return b;
    if (b != b) return a;
    return a > b ? a : b;
}

// lerp math.rs:23-25
inline Float lerp(Float t, Float a, Float b) { return (1.0f - t) * a + t * b; }

// geometries/{vec3,point3,normal}.rs share one arithmetic; one struct serves all three.
struct V3 {
    Float x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(Float a, Float b, Float c) : x(a), y(b), z(c) {}
    Float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    Float& at(int i) { return i == 0 ? x : (i == 1 ? y : z); }
};

inline V3 operator+(const V3& a, const V3& b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(const V3& a, const V3& b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator-(const V3& a) { return V3(-a.x, -a.y, -a.z); }
inline V3 operator*(const V3& a, Float s) { return V3(a.x * s, a.y * s, a.z * s); }
inline V3 operator*(Float s, const V3& a) { return V3(s * a.x, s * a.y, s * a.z); }
// vec3.rs:330-342: division is a multiply by the reciprocal.
inline V3 operator/(const V3& a, Float s) { Float inv = 1.0f / s; return V3(a.x * inv, a.y * inv, a.z * inv); }
inline bool operator==(const V3& a, const V3& b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

// vec3.rs:108-122: cross product evaluated in f64, rounded to f32 per component.
inline V3 cross(const V3& a, const V3& v) {
    double x = a.x, y = a.y, z = a.z;
    double vx = v.x, vy = v.y, vz = v.z;
    return V3((Float)((y * vz) - (z * vy)), (Float)((z * vx) - (x * vz)), (Float)((x * vy) - (y * vx)));
}
// vec3.rs:124-127
inline Float dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline Float abs_dot(const V3& a, const V3& b) { return std::fabs(dot(a, b)); }
// vec3.rs:103-109
inline Float length_squared(const V3& a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline Float length(const V3& a) { return std::sqrt(length_squared(a)); }
// vec3.rs:135-137
inline V3 normalize(const V3& a) { return a / length(a); }
inline V3 vabs(const V3& a) { return V3(std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)); }
// vec3.rs:147-149
inline Float max_component(const V3& a) { return fmax_r(a.x, fmax_r(a.y, a.z)); }
// vec3.rs:159-171
inline int max_dimension(const V3& a) {
    if (a.x > a.y) return a.x > a.z ? 0 : 2;
    return a.y > a.z ? 1 : 2;
}
inline V3 permute(const V3& a, int i, int j, int k) { return V3(a[i], a[j], a[k]); }
// vec3.rs:95-101 (face_forward) and normal.rs:32-38
inline V3 face_forward(const V3& n, const V3& v) { return dot(n, v) < 0.0f ? -n : n; }
inline Float distance_squared(const V3& a, const V3& b) { return length_squared(a - b); }
inline Float distance(const V3& a, const V3& b) { return length(a - b); }

// vec3.rs:25-33
inline void coordinate_system(const V3& v1, V3* v2, V3* v3) {
    if (std::fabs(v1.x) > std::fabs(v1.y))
        *v2 = V3(-v1.z, 0.0f, v1.x) / std::sqrt(v1.x * v1.x + v1.z * v1.z);
    else
        *v2 = V3(0.0f, v1.z, -v1.y) / std::sqrt(v1.y * v1.y + v1.z * v1.z);
    *v3 = cross(v1, *v2);
}

// point3.rs:31-49
inline V3 offset_ray_origin(const V3& p, const V3& perr, const V3& n, const V3& w) {
    V3 offset = dot(vabs(n), perr) * n;
    if (dot(w, n) < 0.0f) offset = -offset;
    V3 po = p + offset;
    for (int i = 0; i < 3; ++i) {
        if (offset[i] > 0.0f) po.at(i) = next_float_up(po[i]);
        else if (offset[i] < 0.0f) po.at(i) = next_float_down(po[i]);
    }
    return po;
}

// ---------------------------------------------------------------- bounds3.rs
struct Bounds3 {
    V3 mn, mx;
    // bounds3.rs:291-298 (Float::MAX / Float::MIN = most negative finite)
    Bounds3() : mn(FLT_MAX, FLT_MAX, FLT_MAX), mx(-FLT_MAX, -FLT_MAX, -FLT_MAX) {}
    // bounds3.rs:18-23
    Bounds3(const V3& a, const V3& b)
        : mn(fmin_r(a.x, b.x), fmin_r(a.y, b.y), fmin_r(a.z, b.z)),
          mx(fmax_r(a.x, b.x), fmax_r(a.y, b.y), fmax_r(a.z, b.z)) {}
    const V3& operator[](int i) const { return i == 0 ? mn : mx; }
};
// bounds3.rs:25-52
inline Bounds3 bunion(const Bounds3& a, const Bounds3& b) {
    Bounds3 r;
    r.mn = V3(fmin_r(a.mn.x, b.mn.x), fmin_r(a.mn.y, b.mn.y), fmin_r(a.mn.z, b.mn.z));
    r.mx = V3(fmax_r(a.mx.x, b.mx.x), fmax_r(a.mx.y, b.mx.y), fmax_r(a.mx.z, b.mx.z));
    return r;
}
// bounds3.rs:54-81
inline Bounds3 bunion(const Bounds3& a, const V3& p) {
    Bounds3 r;
    r.mn = V3(fmin_r(a.mn.x, p.x), fmin_r(a.mn.y, p.y), fmin_r(a.mn.z, p.z));
    r.mx = V3(fmax_r(a.mx.x, p.x), fmax_r(a.mx.y, p.y), fmax_r(a.mx.z, p.z));
    return r;
}
inline V3 diagonal(const Bounds3& b) { return b.mx - b.mn; }
// bounds3.rs:160-163
inline Float surface_area(const Bounds3& b) {
    V3 d = diagonal(b);
    return 2.0f * (d.x * d.y + d.x * d.z + d.y * d.z);
}
// bounds3.rs:175-184
inline int maximum_extent(const Bounds3& b) {
    V3 d = diagonal(b);
    if (d.x > d.y && d.x > d.z) return 0;
    if (d.y > d.z) return 1;
    return 2;
}
// bounds3.rs:194-206 (Vec3 /= Float is a reciprocal multiply)
inline V3 boffset(const Bounds3& b, const V3& p) {
    V3 o = p - b.mn;
    if (b.mx.x > b.mn.x) o.x = o.x * (1.0f / (b.mx.x - b.mn.x));
    if (b.mx.y > b.mn.y) o.y = o.y * (1.0f / (b.mx.y - b.mn.y));
    if (b.mx.z > b.mn.z) o.z = o.z * (1.0f / (b.mx.z - b.mn.z));
    return o;
}

struct Ray {
    V3 o, d;
    Float t_max;
    Float time;
};

// bounds3.rs:246-288
inline bool intersect_range_precomputed(const Bounds3& b, const Ray& ray, const V3& inv, const int neg[3]) {
    const Float pad = 1.0f + 2.0f * gamma(3.0f);
    Float t_min = (b[neg[0]].x - ray.o.x) * inv.x;
    Float t_max = (b[1 - neg[0]].x - ray.o.x) * inv.x;
    Float ty_min = (b[neg[1]].y - ray.o.y) * inv.y;
    Float ty_max = (b[1 - neg[1]].y - ray.o.y) * inv.y;
    t_max *= pad;
    ty_max *= pad;
    if (t_min > ty_max || ty_min > t_max) return false;
    if (ty_min > t_min) t_min = ty_min;
    if (ty_max < t_max) t_max = ty_max;
    Float tz_min = (b[neg[2]].z - ray.o.z) * inv.z;
    Float tz_max = (b[1 - neg[2]].z - ray.o.z) * inv.z;
    tz_max *= pad;
    if (t_min > tz_max || tz_min > t_max) return false;
    if (tz_min > t_min) t_min = tz_min;
    if (tz_max < t_max) t_max = tz_max;
    return (t_min < ray.t_max) && (t_max > 0.0f);
}

// ------------------------------------------------------------------ efloat.rs
struct EFloat {
    Float v, low, high;
    EFloat() : v(0), low(0), high(0) {}
    // efloat.rs:19-33
    EFloat(Float val, Float err) : v(val) {
        if (err == 0.0f) { low = val; high = val; }
        else { low = next_float_down(val - err); high = next_float_up(val + err); }
    }
    static EFloat raw(Float v, Float lo, Float hi) { EFloat e; e.v = v; e.low = lo; e.high = hi; return e; }
};
inline bool operator==(const EFloat& a, const EFloat& b) { return a.v == b.v && a.low == b.low && a.high == b.high; }
// efloat.rs:144-160
inline EFloat operator+(const EFloat& a, const EFloat& b) {
    return EFloat::raw(a.v + b.v, next_float_down(a.low + b.low), next_float_up(a.high + b.high));
}
// efloat.rs:170-184
inline EFloat operator-(const EFloat& a, const EFloat& b) {
    return EFloat::raw(a.v - b.v, next_float_down(a.low - b.high), next_float_up(a.high - b.low));
}
// efloat.rs:202-233
inline EFloat operator*(const EFloat& a, const EFloat& b) {
    Float p0 = a.low * b.low, p1 = a.high * b.low, p2 = a.low * b.high, p3 = a.high * b.high;
    return EFloat::raw(a.v * b.v,
                       next_float_down(fmin_r(fmin_r(fmin_r(p0, p1), p2), p3)),
                       next_float_up(fmax_r(fmax_r(fmax_r(p0, p1), p2), p3)));
}
// efloat.rs:243-285
inline EFloat operator/(const EFloat& a, const EFloat& b) {
    Float q0 = a.low / b.low, q1 = a.high / b.low, q2 = a.low / b.high, q3 = a.high / b.high;
    bool straddle = b.low < 0.0f && b.high > 0.0f;
    Float lo = straddle ? -INFINITY : next_float_down(fmin_r(fmin_r(fmin_r(q0, q1), q2), q3));
    Float hi = straddle ? INFINITY : next_float_up(fmax_r(fmax_r(fmax_r(q0, q1), q2), q3));
    return EFloat::raw(a.v / b.v, lo, hi);
}
inline EFloat operator*(Float s, const EFloat& b) { return EFloat(s, 0.0f) * b; }

// efloat.rs:35-59
inline bool quadratic(const EFloat& a, const EFloat& b, const EFloat& c, EFloat* t0, EFloat* t1) {
    double discrim = (double)b.v * (double)b.v - 4.0 * (double)a.v * (double)c.v;
    if (discrim < 0.0) return false;
    Float root = (Float)std::sqrt(discrim);
    EFloat rootf(root, MACHINE_EPSILON * root);
    EFloat q = (b.v < 0.0f) ? -0.5f * (b - rootf) : -0.5f * (b + rootf);
    *t0 = q / a;
    *t1 = c / q;
    if (t0->v > t1->v) { EFloat tmp = *t0; *t0 = *t1; *t1 = tmp; }
    return true;
}

// ------------------------------------------------------- mat4.rs / transform.rs
struct Mat4 {
    Float m[4][4];
    Mat4() { for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) m[i][j] = (i == j) ? 1.0f : 0.0f; }
};
// mat4.rs:37-48
inline Mat4 mul(const Mat4& a, const Mat4& b) {
    Mat4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            r.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j] + a.m[i][3] * b.m[3][j];
    return r;
}
// mat4.rs:71-143 (Gauss-Jordan with full pivoting)
inline Mat4 inverse(const Mat4& in) {
    Float mi[4][4];
    std::memcpy(mi, in.m, sizeof(mi));
    int ic[4] = {0, 0, 0, 0}, ir[4] = {0, 0, 0, 0}, piv[4] = {0, 0, 0, 0};
    for (int i = 0; i < 4; ++i) {
        int row = 0, col = 0;
        Float big = 0.0f;
        for (int j = 0; j < 4; ++j) {
            if (piv[j] != 1) {
                for (int k = 0; k < 4; ++k) {
                    if (piv[k] == 0) {
                        if (std::fabs(mi[j][k]) >= big) { big = std::fabs(mi[j][k]); row = j; col = k; }
                    }
                }
            }
        }
        piv[col] += 1;
        if (row != col)
            for (int k = 0; k < 4; ++k) { Float t = mi[row][k]; mi[row][k] = mi[col][k]; mi[col][k] = t; }
        ir[i] = row; ic[i] = col;
        Float pinv = 1.0f / mi[col][col];
        mi[col][col] = 1.0f;
        for (int j = 0; j < 4; ++j) mi[col][j] *= pinv;
        for (int j = 0; j < 4; ++j) {
            if (j != col) {
                Float save = mi[j][col];
                mi[j][col] = 0.0f;
                for (int k = 0; k < 4; ++k) mi[j][k] -= mi[col][k] * save;
            }
        }
    }
    for (int j = 3; j >= 0; --j) {
        if (ir[j] != ic[j])
            for (int k = 0; k < 4; ++k) { Float t = mi[k][ir[j]]; mi[k][ir[j]] = mi[k][ic[j]]; mi[k][ic[j]] = t; }
    }
    Mat4 r;
    std::memcpy(r.m, mi, sizeof(mi));
    return r;
}

struct Transform {
    Mat4 m, minv;
};
inline Mat4 mat4_from(const Float v[16]) { Mat4 r; for (int i = 0; i < 16; ++i) r.m[i / 4][i % 4] = v[i]; return r; }
// transform.rs:1629-1637
inline Transform tmul(const Transform& a, const Transform& b) { return Transform{mul(a.m, b.m), mul(b.minv, a.minv)}; }
inline Transform tinverse(const Transform& t) { return Transform{t.minv, t.m}; }
// transform.rs:87-97
inline Transform translate(const V3& d) {
    Transform t;
    t.m.m[0][3] = d.x; t.m.m[1][3] = d.y; t.m.m[2][3] = d.z;
    t.minv.m[0][3] = -d.x; t.minv.m[1][3] = -d.y; t.minv.m[2][3] = -d.z;
    return t;
}
// transform.rs:126-150
inline Transform scale(Float x, Float y, Float z) {
    Transform t;
    t.m.m[0][0] = x; t.m.m[1][1] = y; t.m.m[2][2] = z;
    t.minv.m[0][0] = 1.0f / x; t.minv.m[1][1] = 1.0f / y; t.minv.m[2][2] = 1.0f / z;
    return t;
}
// Rust f32::to_radians: self * (PI / 180.0)
inline Float to_radians(Float deg) { return deg * (PI / 180.0f); }
// transform.rs:48-73
inline Transform perspective(Float fov, Float n, Float f) {
    Mat4 p;
    p.m[2][2] = f / (f - n);
    p.m[2][3] = -f * n / (f - n);
    p.m[3][2] = 1.0f;
    p.m[3][3] = 0.0f;
    Transform persp{p, inverse(p)};
    Float inv_tan = 1.0f / std::tan(to_radians(fov) / 2.0f);
    return tmul(scale(inv_tan, inv_tan, 1.0f), persp);
}
// transform.rs:152-189
inline Transform look_at(const V3& pos, const V3& look, const V3& up) {
    Mat4 c2w;
    c2w.m[0][3] = pos.x; c2w.m[1][3] = pos.y; c2w.m[2][3] = pos.z; c2w.m[3][3] = 1.0f;
    V3 dir = normalize(look - pos);
    if (length(cross(normalize(up), dir)) == 0.0f) return Transform();
    V3 right = normalize(cross(normalize(up), dir));
    V3 new_up = cross(dir, right);
    c2w.m[0][0] = right.x; c2w.m[1][0] = right.y; c2w.m[2][0] = right.z; c2w.m[3][0] = 0.0f;
    c2w.m[0][1] = new_up.x; c2w.m[1][1] = new_up.y; c2w.m[2][1] = new_up.z; c2w.m[3][1] = 0.0f;
    c2w.m[0][2] = dir.x; c2w.m[1][2] = dir.y; c2w.m[2][2] = dir.z; c2w.m[3][2] = 0.0f;
    return Transform{c2w, inverse(c2w)};
}
// transform.rs:218-224
inline bool swaps_handedness(const Transform& t) {
    const Float (*m)[4] = t.m.m;
    Float det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
                m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    return det < 0.0f;
}
inline bool is_identity(const Transform& t) {
    for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) if (t.m.m[i][j] != (i == j ? 1.0f : 0.0f)) return false;
    return true;
}
// point3.rs:51-69
inline V3 xform_point(const Transform& t, const V3& p) {
    const Float (*m)[4] = t.m.m;
    Float xp = m[0][0] * p.x + m[0][1] * p.y + m[0][2] * p.z + m[0][3];
    Float yp = m[1][0] * p.x + m[1][1] * p.y + m[1][2] * p.z + m[1][3];
    Float zp = m[2][0] * p.x + m[2][1] * p.y + m[2][2] * p.z + m[2][3];
    Float wp = m[3][0] * p.x + m[3][1] * p.y + m[3][2] * p.z + m[3][3];
    if (wp == 1.0f) return V3(xp, yp, zp);
    return V3(xp, yp, zp) / wp;
}
// point3.rs:71-104
inline V3 xform_point_err(const Transform& t, const V3& p, V3* err) {
    const Float (*m)[4] = t.m.m;
    Float xp = (m[0][0] * p.x + m[0][1] * p.y) + (m[0][2] * p.z + m[0][3]);
    Float yp = (m[1][0] * p.x + m[1][1] * p.y) + (m[1][2] * p.z + m[1][3]);
    Float zp = (m[2][0] * p.x + m[2][1] * p.y) + (m[2][2] * p.z + m[2][3]);
    Float wp = (m[3][0] * p.x + m[3][1] * p.y) + (m[3][2] * p.z + m[3][3]);
    Float xs = std::fabs(m[0][0] * p.x) + std::fabs(m[0][1] * p.y) + std::fabs(m[0][2] * p.z) + std::fabs(m[0][3]);
    Float ys = std::fabs(m[1][0] * p.x) + std::fabs(m[1][1] * p.y) + std::fabs(m[1][2] * p.z) + std::fabs(m[1][3]);
    Float zs = std::fabs(m[2][0] * p.x) + std::fabs(m[2][1] * p.y) + std::fabs(m[2][2] * p.z) + std::fabs(m[2][3]);
    *err = gamma(3.0f) * V3(xs, ys, zs);
    if (wp == 1.0f) return V3(xp, yp, zp);
    return V3(xp, yp, zp) / wp;
}
// point3.rs:106-157
inline V3 xform_point_perr(const Transform& t, const V3& p, const V3& perr, V3* abs_err) {
    const Float (*m)[4] = t.m.m;
    Float xp = (m[0][0] * p.x + m[0][1] * p.y) + (m[0][2] * p.z + m[0][3]);
    Float yp = (m[1][0] * p.x + m[1][1] * p.y) + (m[1][2] * p.z + m[1][3]);
    Float zp = (m[2][0] * p.x + m[2][1] * p.y) + (m[2][2] * p.z + m[2][3]);
    Float wp = (m[3][0] * p.x + m[3][1] * p.y) + (m[3][2] * p.z + m[3][3]);
    Float g3 = gamma(3.0f);
    Float e[3];
    for (int r = 0; r < 3; ++r) {
        e[r] = (g3 + 1.0f) * (std::fabs(m[r][0]) * perr.x + std::fabs(m[r][1]) * perr.y + std::fabs(m[r][2]) * perr.z) +
               g3 * (std::fabs(m[r][0] * p.x) + std::fabs(m[r][1] * p.y) + std::fabs(m[r][2] * p.z) + std::fabs(m[r][3]));
    }
    *abs_err = V3(e[0], e[1], e[2]);
    if (wp == 1.0f) return V3(xp, yp, zp);
    return V3(xp, yp, zp) / wp;
}
// vec3.rs:35-44
inline V3 xform_vec(const Transform& t, const V3& v) {
    const Float (*m)[4] = t.m.m;
    return V3(m[0][0] * v.x + m[0][1] * v.y + m[0][2] * v.z,
              m[1][0] * v.x + m[1][1] * v.y + m[1][2] * v.z,
              m[2][0] * v.x + m[2][1] * v.y + m[2][2] * v.z);
}
// vec3.rs:46-64
inline V3 xform_vec_err(const Transform& t, const V3& v, V3* e) {
    const Float (*m)[4] = t.m.m;
    Float g3 = gamma(3.0f);
    e->x = g3 * (std::fabs(m[0][0] * v.x) + std::fabs(m[0][1] * v.y) + std::fabs(m[0][2] * v.z));
    e->y = g3 * (std::fabs(m[1][0] * v.x) + std::fabs(m[1][1] * v.y) + std::fabs(m[1][2] * v.z));
    e->z = g3 * (std::fabs(m[2][0] * v.x) + std::fabs(m[2][1] * v.y) + std::fabs(m[2][2] * v.z));
    return xform_vec(t, v);
}
// normal.rs:21-30 (uses the inverse matrix, transposed)
inline V3 xform_normal(const Transform& t, const V3& n) {
    const Float (*mi)[4] = t.minv.m;
    return V3(mi[0][0] * n.x + mi[1][0] * n.y + mi[2][0] * n.z,
              mi[0][1] * n.x + mi[1][1] * n.y + mi[2][1] * n.z,
              mi[0][2] * n.x + mi[1][2] * n.y + mi[2][2] * n.z);
}
// ray.rs:50-76
inline Ray xform_ray(const Transform& t, const Ray& r) {
    V3 oerr;
    V3 o = xform_point_err(t, r.o, &oerr);
    V3 d = xform_vec(t, r.d);
    Float l2 = length_squared(d);
    Float t_max = r.t_max;
    if (l2 > 0.0f) {
        Float dt = dot(vabs(d), oerr) / l2;
        o = o + d * dt;
        t_max -= dt;
    }
    return Ray{o, d, t_max, r.time};
}
// ray.rs:78-106 (t_max deliberately not reduced, as in the reference)
inline Ray xform_ray_err(const Transform& t, const Ray& r, V3* oerr, V3* derr) {
    V3 o = xform_point_err(t, r.o, oerr);
    V3 d = xform_vec_err(t, r.d, derr);
    Float l2 = length_squared(d);
    if (l2 > 0.0f) {
        Float dt = dot(vabs(d), *oerr) / l2;
        o = o + d * dt;
    }
    return Ray{o, d, r.t_max, r.time};
}

}  // namespace orc
