// Oracle (test infrastructure only): triangle and sphere intersection restated
// from bonsai's shapes/triangle.rs and shapes/sphere.rs, plus the surface
// interaction each builds (base/interaction.rs).
#pragma once

#include "rt.h"

namespace orc {

// The parts of base/interaction.rs:47-54 the path integrator reads.
struct Isect {
    V3 p, perr, n;        // point, point_error, geometric normal
    V3 ns, dpdu_s;        // shading normal, shading dpdu (BSDF frame)
    V3 wo;                // -ray.direction
    Float u, v;           // surface uv
    Float time;
};

// ------------------------------------------------------------ triangle.rs
struct TriVerts { V3 p0, p1, p2; };

// Default parameterisation when the mesh has no uv array: triangle.rs:107-117.
static const Float TRI_UV[3][2] = {{0.0f, 0.0f}, {1.0f, 0.0f}, {1.0f, 1.0f}};

// triangle.rs:135-240 and 390-491: the watertight predicate shared by
// intersect and intersect_test. Returns the barycentrics and t on success.
inline bool tri_core(const TriVerts& tv, const Ray& ray, Float* b_out, Float* t_out) {
    const V3 &p0 = tv.p0, &p1 = tv.p1, &p2 = tv.p2;
    V3 p0t = p0 - ray.o, p1t = p1 - ray.o, p2t = p2 - ray.o;
    int kz = max_dimension(vabs(ray.d));
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    V3 d = permute(ray.d, kx, ky, kz);
    p0t = permute(p0t, kx, ky, kz);
    p1t = permute(p1t, kx, ky, kz);
    p2t = permute(p2t, kx, ky, kz);
    Float sx = -d.x / d.z, sy = -d.y / d.z, sz = 1.0f / d.z;
    p0t.x += sx * p0t.z; p0t.y += sy * p0t.z;
    p1t.x += sx * p1t.z; p1t.y += sy * p1t.z;
    p2t.x += sx * p2t.z; p2t.y += sy * p2t.z;
    Float e0 = p1t.x * p2t.y - p1t.y * p2t.x;
    Float e1 = p2t.x * p0t.y - p2t.y * p0t.x;
    Float e2 = p0t.x * p1t.y - p0t.y * p1t.x;
    // triangle.rs:179-190: f64 fallback (note the operand order, kept as written).
    if (e0 == 0.0f || e1 == 0.0f || e2 == 0.0f) {
        double p2txp1ty = (double)p2t.x * (double)p1t.y;
        double p2typ1tx = (double)p2t.y * (double)p1t.x;
        e0 = (Float)(p2typ1tx - p2txp1ty);
        double p0txp2ty = (double)p0t.x * (double)p2t.y;
        double p0typ2tx = (double)p0t.y * (double)p2t.x;
        e1 = (Float)(p0typ2tx - p0txp2ty);
        double p1txp0ty = (double)p1t.x * (double)p0t.y;
        double p1typ0tx = (double)p1t.y * (double)p0t.x;
        e2 = (Float)(p1typ0tx - p1txp0ty);
    }
    if ((e0 < 0.0f || e1 < 0.0f || e2 < 0.0f) && (e0 > 0.0f || e1 > 0.0f || e2 > 0.0f)) return false;
    Float det = e0 + e1 + e2;
    if (det == 0.0f) return false;
    p0t.z *= sz; p1t.z *= sz; p2t.z *= sz;
    Float t_scaled = e0 * p0t.z + e1 * p1t.z + e2 * p2t.z;
    if (det < 0.0f && (t_scaled >= 0.0f || t_scaled < ray.t_max * det)) return false;
    else if (det > 0.0f && (t_scaled <= 0.0f || t_scaled > ray.t_max * det)) return false;
    Float inv_det = 1.0f / det;
    Float b0 = e0 * inv_det, b1 = e1 * inv_det, b2 = e2 * inv_det;
    Float t = t_scaled * inv_det;
    Float max_zt = max_component(vabs(V3(p0t.z, p1t.z, p2t.z)));
    Float delta_z = gamma(3.0f) * max_zt;
    Float max_xt = max_component(vabs(V3(p0t.x, p1t.x, p2t.x)));
    Float max_yt = max_component(vabs(V3(p0t.y, p1t.y, p2t.y)));
    Float delta_x = gamma(5.0f) * (max_xt + max_zt);
    Float delta_y = gamma(5.0f) * (max_yt + max_zt);
    Float delta_e = 2.0f * (gamma(2.0f) * max_xt * max_yt + delta_y * max_xt + delta_x * max_yt);
    Float max_e = max_component(vabs(V3(e0, e1, e2)));
    Float delta_t = 3.0f * (gamma(3.0f) * max_e * max_zt + delta_e * max_zt + delta_z * max_e) * std::fabs(inv_det);
    if (t <= delta_t) return false;
    b_out[0] = b0; b_out[1] = b1; b_out[2] = b2;
    *t_out = t;
    return true;
}

// triangle.rs:241-268 / 493-517: partial derivatives from the (default) uvs,
// with the degenerate-triangle exit. Returns false for a degenerate triangle.
inline bool tri_dpduv(const TriVerts& tv, V3* dpdu, V3* dpdv) {
    const V3 &p0 = tv.p0, &p1 = tv.p1, &p2 = tv.p2;
    Float duv02[2] = {TRI_UV[0][0] - TRI_UV[2][0], TRI_UV[0][1] - TRI_UV[2][1]};
    Float duv12[2] = {TRI_UV[1][0] - TRI_UV[2][0], TRI_UV[1][1] - TRI_UV[2][1]};
    V3 dp02 = p0 - p2, dp12 = p1 - p2;
    Float det = duv02[0] * duv12[1] - duv02[1] * duv12[0];
    bool degenerate_uv = std::fabs(det) < 1e-8f;
    *dpdu = V3(); *dpdv = V3();
    if (!degenerate_uv) {
        Float inv = 1.0f / det;
        *dpdu = (duv12[1] * dp02 - duv02[1] * dp12) * inv;
        *dpdv = (-duv12[0] * dp02 + duv02[0] * dp12) * inv;
    }
    if (degenerate_uv || length_squared(cross(*dpdu, *dpdv)) == 0.0f) {
        V3 ng = cross(p2 - p0, p1 - p0);
        if (length_squared(ng) == 0.0f) return false;
        coordinate_system(normalize(ng), dpdu, dpdv);
    }
    return true;
}

// triangle.rs:390-520
inline bool tri_intersect_test(const TriVerts& tv, const Ray& ray, Float* t_out) {
    Float b[3], t;
    if (!tri_core(tv, ray, b, &t)) return false;
    V3 dpdu, dpdv;
    if (!tri_dpduv(tv, &dpdu, &dpdv)) return false;
    if (t_out) *t_out = t;
    return true;
}

// triangle.rs:135-388 (mesh without normals/tangents/uvs). `flip` is
// reverse_orientation ^ transform_swaps_handedness.
inline bool tri_intersect(const TriVerts& tv, const Ray& ray, bool flip, Float* t_hit, Float* bary, Isect* it) {
    Float b[3], t;
    if (!tri_core(tv, ray, b, &t)) return false;
    V3 dpdu, dpdv;
    if (!tri_dpduv(tv, &dpdu, &dpdv)) return false;
    const V3 &p0 = tv.p0, &p1 = tv.p1, &p2 = tv.p2;
    if (it) {
        Float xs = std::fabs(b[0] * p0.x) + std::fabs(b[1] * p1.x) + std::fabs(b[2] * p2.x);
        Float ys = std::fabs(b[0] * p0.y) + std::fabs(b[1] * p1.y) + std::fabs(b[2] * p2.y);
        Float zs = std::fabs(b[0] * p0.z) + std::fabs(b[1] * p1.z) + std::fabs(b[2] * p2.z);
        it->perr = gamma(7.0f) * V3(xs, ys, zs);
        it->p = b[0] * p0 + b[1] * p1 + b[2] * p2;
        it->u = b[0] * TRI_UV[0][0] + b[1] * TRI_UV[1][0] + b[2] * TRI_UV[2][0];
        it->v = b[0] * TRI_UV[0][1] + b[1] * TRI_UV[1][1] + b[2] * TRI_UV[2][1];
        it->wo = -ray.d;
        it->time = ray.time;
        // triangle.rs:300-307: the geometric normal replaces the one built by Interaction::new.
        V3 dp02 = p0 - p2, dp12 = p1 - p2;
        V3 n = normalize(cross(dp02, dp12));
        if (flip) n = -n;
        it->n = n;
        it->ns = n;
        it->dpdu_s = dpdu;
    }
    if (bary) { bary[0] = b[0]; bary[1] = b[1]; bary[2] = b[2]; }
    *t_hit = t;
    return true;
}

// triangle.rs:561-566
inline Float tri_area(const TriVerts& tv) { return 0.5f * length(cross(tv.p1 - tv.p0, tv.p2 - tv.p0)); }

// sampling.rs:233-236
inline void uniform_sample_triangle(Float u0, Float u1, Float* b0, Float* b1) {
    Float s = std::sqrt(u0);
    *b0 = 1.0f - s;
    *b1 = u1 * s;
}

// triangle.rs:522-559 (no mesh normals). Fills point, error and normal.
inline void tri_sample(const TriVerts& tv, bool flip, Float u0, Float u1, V3* p, V3* perr, V3* n, Float* pdf) {
    Float b0, b1;
    uniform_sample_triangle(u0, u1, &b0, &b1);
    Float b2 = 1.0f - b0 - b1;
    *p = b0 * tv.p0 + b1 * tv.p1 + b2 * tv.p2;
    V3 nn = normalize(cross(tv.p1 - tv.p0, tv.p2 - tv.p0));
    if (flip) nn = nn * -1.0f;
    *n = nn;
    V3 pabs = vabs(b0 * tv.p0) + vabs(b1 * tv.p1) + vabs(b2 * tv.p2);
    *perr = gamma(6.0f) * pabs;
    *pdf = 1.0f / tri_area(tv);
}

// ------------------------------------------------------------ sphere.rs
struct Sphere {
    Transform o2w, w2o;
    bool reverse_orientation, swaps;
    Float radius, z_min, z_max, theta_min, theta_max, phi_max;
};

// Rust f32::clamp
inline Float clampf(Float v, Float lo, Float hi) { return v < lo ? lo : (v > hi ? hi : v); }

// sphere.rs:39-65
inline Sphere make_sphere(const Transform& o2w, bool reverse, Float radius, Float zmin, Float zmax, Float phimax_deg) {
    Sphere s;
    s.o2w = o2w;
    s.w2o = is_identity(o2w) ? o2w : tinverse(o2w);
    s.swaps = swaps_handedness(o2w);
    s.reverse_orientation = reverse;
    s.radius = radius;
    s.z_min = clampf(fmin_r(zmin, zmax), -radius, radius);
    s.z_max = clampf(fmax_r(zmin, zmax), -radius, radius);
    s.theta_min = std::acos(clampf(fmin_r(zmin, zmax) / radius, -1.0f, 1.0f));
    s.theta_max = std::acos(clampf(fmax_r(zmin, zmax) / radius, -1.0f, 1.0f));
    s.phi_max = to_radians(clampf(phimax_deg, 0.0f, 360.0f));
    return s;
}

// sphere.rs:76-78 with transform.rs:75-85
inline Bounds3 sphere_world_bounds(const Sphere& s) {
    V3 mn(-s.radius, -s.radius, s.z_min), mx(s.radius, s.radius, s.z_max);
    Bounds3 ob(mn, mx);
    const Transform& t = s.o2w;
    Bounds3 r;
    V3 c0 = xform_point(t, V3(ob.mn.x, ob.mn.y, ob.mn.z));
    r.mn = c0; r.mx = c0;
    r = bunion(r, xform_point(t, V3(ob.mx.x, ob.mn.y, ob.mn.z)));
    r = bunion(r, xform_point(t, V3(ob.mn.x, ob.mx.y, ob.mn.z)));
    r = bunion(r, xform_point(t, V3(ob.mn.x, ob.mn.y, ob.mx.z)));
    r = bunion(r, xform_point(t, V3(ob.mn.x, ob.mx.y, ob.mx.z)));
    r = bunion(r, xform_point(t, V3(ob.mx.x, ob.mx.y, ob.mn.z)));
    r = bunion(r, xform_point(t, V3(ob.mx.x, ob.mn.y, ob.mx.z)));
    r = bunion(r, xform_point(t, V3(ob.mx.x, ob.mx.y, ob.mx.z)));
    return r;
}

// Refine + phi for a candidate t (sphere.rs:119-130).
inline void sphere_point_phi(const Sphere& s, const Ray& r, Float t, V3* p, Float* phi) {
    V3 ph = r.o + r.d * t;
    ph = ph * (s.radius / distance(ph, V3()));
    if (ph.x == 0.0f && ph.y == 0.0f) ph.x = 1e-5f * s.radius;
    Float f = std::atan2(ph.y, ph.x);
    if (f < 0.0f) f += 2.0f * PI;
    *p = ph; *phi = f;
}

inline bool sphere_clipped(const Sphere& s, const V3& p, Float phi) {
    return (s.z_min > -s.radius && p.z < s.z_min) || (s.z_max < s.radius && p.z > s.z_max) || phi > s.phi_max;
}

// sphere.rs:80-163 / 239-326: the shared hit decision. Returns the object-space
// ray, the accepted t interval and the refined object-space point/phi.
inline bool sphere_core(const Sphere& s, const Ray& wray, Ray* oray, EFloat* thit, V3* phit, Float* phi) {
    V3 oerr, derr;
    Ray ray = xform_ray_err(s.w2o, wray, &oerr, &derr);
    EFloat ox(ray.o.x, oerr.x), oy(ray.o.y, oerr.y), oz(ray.o.z, oerr.z);
    EFloat dx(ray.d.x, derr.x), dy(ray.d.y, derr.y), dz(ray.d.z, derr.z);
    EFloat a = dx * dx + dy * dy + dz * dz;
    EFloat b = 2.0f * (dx * ox + dy * oy + dz * oz);
    EFloat c = ox * ox + oy * oy + oz * oz - EFloat(s.radius, 0.0f) * EFloat(s.radius, 0.0f);
    EFloat t0, t1;
    if (!quadratic(a, b, c, &t0, &t1)) return false;
    if (t0.high > ray.t_max || t1.low <= 0.0f) return false;
    EFloat ts = t0;
    if (ts.low <= 0.0f) {
        ts = t1;
        if (ts.high > ray.t_max) return false;
    }
    V3 p; Float f;
    sphere_point_phi(s, ray, ts.v, &p, &f);
    if (sphere_clipped(s, p, f)) {
        if (ts == t1) return false;
        if (t1.high > ray.t_max) return false;
        ts = t1;
        sphere_point_phi(s, ray, ts.v, &p, &f);
        if (sphere_clipped(s, p, f)) return false;
    }
    *oray = ray; *thit = ts; *phit = p; *phi = f;
    return true;
}

inline bool sphere_intersect_test(const Sphere& s, const Ray& r, Float* t_out) {
    Ray oray; EFloat th; V3 p; Float phi;
    if (!sphere_core(s, r, &oray, &th, &p, &phi)) return false;
    if (t_out) *t_out = th.v;
    return true;
}

// sphere.rs:80-237: hit plus interaction (uv, dpdu/dpdv, normal) in world space.
inline bool sphere_intersect(const Sphere& s, const Ray& r, Float* t_hit, Float* uv, Isect* it) {
    Ray ray; EFloat th; V3 ph; Float phi;
    if (!sphere_core(s, r, &ray, &th, &ph, &phi)) return false;
    Float u = phi / s.phi_max;
    Float theta = std::acos(clampf(ph.z / s.radius, -1.0f, 1.0f));
    Float v = (theta - s.theta_min) / (s.theta_max - s.theta_min);
    if (uv) { uv[0] = u; uv[1] = v; }
    if (it) {
        Float z_radius = std::sqrt(ph.x * ph.x + ph.y * ph.y);
        Float inv_z_radius = 1.0f / z_radius;
        Float cos_phi = ph.x * inv_z_radius, sin_phi = ph.y * inv_z_radius;
        V3 dpdu(-s.phi_max * ph.y, s.phi_max * ph.x, 0.0f);
        V3 dpdv = (s.theta_max - s.theta_min) * V3(ph.z * cos_phi, ph.z * sin_phi, -s.radius * std::sin(theta));
        V3 perr = gamma(5.0f) * vabs(ph);
        // Interaction::new (interaction.rs:65-69)
        V3 n = normalize(cross(dpdu, dpdv));
        if (s.reverse_orientation ^ s.swaps) n = n * -1.0f;
        // Interaction::transform (interaction.rs:268-300)
        V3 aerr;
        it->p = xform_point_perr(s.o2w, ph, perr, &aerr);
        it->perr = aerr;
        it->wo = xform_vec(s.o2w, -ray.d);
        it->n = normalize(xform_normal(s.o2w, n));
        it->ns = face_forward(normalize(xform_normal(s.o2w, n)), it->n);
        it->dpdu_s = xform_vec(s.o2w, dpdu);
        it->u = u; it->v = v;
        it->time = r.time;
    }
    *t_hit = th.v;
    return true;
}

inline Float sphere_area(const Sphere& s) { return s.phi_max * s.radius * (s.z_max - s.z_min); }

}  // namespace orc
