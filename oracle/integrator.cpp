// Oracle (test infrastructure only): see integrator.h.
#include "integrator.h"

#include <atomic>
#include <thread>

namespace orc {

namespace {

const uint32_t BSDF_REFLECTION = 1, BSDF_TRANSMISSION = 2, BSDF_DIFFUSE = 4, BSDF_GLOSSY = 8, BSDF_SPECULAR = 16;
const uint32_t BSDF_ALL = 31;

// ------------------------------------------------------------ bxdf.rs helpers
inline Float cos2_theta(const V3& w) { return w.z * w.z; }
inline Float abs_cos_theta(const V3& w) { return std::fabs(w.z); }
inline Float sin2_theta(const V3& w) { return fmax_r(0.0f, 1.0f - cos2_theta(w)); }
inline Float sin_theta(const V3& w) { return std::sqrt(sin2_theta(w)); }
inline Float tan_theta(const V3& w) { return sin_theta(w) / w.z; }
inline Float tan2_theta(const V3& w) { return sin2_theta(w) / cos2_theta(w); }
inline Float cos_phi(const V3& w) { Float v = sin_theta(w); return v == 0.0f ? 1.0f : clampf(w.x / v, -1.0f, 1.0f); }
inline Float sin_phi(const V3& w) { Float v = sin_theta(w); return v == 0.0f ? 0.0f : clampf(w.y / v, -1.0f, 1.0f); }
inline Float cos2_phi(const V3& w) { return cos_phi(w) * cos_phi(w); }
inline Float sin2_phi(const V3& w) { return sin_phi(w) * sin_phi(w); }
inline bool same_hemisphere(const V3& a, const V3& b) { return a.z * b.z > 0.0f; }
// bxdf.rs:144-146
inline V3 reflect(const V3& wo, const V3& n) { return -wo + 2.0f * dot(wo, n) * n; }

// sampling.rs:238-242
inline V3 cosine_sample_hemisphere(Float u0, Float u1) {
    Float dx, dy;
    concentric_sample_disk(u0, u1, &dx, &dy);
    Float z = std::sqrt(fmax_r(0.0f, 1.0f - dx * dx - dy * dy));
    return V3(dx, dy, z);
}

// bxdf.rs:190-217
RGB fresnel_conductor(Float cos_theta_i, const RGB& eta_i, const RGB& eta_t, const RGB& k) {
    RGB c(clampf(cos_theta_i, -1.0f, 1.0f));
    RGB eta = eta_t / eta_i;
    RGB eta_k = k / eta_i;
    RGB c2 = c * c;
    RGB s2 = RGB(1.0f) - c2;
    RGB eta2 = eta * eta;
    RGB etak2 = eta_k * eta_k;
    RGB t0 = eta2 - etak2 - s2;
    RGB a2b2 = sqrt_rgb(t0 * t0 + eta2 * etak2 * 4.0f);
    RGB t1 = a2b2 + c2;
    RGB a = sqrt_rgb((a2b2 + t0) * 0.5f);
    RGB t2 = c * a * 2.0f;
    RGB rs = (t1 - t2) / (t1 + t2);
    RGB t3 = c2 * a2b2 + s2 * s2;
    RGB t4 = t2 * s2;
    RGB rp = rs * (t3 - t4) / (t3 + t4);
    return (rp + rs) * 0.5f;
}

// base/bxdf.rs:164-187
Float fresnel_dielectric(Float cos_theta_i, Float eta_i, Float eta_t, bool fold) {
    cos_theta_i = clampf(cos_theta_i, -1.0f, 1.0f);
    if (!(cos_theta_i > 0.0f)) {
        std::swap(eta_i, eta_t);
        // the reference keeps the sign here (a pole at cos_i = -0.5547 for 1.5 / 1.0); the
        // corrected mode folds it, as the Fresnel equations do for light from the other side
        if (fold) cos_theta_i = std::fabs(cos_theta_i);
    }
    Float sin_theta_i = std::sqrt(fmax_r(0.0f, 1.0f - cos_theta_i * cos_theta_i));
    Float sin_theta_t = eta_i / eta_t * sin_theta_i;
    if (sin_theta_t >= 1.0f) return 1.0f;
    Float cos_theta_t = std::sqrt(fmax_r(0.0f, 1.0f - sin_theta_t * sin_theta_t));
    Float par = ((eta_t * cos_theta_i) - (eta_i * cos_theta_t)) / ((eta_t * cos_theta_i) + (eta_i * cos_theta_t));
    Float perp = ((eta_i * cos_theta_i) - (eta_t * cos_theta_t)) / ((eta_i * cos_theta_i) + (eta_t * cos_theta_t));
    return (par * par + perp * perp) / 2.0f;
}

// microfacet.rs:45-53
Float roughness_to_alpha(Float roughness) {
    roughness = fmax_r(roughness, 1e-3f);
    Float x = std::log(roughness);
    return 1.62142f + 0.819955f * x + 0.1734f * x * x + 0.0171201f * x * x * x + 0.000640711f * x * x * x * x;
}

// ---------------------------------------------- Trowbridge-Reitz (microfacet.rs)
struct TR {
    Float ax, ay;
    Mode mode;
};
// microfacet.rs:116-128
Float tr_d(const TR& d, const V3& wh) {
    Float t2 = tan2_theta(wh);
    if (std::isinf(t2)) return 0.0f;
    Float cos4 = cos2_theta(wh) * cos2_theta(wh);
    Float e = (cos2_phi(wh) / (d.ax * d.ax) + sin2_phi(wh) / (d.ay * d.ay)) * t2;
    return 1.0f / (PI * d.ax * d.ay * cos4 * (1.0f + e) * (1.0f + e));
}
// microfacet.rs:130-142
Float tr_lambda(const TR& d, const V3& w) {
    Float att = std::fabs(tan_theta(w));
    if (std::isinf(att)) return 0.0f;
    Float alpha = std::sqrt(cos2_phi(w) * d.ax * d.ax + sin2_phi(w) * d.ay * d.ay);
    Float a2t2 = (alpha * att) * (alpha * att);
    return (-1.0f + std::sqrt(1.0f + a2t2)) / 2.0f;
}
// microfacet.rs:17-19 (Q6: 1/(1/Lambda) = Lambda); corrected: 1/(1+Lambda)
Float tr_g1(const TR& d, const V3& w) {
    if (d.mode == MODE_REF) return 1.0f / (1.0f / tr_lambda(d, w));
    return 1.0f / (1.0f + tr_lambda(d, w));
}
// microfacet.rs:21-23
Float tr_g(const TR& d, const V3& wo, const V3& wi) { return 1.0f / (1.0f + tr_lambda(d, wo) + tr_lambda(d, wi)); }
// microfacet.rs:27-29
Float tr_pdf(const TR& d, const V3& wo, const V3& wh) { return tr_d(d, wh) * tr_g1(d, wo) * abs_dot(wo, wh) / abs_cos_theta(wo); }

// microfacet.rs:55-97 (Q7: the reference keeps slope_x_1 in both branches)
void sample11(Mode mode, Float cos_theta, Float u1, Float u2, Float* sx, Float* sy) {
    if (cos_theta > 0.9999f) {
        Float r = std::sqrt(u1 / (1.0f - u1));
        Float phi = 6.28318530718f * u2;
        *sx = r * std::cos(phi);
        *sy = r * std::sin(phi);
        return;
    }
    Float sin_t = std::sqrt(fmax_r(1.0f - cos_theta * cos_theta, 0.0f));
    Float tan_t = sin_t / cos_theta;
    Float a = 1.0f / tan_t;
    Float g1 = 2.0f / (1.0f + std::sqrt(1.0f + 1.0f / (a * a)));
    a = 2.0f * u1 / g1 - 1.0f;
    Float temp = 1.0f / (a * a - 1.0f);
    if (temp > 1e10f) temp = 1e10f;
    Float b = tan_t;
    Float dd = std::sqrt(fmax_r(b * b * temp * temp - (a * a - b * b) * temp, 0.0f));
    Float x1 = b * temp - dd;
    Float x2 = b * temp + dd;
    Float slope_x;
    if (a < 0.0f || x2 > 1.0f / tan_t) slope_x = x1;
    else slope_x = (mode == MODE_REF) ? x1 : x2;
    Float s;
    if (u2 > 0.5f) { s = 1.0f; u2 = 2.0f * (u2 - 0.5f); }
    else { s = -1.0f; u2 = 2.0f * (0.5f - u2); }
    Float z = (u2 * (u2 * (u2 * 0.27385f - 0.73369f) + 0.46341f)) /
              (u2 * (u2 * (u2 * 0.093073f + 0.309420f) - 1.000000f) + 0.597999f);
    *sx = slope_x;
    *sy = s * z * std::sqrt(1.0f + slope_x * slope_x);
}

// microfacet.rs:99-112
V3 tr_sample_vn(Mode mode, const V3& wi, Float ax, Float ay, Float u1, Float u2) {
    V3 ws = normalize(V3(ax * wi.x, ay * wi.y, wi.z));
    Float sx, sy;
    sample11(mode, ws.z, u1, u2, &sx, &sy);
    Float tmp = cos_phi(ws) * sx - sin_phi(ws) * sy;
    sy = sin_phi(ws) * sx + cos_phi(ws) * sy;
    sx = tmp;
    sx = ax * sx;
    sy = ay * sy;
    return normalize(V3(-sx, -sy, 1.0f));
}
// microfacet.rs:144-153
V3 tr_sample(const TR& d, const V3& wo, Float u0, Float u1) {
    bool flip = wo.z < 0.0f;
    V3 w = flip ? -wo : wo;
    V3 wh = tr_sample_vn(d.mode, w, d.ax, d.ay, u0, u1);
    return flip ? -wh : wh;
}

// ------------------------------------------------------------------- BSDF
struct Lobe {
    uint32_t type = 0;  // BxDF flags; 0 = no lobe
    RGB r;
    TR tr;
    RGB eta, k;               // FresnelConductor (eta_i = 1)
    bool dielectric = false;  // FresnelDielectric(eta_di, eta_dt) instead (fresnel.rs:19-22, 50-54)
    Float eta_di = 1.0f, eta_dt = 1.0f;
};

struct BxSample {
    V3 wi;
    RGB f;
    Float pdf = 0.0f;
    uint32_t type = 0;
};

// bxdfs/lambertian.rs:43-45, bxdf.rs:33-45, 72-78; bxdfs/microfacet.rs:69-147
RGB lobe_f(const Lobe& l, const V3& wo, const V3& wi) {
    if (l.type & BSDF_DIFFUSE) return l.r * (1.0f / PI);
    Float co = abs_cos_theta(wo), ci = abs_cos_theta(wi);
    V3 wh = wi + wo;
    if (ci == 0.0f || co == 0.0f) return RGB();
    if (wh.x == 0.0f && wh.y == 0.0f && wh.z == 0.0f) return RGB();
    wh = normalize(wh);
    const Float cwh = dot(wi, face_forward(wh, V3(0.0f, 0.0f, 1.0f)));
    RGB F = l.dielectric ? RGB(fresnel_dielectric(cwh, l.eta_di, l.eta_dt, l.tr.mode != MODE_REF))
                         : fresnel_conductor(std::fabs(cwh), RGB(1.0f), l.eta, l.k);
    return l.r * tr_d(l.tr, wh) * tr_g(l.tr, wo, wi) * F / (4.0f * ci * co);
}
Float lobe_pdf(const Lobe& l, const V3& wo, const V3& wi) {
    if (l.type & BSDF_DIFFUSE) return same_hemisphere(wo, wi) ? abs_cos_theta(wi) * (1.0f / PI) : 0.0f;
    if (!same_hemisphere(wo, wi)) return 0.0f;
    V3 wh = normalize(wo + wi);
    return tr_pdf(l.tr, wo, wh) / (4.0f * dot(wo, wh));
}
BxSample lobe_sample(const Lobe& l, const V3& wo, Float u0, Float u1) {
    BxSample s;
    if (l.type & BSDF_DIFFUSE) {
        V3 wi = cosine_sample_hemisphere(u0, u1);
        if (wo.z < 0.0f) wi.z *= -1.0f;
        s.wi = wi;
        s.f = lobe_f(l, wo, wi);
        s.pdf = lobe_pdf(l, wo, wi);
        return s;
    }
    if (wo.z == 0.0f) return s;
    V3 wh = tr_sample(l.tr, wo, u0, u1);
    if (dot(wo, wh) < 0.0f) return s;
    V3 wi = reflect(wo, wh);
    if (!same_hemisphere(wo, wi)) return s;
    s.wi = wi;
    s.f = lobe_f(l, wo, wi);
    s.pdf = tr_pdf(l.tr, wo, wh) / (4.0f * dot(wo, wh));
    return s;
}

inline bool matches(uint32_t type, uint32_t flags) { return (type & flags) == type; }

// bsdf.rs:16-259 with up to two lobes (plastic: Lambertian, then microfacet).
struct BSDF {
    V3 ns, ng, ss, ts;
    Lobe lobes[2];
    int n = 0;
    V3 to_local(const V3& v) const { return V3(dot(v, ss), dot(v, ts), dot(v, ns)); }
    V3 to_world(const V3& v) const {
        return V3(ss.x * v.x + ts.x * v.y + ns.x * v.z, ss.y * v.x + ts.y * v.y + ns.y * v.z,
                  ss.z * v.x + ts.z * v.y + ns.z * v.z);
    }
    void add(const Lobe& l) { lobes[n++] = l; }
    int num_components(uint32_t flags) const {
        int c = 0;
        for (int i = 0; i < n; ++i) c += matches(lobes[i].type, flags) ? 1 : 0;
        return c;
    }
    // bsdf.rs:84-104
    RGB f(const V3& wo_w, const V3& wi_w, uint32_t flags) const {
        V3 wo = to_local(wo_w), wi = to_local(wi_w);
        if (wo.z == 0.0f) return RGB();
        bool refl = dot(wi_w, ng) * dot(wo_w, ng) > 0.0f;
        RGB r;
        for (int i = 0; i < n; ++i) {
            const Lobe& l = lobes[i];
            if (matches(l.type, flags) &&
                ((refl && (l.type & BSDF_REFLECTION)) || (!refl && (l.type & BSDF_TRANSMISSION))))
                r = r + lobe_f(l, wo, wi);
        }
        return r;
    }
    // bsdf.rs:234-258
    Float pdf(const V3& wo_w, const V3& wi_w, uint32_t flags) const {
        if (n == 0) return 0.0f;
        V3 wo = to_local(wo_w), wi = to_local(wi_w);
        if (wo.z == 0.0f) return 0.0f;
        Float ret = 0.0f, comps = 0.0f;
        for (int i = 0; i < n; ++i)
            if (matches(lobes[i].type, flags)) {
                comps += 1.0f;
                ret += lobe_pdf(lobes[i], wo, wi);
            }
        return comps > 0.0f ? ret / comps : 0.0f;
    }
    // bsdf.rs:140-232: pick the component-th matching lobe with u0, remap u0, sample it; for a
    // non-specular pick, the pdf averages every matching lobe and f sums them
    BxSample sample(const V3& wo_w, Float u0, Float u1, uint32_t flags) const {
        BxSample out;
        int m = num_components(flags);
        if (m == 0) return out;
        uint32_t comp = (uint32_t)std::floor(u0 * (Float)m);
        if (comp > (uint32_t)m - 1) comp = (uint32_t)m - 1;
        int chosen = -1;
        uint32_t count = comp;
        for (int i = 0; i < n; ++i) {
            if (matches(lobes[i].type, flags) && count == 0) {
                chosen = i;
                break;
            }
            --count;
        }
        const Lobe& b = lobes[chosen];
        Float ur = fmin_r(u0 * (Float)m - (Float)comp, ONE_MINUS_EPSILON);
        V3 wo = to_local(wo_w);
        if (wo.z == 0.0f) return out;
        BxSample s = lobe_sample(b, wo, ur, u1);
        uint32_t stype = b.type;
        if (s.pdf == 0.0f) return out;
        V3 wi_w = to_world(s.wi);
        if (!(b.type & BSDF_SPECULAR) && m > 1)
            for (int i = 0; i < n; ++i)
                if (i != chosen && matches(lobes[i].type, flags)) s.pdf += lobe_pdf(lobes[i], wo, s.wi);
        if (m > 1) s.pdf /= (Float)m;
        if (!(b.type & BSDF_SPECULAR)) {
            bool refl = dot(wi_w, ng) * dot(wo_w, ng) > 0.0f;
            RGB f;
            for (int i = 0; i < n; ++i) {
                const Lobe& l = lobes[i];
                if (matches(l.type, flags) &&
                    ((refl && (l.type & BSDF_REFLECTION)) || (!refl && (l.type & BSDF_TRANSMISSION))))
                    f = f + lobe_f(l, wo, s.wi);
            }
            s.f = f;
        }
        out.wi = wi_w;
        out.f = s.f;
        out.pdf = s.pdf;
        out.type = stype;
        return out;
    }
};

// matte.rs:20-41 (sigma = 0) / conductor / plastic.rs:24-52; BSDF::new (bsdf.rs:33-52)
BSDF make_bsdf(const Isect& it, const Material& m, Mode mode) {
    BSDF b;
    b.ns = it.ns;
    b.ng = it.n;
    b.ss = normalize(it.dpdu_s);
    b.ts = cross(b.ns, b.ss);
    auto clamp0 = [](const RGB& c) {
        return RGB(clampf(c[0], 0.0f, INFINITY), clampf(c[1], 0.0f, INFINITY), clampf(c[2], 0.0f, INFINITY));
    };
    if (m.type == MAT_MATTE || m.type == MAT_PLASTIC) {
        RGB kd = clamp0(m.r);
        if (!is_black(kd)) {
            Lobe l;
            l.type = BSDF_REFLECTION | BSDF_DIFFUSE;
            l.r = kd;
            b.add(l);
        }
    }
    if (m.type == MAT_PLASTIC) {
        RGB ks = clamp0(m.eta);
        if (!is_black(ks)) {
            Float rough = m.alpha;
            if (m.k[0] != 0.0f) rough = roughness_to_alpha(rough);
            Lobe l;
            l.type = BSDF_REFLECTION | BSDF_GLOSSY;
            l.r = ks;
            l.tr = TR{fmax_r(rough, 0.001f), fmax_r(rough, 0.001f), mode};
            l.dielectric = true;
            l.eta_di = 1.5f;
            l.eta_dt = 1.0f;
            b.add(l);
        }
    } else if (m.type == MAT_CONDUCTOR) {
        Lobe l;
        l.type = BSDF_REFLECTION | BSDF_GLOSSY;
        l.r = m.r;
        l.tr = TR{fmax_r(m.alpha, 0.001f), fmax_r(m.alpha, 0.001f), mode};
        l.eta = m.eta;
        l.k = m.k;
        b.add(l);
    }
    return b;
}

// sampling.rs:248-252
inline Float power_heuristic(Float nf, Float fp, Float ng, Float gp) {
    Float f = nf * fp, g = ng * gp;
    return (f * f) / (f * f + g * g);
}

// Interaction::spawn_ray (interaction.rs:115-121). Q2: t_max = 1 - 1e-4 in REF.
inline Ray spawn_ray(const Isect& it, const V3& d, Mode mode) {
    V3 o = offset_ray_origin(it.p, it.perr, it.n, d);
    return Ray{o, d, mode == MODE_REF ? 1.0f - 0.0001f : INFINITY, it.time};
}
// interaction.rs:132-143
inline Ray spawn_ray_to(const V3& p0, const V3& e0, const V3& n0, Float time, const V3& p1, const V3& e1,
                        const V3& n1, V3* target = nullptr) {
    V3 o = offset_ray_origin(p0, e0, n0, p1 - p0);
    V3 t = offset_ray_origin(p1, e1, n1, o - p1);
    if (target) *target = t;
    return Ray{o, t - o, 1.0f - 0.0001f, time};
}

// diffuse.rs:163-169
inline RGB emission(const AreaLight& l, const V3& n, const V3& w) {
    return (l.double_sided || dot(n, w) > 0.0f) ? l.radiance : RGB();
}

struct LightSample {
    RGB radiance;
    V3 wi;
    Float pdf = 0.0f;
    V3 lp, lerr, ln;  // sampled point on the light (visibility target)
    bool vis = false;
};

// sampling.rs:211-216
inline V3 uniform_sample_sphere(Float u0, Float u1) {
    Float z = 1.0f - 2.0f * u0;
    Float r = std::sqrt(fmax_r(1.0f - z * z, 0.0f));
    Float phi = 2.0f * PI * u1;
    return V3(r * std::cos(phi), r * std::sin(phi), z);
}

// Sphere::sample (sphere.rs:328-354): a uniform point on the whole sphere (also a partial one).
// The reference transforms the normal of the unreprojected point without normalising it, so in
// REF its length is the radius (for a transform without scale); CORRECTED normalises it (pbrt-v3).
void sphere_sample_area(const Sphere& sp, Float u0, Float u1, Mode mode, V3* p, V3* perr, V3* n) {
    V3 obj = V3(0.0f, 0.0f, 0.0f) + sp.radius * uniform_sample_sphere(u0, u1);
    V3 nn = xform_normal(sp.o2w, obj);
    if (mode == MODE_CORRECTED) nn = normalize(nn);
    if (sp.reverse_orientation) nn = nn * -1.0f;
    obj = obj * (sp.radius / distance(obj, V3(0.0f, 0.0f, 0.0f)));
    V3 oerr = gamma(5.0f) * vabs(obj);
    *p = xform_point_perr(sp.o2w, obj, oerr, perr);
    *n = nn;
}

// The centre test of sphere.rs:356-363: whether the origin a ray from `it` toward the centre would
// have lies inside the sphere.
inline bool sphere_contains_origin(const Sphere& sp, const Isect& it, V3* center) {
    *center = xform_point(sp.o2w, V3(0.0f, 0.0f, 0.0f));
    V3 o = offset_ray_origin(it.p, it.perr, it.n, *center - it.p);
    return distance_squared(o, *center) <= sp.radius * sp.radius;
}

// sampling.rs:229-231
inline Float uniform_cone_pdf(Float cos_max) { return 1.0f / (2.0f * PI * (1.0f - cos_max)); }

// CORRECTED (pbrt-v3 Sphere::Sample(ref, u)) from outside: a direction uniform in the cone the
// sphere subtends, and the point of the sphere it sees first. Spheres are taken unscaled.
void sphere_sample_cone(const Sphere& sp, const V3& c, const V3& ref_p, Float u0, Float u1, V3* p, V3* perr, V3* n,
                        Float* pdf) {
    Float dc = distance(ref_p, c);
    Float inv_dc = 1.0f / dc;
    V3 wc = (c - ref_p) * inv_dc;
    V3 wcx, wcy;
    coordinate_system(wc, &wcx, &wcy);
    Float sin_max = sp.radius * inv_dc;
    Float sin_max2 = sin_max * sin_max;
    Float inv_sin_max = 1.0f / sin_max;
    Float cos_max = std::sqrt(fmax_r(0.0f, 1.0f - sin_max2));
    Float cos_t = (cos_max - 1.0f) * u0 + 1.0f;
    Float sin_t2 = 1.0f - cos_t * cos_t;
    if (sin_max2 < 0.00068523f) {  // sin^2(1.5 deg): small cones in the sin^2 parameterisation
        sin_t2 = sin_max2 * u0;
        cos_t = std::sqrt(1.0f - sin_t2);
    }
    // angle alpha at the centre between -wc and the point the sampled direction meets
    Float cos_a = sin_t2 * inv_sin_max + cos_t * std::sqrt(fmax_r(0.0f, 1.0f - sin_t2 * inv_sin_max * inv_sin_max));
    Float sin_a = std::sqrt(fmax_r(0.0f, 1.0f - cos_a * cos_a));
    Float phi = u1 * 2.0f * PI;
    V3 nw = (sin_a * std::cos(phi)) * (-wcx) + (sin_a * std::sin(phi)) * (-wcy) + cos_a * (-wc);
    *p = c + sp.radius * nw;
    *perr = gamma(5.0f) * vabs(*p);
    *n = sp.reverse_orientation ? nw * -1.0f : nw;
    *pdf = uniform_cone_pdf(cos_max);
}

// diffuse.rs:134-167 with shape.rs:17-34 over triangle.rs:522-559 or sphere.rs:328-354
// (CORRECTED spheres seen from outside: the cone sampling above)
LightSample light_sample_point(const ShadeScene& sc, const AreaLight& l, const Isect& it, Float u0, Float u1,
                               Mode mode) {
    LightSample ls;
    V3 p, perr, n;
    Float pdf;
    bool solid_angle = false;  // pdf already per solid angle
    if (sc.geom->is_sphere(l.prim)) {
        const Sphere& sp = sc.geom->spheres[l.prim - sc.geom->tris.size()];
        V3 c;
        if (mode == MODE_CORRECTED && !sphere_contains_origin(sp, it, &c)) {
            sphere_sample_cone(sp, c, it.p, u0, u1, &p, &perr, &n, &pdf);
            solid_angle = true;
        } else {
            sphere_sample_area(sp, u0, u1, mode, &p, &perr, &n);
            pdf = 1.0f / sphere_area(sp);
        }
    } else {
        const TriVerts& tv = sc.geom->tris[l.prim];
        tri_sample(tv, sc.geom->tri_flip[l.prim] != 0, u0, u1, &p, &perr, &n, &pdf);
    }
    if (!solid_angle) {
        V3 wi = p - it.p;
        if (length_squared(wi) == 0.0f) {
            pdf = 0.0f;
        } else {
            wi = normalize(wi);
            // Q3: the reference replaces the area pdf by the solid-angle factor
            Float conv = distance_squared(it.p, p) / abs_dot(n, -wi);
            pdf = (mode == MODE_REF) ? conv : pdf * conv;
            if (std::isinf(pdf)) pdf = 0.0f;
        }
    }
    if (pdf == 0.0f || length_squared(p - it.p) == 0.0f) return ls;
    ls.wi = normalize(p - it.p);
    // Q4: the reference evaluates emission toward +wi
    ls.radiance = emission(l, n, mode == MODE_REF ? ls.wi : -ls.wi);
    ls.pdf = pdf;
    ls.lp = p; ls.lerr = perr; ls.ln = n;
    ls.vis = true;
    return ls;
}

// shape.rs:40-57 on the light's triangle; sphere.rs:356-373 on a sphere (the cone pdf from
// outside, whatever wi, as the reference and pbrt-v3 have it)
Float light_point_pdf(const ShadeScene& sc, const AreaLight& l, const Isect& it, const V3& wi, Mode mode) {
    if (sc.geom->is_sphere(l.prim)) {
        const Sphere& sp = sc.geom->spheres[l.prim - sc.geom->tris.size()];
        V3 c;
        if (!sphere_contains_origin(sp, it, &c)) {
            Float sin2 = sp.radius * sp.radius / distance_squared(it.p, c);
            Float cos_max = std::sqrt(fmax_r(0.0f, 1.0f - sin2));
            return uniform_cone_pdf(cos_max);
        }
        Ray r = spawn_ray(it, wi, mode);
        Float t;
        Isect hit;
        if (!sphere_intersect(sp, r, &t, nullptr, &hit)) return 0.0f;
        Float pdf = distance_squared(it.p, hit.p) / (abs_dot(hit.n, -wi) * sphere_area(sp));
        if (std::isinf(pdf)) pdf = 0.0f;
        return pdf;
    }
    Ray r = spawn_ray(it, wi, mode);
    const TriVerts& tv = sc.geom->tris[l.prim];
    Float t;
    Isect hit;
    if (!tri_intersect(tv, r, sc.geom->tri_flip[l.prim] != 0, &t, nullptr, &hit)) return 0.0f;
    Float pdf = distance_squared(it.p, hit.p) / (abs_dot(hit.n, -wi) * tri_area(tv));
    if (std::isinf(pdf)) pdf = 0.0f;
    return pdf;
}

inline bool scene_intersect(const ShadeScene& sc, Ray* r, Isect* it, uint32_t* prim) {
    Hit h;
    bool ok = bvh_intersect(*sc.geom, *sc.bvh, r, &h, it, nullptr);
    *prim = h.prim;
    return ok;
}
inline bool scene_occluded(const ShadeScene& sc, const Ray& r) { return bvh_intersect_test(*sc.geom, *sc.bvh, r, nullptr, nullptr); }

// integrator.rs:330-437
RGB estimate_direct(const ShadeScene& sc, const Isect& it, const BSDF& bsdf, const AreaLight& light, uint32_t light_idx,
                    Float us0, Float us1, Float ul0, Float ul1, Mode mode) {
    const uint32_t flags = BSDF_ALL & ~BSDF_SPECULAR;
    RGB out;
    LightSample ls = light_sample_point(sc, light, it, ul0, ul1, mode);
    Float scattering_pdf = 0.0f;
    if (ls.pdf > 0.0f && !is_black(ls.radiance)) {
        RGB f = bsdf.f(it.wo, ls.wi, flags) * abs_dot(ls.wi, bsdf.ns);
        scattering_pdf = bsdf.pdf(it.wo, ls.wi, flags);
        if (!is_black(f)) {
            Ray r = spawn_ray_to(it.p, it.perr, it.n, it.time, ls.lp, ls.lerr, ls.ln);
            if (scene_occluded(sc, r)) ls.radiance = RGB();
            if (!is_black(ls.radiance)) {
                Float w = power_heuristic(1.0f, ls.pdf, 1.0f, scattering_pdf);
                out = out + f * ls.radiance * w / ls.pdf;
            }
        }
    }
    // BSDF sampling branch
    BxSample bs = bsdf.sample(it.wo, us0, us1, flags);
    RGB f = bs.f * abs_dot(bs.wi, bsdf.ns);
    bool sampled_specular = (bs.type & BSDF_SPECULAR) != 0;
    if (mode == MODE_REF) {
        // Q5: traces the light-sample direction and divides by its BSDF pdf
        if (!is_black(f) && scattering_pdf > 0.0f) {
            Float w = 1.0f;
            if (!sampled_specular) {
                Float lp = light_point_pdf(sc, light, it, ls.wi, mode);
                if (lp == 0.0f) return out;
                w = power_heuristic(1.0f, scattering_pdf, 1.0f, lp);
            }
            Ray r = spawn_ray(it, ls.wi, mode);
            Isect lit;
            uint32_t prim;
            RGB L;
            if (scene_intersect(sc, &r, &lit, &prim)) {
                int32_t li = sc.prim_light[prim];
                if (li >= 0) L = emission(sc.lights[li], lit.n, -ls.wi);
            }
            if (!is_black(L)) out = out + f * L * w / scattering_pdf;
        }
    } else {
        Float spdf = bs.pdf;
        if (!is_black(f) && spdf > 0.0f) {
            Float w = 1.0f;
            if (!sampled_specular) {
                Float lp = light_point_pdf(sc, light, it, bs.wi, mode);
                if (lp == 0.0f) return out;
                w = power_heuristic(1.0f, spdf, 1.0f, lp);
            }
            Ray r = spawn_ray(it, bs.wi, mode);
            Isect lit;
            uint32_t prim;
            RGB L;
            if (scene_intersect(sc, &r, &lit, &prim)) {
                int32_t li = sc.prim_light[prim];
                if (li == (int32_t)light_idx) L = emission(sc.lights[li], lit.n, -bs.wi);
            }
            if (!is_black(L)) out = out + f * L * w / spdf;
        }
    }
    return out;
}

}  // namespace

// path.rs:51-146
RGB radiance(const ShadeScene& sc, const Ray& camera_ray, SampleStream& st, const RenderParams& p) {
    const Mode mode = p.mode;
    RGB L, beta(1.0f);
    Ray ray = camera_ray;
    bool specular = false;
    uint32_t bounces = 0;
    for (;;) {
        Isect it;
        uint32_t prim = 0xFFFFFFFFu;
        bool hit = scene_intersect(sc, &ray, &it, &prim);
        if (bounces == 0 || specular) {
            if (hit) {
                int32_t li = sc.prim_light[prim];
                if (li >= 0) L = L + beta * emission(sc.lights[li], it.n, -ray.d);
            }
        }
        if (!hit || bounces >= p.max_depth) break;
        const Material& m = sc.materials[sc.prim_material[prim]];
        BSDF bsdf = make_bsdf(it, m, mode);
        if (bsdf.num_components(BSDF_ALL & ~BSDF_SPECULAR) > 0 && !sc.lights.empty()) {
            // integrator.rs:309-328
            Float n = (Float)sc.lights.size();
            Float u = get_1d(st);
            uint32_t idx = (uint32_t)fmin_r(u * n, n - 1.0f);
            Float light_pdf = 1.0f / n;
            Float ul0 = get_1d(st), ul1 = get_1d(st);
            Float us0 = get_1d(st), us1 = get_1d(st);
            RGB ld = estimate_direct(sc, it, bsdf, sc.lights[idx], idx, us0, us1, ul0, ul1, mode) / light_pdf;
            L = L + beta * ld;
        } else if (bsdf.num_components(BSDF_ALL & ~BSDF_SPECULAR) > 0) {
            // no lights: uniform_sample_one_light returns before drawing (integrator.rs:314-316)
        }
        V3 wo = -ray.d;
        Float b0 = get_1d(st), b1 = get_1d(st);
        BxSample bs = bsdf.sample(wo, b0, b1, BSDF_ALL);
        if (is_black(bs.f) || bs.pdf == 0.0f) break;
        beta = beta * (bs.f * abs_dot(bs.wi, bsdf.ns) / bs.pdf);
        specular = (bs.type & BSDF_SPECULAR) != 0;
        RGB rr_beta = beta * 1.0f;
        if (max_component_value(rr_beta) < p.rr_threshold && bounces > 3) {
            Float q = fmax_r(1.0f - max_component_value(rr_beta), 0.5f);
            if (get_1d(st) < q) break;
            beta = beta / (1.0f - q);
        }
        bounces += 1;
        if (mode == MODE_CORRECTED) ray = spawn_ray(it, bs.wi, mode);  // Q1: the reference keeps the camera ray
    }
    return L;
}

// direct.rs:72-113. Every primitive has a material, so the BSDF always exists (the
// reference's bsdf.is_none() continuation never runs); area lights return no
// radiance for escaping rays (light.rs:32-34); the supported materials have no
// specular lobes, so specular_reflect / specular_transmit (integrator.rs:143-271)
// add nothing and draw only samples no later step uses.
RGB radiance_direct(const ShadeScene& sc, const Ray& camera_ray, SampleStream& st, const RenderParams& p) {
    const Mode mode = p.mode;
    Ray ray = camera_ray;
    Isect it;
    uint32_t prim = 0xFFFFFFFFu;
    if (!scene_intersect(sc, &ray, &it, &prim)) return RGB();
    const Material& m = sc.materials[sc.prim_material[prim]];
    BSDF bsdf = make_bsdf(it, m, mode);
    RGB L;
    // it.emitted_radiance(&it.direction)
    int32_t hit_light = sc.prim_light[prim];
    if (hit_light >= 0) L = L + emission(sc.lights[hit_light], it.n, it.wo);
    if (!sc.lights.empty()) {
        if (p.strategy == SAMPLE_ALL_LIGHTS) {
            // integrator.rs:274-306 with Light::sample_count() = 1 (light.rs:52-54)
            RGB all;
            for (uint32_t i = 0; i < (uint32_t)sc.lights.size(); ++i) {
                Float ul0 = get_1d(st), ul1 = get_1d(st);
                Float us0 = get_1d(st), us1 = get_1d(st);
                RGB direct = RGB() + estimate_direct(sc, it, bsdf, sc.lights[i], i, us0, us1, ul0, ul1, mode);
                all = all + direct / 1.0f;
            }
            L = L + all;
        } else {
            // integrator.rs:309-328
            Float n = (Float)sc.lights.size();
            Float u = get_1d(st);
            uint32_t idx = (uint32_t)fmin_r(u * n, n - 1.0f);
            Float light_pdf = 1.0f / n;
            Float ul0 = get_1d(st), ul1 = get_1d(st);
            Float us0 = get_1d(st), us1 = get_1d(st);
            L = L + estimate_direct(sc, it, bsdf, sc.lights[idx], idx, us0, us1, ul0, ul1, mode) / light_pdf;
        }
    }
    return L;
}

// whitted.rs:41-93, with the same remarks as radiance_direct.
RGB radiance_whitted(const ShadeScene& sc, const Ray& camera_ray, SampleStream& st, const RenderParams& p) {
    const Mode mode = p.mode;
    Ray ray = camera_ray;
    Isect it;
    uint32_t prim = 0xFFFFFFFFu;
    if (!scene_intersect(sc, &ray, &it, &prim)) return RGB();
    const Material& m = sc.materials[sc.prim_material[prim]];
    BSDF bsdf = make_bsdf(it, m, mode);
    RGB L;
    int32_t hit_light = sc.prim_light[prim];
    if (hit_light >= 0) L = L + emission(sc.lights[hit_light], it.n, it.wo);
    for (uint32_t i = 0; i < (uint32_t)sc.lights.size(); ++i) {
        Float u0 = get_1d(st), u1 = get_1d(st);
        LightSample ls = light_sample_point(sc, sc.lights[i], it, u0, u1, mode);
        if (is_black(ls.radiance) || ls.pdf == 0.0f) continue;
        RGB f = bsdf.f(it.wo, ls.wi, BSDF_ALL);
        if (ls.vis && !is_black(f)) {
            Ray r = spawn_ray_to(it.p, it.perr, it.n, it.time, ls.lp, ls.lerr, ls.ln);
            if (!scene_occluded(sc, r)) L = L + f * ls.radiance * abs_dot(ls.wi, bsdf.ns) / ls.pdf;
        }
    }
    return L;
}

void render(const ShadeScene& sc, const PerspCamera& cam, const RenderParams& p, int threads, double* rgb_sum,
            double* weight_sum) {
    const int TILE = 16;
    const int W = (int)p.width, H = (int)p.height;
    const int ntx = (W + TILE - 1) / TILE, nty = (H + TILE - 1) / TILE;
    const int wx0 = (int)p.x0, wy0 = (int)p.y0, wx1 = (int)p.x1, wy1 = (int)p.y1;
    const int ww = wx1 - wx0;
    // one accumulation band per tile row: rows [ty*16 - 1, ty*16 + 16)
    std::vector<std::vector<double>> bands(nty);
    std::atomic<int> next(0);
    auto work = [&]() {
        for (;;) {
            int ty = next.fetch_add(1);
            if (ty >= nty) break;
            int ty0 = ty * TILE, ty1 = std::min(ty0 + TILE, H);
            if (ty1 <= wy0 || ty0 > wy1) continue;
            int band0 = ty0 - 1;
            std::vector<double>& band = bands[ty];
            band.assign((size_t)(TILE + 1) * ww * 4, 0.0);
            for (int tx = 0; tx < ntx; ++tx) {
                int tx0 = tx * TILE, tx1 = std::min(tx0 + TILE, W);
                // film.rs:121-136: tile pixel bounds [min - 1, max) (no +1 on max), clipped to the film
                int bx0 = std::max(tx0 - 1, 0), by0 = std::max(ty0 - 1, 0), bx1 = tx1, by1 = ty1;
                for (int py = ty0; py < ty1; ++py) {
                    for (int px = tx0; px < tx1; ++px) {
                        if (px < wx0 || px >= wx1 || py < wy0 || py >= wy1) continue;
                        for (uint32_t s = 0; s < p.spp; ++s) {
                            SampleStream st = make_stream(p.seed, (uint64_t)py * W + px, s);
                            Float fx = (Float)px + get_1d(st), fy = (Float)py + get_1d(st);
                            Float lu = get_1d(st), lv = get_1d(st);
                            Float tu = get_1d(st);
                            Ray r = camera_generate_ray(cam, fx, fy, lu, lv, tu);
                            RGB L = p.integrator == INTEGRATOR_DIRECT    ? radiance_direct(sc, r, st, p)
                                    : p.integrator == INTEGRATOR_WHITTED ? radiance_whitted(sc, r, st, p)
                                                                         : radiance(sc, r, st, p);
                            // integrator.rs:92-117
                            Float y = luminance(L);
                            if (L[0] != L[0] || L[1] != L[1] || L[2] != L[2]) L = RGB();
                            else if (y < -1e-5f) L = RGB();
                            else if (std::isinf(y)) L = RGB();
                            // film.rs:232-284 (box filter, radius 0.5)
                            Float fpx = fx - 0.5f, fpy = fy - 0.5f;
                            int p0x = std::max((int)std::ceil(fpx - 0.5f), bx0);
                            int p0y = std::max((int)std::ceil(fpy - 0.5f), by0);
                            int p1x = std::min((int)std::floor(fpx + 0.5f) + 1, bx1);
                            int p1y = std::min((int)std::floor(fpy + 0.5f) + 1, by1);
                            for (int yy = p0y; yy < p1y; ++yy)
                                for (int xx = p0x; xx < p1x; ++xx) {
                                    if (xx < wx0 || xx >= wx1) continue;
                                    double* q = &band[((size_t)(yy - band0) * ww + (xx - wx0)) * 4];
                                    q[0] += (double)L[0]; q[1] += (double)L[1]; q[2] += (double)L[2];
                                    q[3] += 1.0;
                                }
                        }
                    }
                }
            }
        }
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < std::max(1, threads); ++t) pool.emplace_back(work);
    for (auto& t : pool) t.join();
    const int wh = wy1 - wy0;
    for (size_t i = 0; i < (size_t)ww * wh; ++i) { rgb_sum[3 * i] = rgb_sum[3 * i + 1] = rgb_sum[3 * i + 2] = 0; weight_sum[i] = 0; }
    for (int ty = 0; ty < nty; ++ty) {
        if (bands[ty].empty()) continue;
        int band0 = ty * TILE - 1;
        for (int r = 0; r < TILE + 1; ++r) {
            int yy = band0 + r;
            if (yy < wy0 || yy >= wy1) continue;
            for (int xx = 0; xx < ww; ++xx) {
                const double* q = &bands[ty][((size_t)r * ww + xx) * 4];
                size_t o = (size_t)(yy - wy0) * ww + xx;
                rgb_sum[3 * o] += q[0]; rgb_sum[3 * o + 1] += q[1]; rgb_sum[3 * o + 2] += q[2];
                weight_sum[o] += q[3];
            }
        }
    }
}

BsdfProbe bsdf_probe(const Material& m, Mode mode, const V3& wo, const V3& wi, Float u0, Float u1) {
    Isect it{};
    it.n = V3(0.0f, 0.0f, 1.0f);
    it.ns = it.n;
    it.dpdu_s = V3(1.0f, 0.0f, 0.0f);
    BSDF b = make_bsdf(it, m, mode);
    BsdfProbe r;
    r.f = b.f(wo, wi, BSDF_ALL);
    r.pdf = b.pdf(wo, wi, BSDF_ALL);
    BxSample s = b.sample(wo, u0, u1, BSDF_ALL);
    r.swi = s.wi;
    r.sf = s.f;
    r.spdf = s.pdf;
    r.stype = s.type;
    r.ncomp = b.num_components(BSDF_ALL);
    return r;
}

RGB probe_fresnel_conductor(Float cos_theta_i, const RGB& eta, const RGB& k) {
    return fresnel_conductor(cos_theta_i, RGB(1.0f), eta, k);
}

void probe_tr(Mode mode, Float alpha, const V3& wo, const V3& wh, Float out[4]) {
    TR d{fmax_r(alpha, 0.001f), fmax_r(alpha, 0.001f), mode};
    out[0] = tr_d(d, wh);
    out[1] = tr_lambda(d, wo);
    out[2] = tr_g1(d, wo);
    out[3] = tr_pdf(d, wo, wh);
}

void probe_sample11(Mode mode, Float cos_theta, Float u1, Float u2, Float out[2]) {
    sample11(mode, cos_theta, u1, u2, &out[0], &out[1]);
}

V3 probe_tr_sample(Mode mode, Float alpha, const V3& wo, Float u1, Float u2) {
    TR d{fmax_r(alpha, 0.001f), fmax_r(alpha, 0.001f), mode};
    return tr_sample(d, wo, u1, u2);
}

Ray probe_spawn_ray(const V3& p, const V3& perr, const V3& n, Float time, const V3& d, Mode mode) {
    Isect it{};
    it.p = p;
    it.perr = perr;
    it.n = n;
    it.time = time;
    return spawn_ray(it, d, mode);
}

Ray probe_spawn_ray_to(const V3& p0, const V3& e0, const V3& n0, Float time, const V3& p1, const V3& e1,
                       const V3& n1, V3* target) {
    return spawn_ray_to(p0, e0, n0, time, p1, e1, n1, target);
}

void probe_light_sample(const ShadeScene& sc, uint32_t light, Mode mode, const V3& p, const V3& perr, const V3& n,
                        Float u0, Float u1, Float out[10]) {
    Isect it{};
    it.p = p;
    it.perr = perr;
    it.n = n;
    it.ns = n;
    LightSample ls = light_sample_point(sc, sc.lights[light], it, u0, u1, mode);
    out[0] = ls.lp.x; out[1] = ls.lp.y; out[2] = ls.lp.z;
    out[3] = ls.ln.x; out[4] = ls.ln.y; out[5] = ls.ln.z;
    out[6] = ls.pdf;
    out[7] = ls.radiance[0]; out[8] = ls.radiance[1]; out[9] = ls.radiance[2];
}

Float probe_light_pdf(const ShadeScene& sc, uint32_t light, Mode mode, const V3& p, const V3& perr, const V3& n,
                      const V3& wi) {
    Isect it{};
    it.p = p;
    it.perr = perr;
    it.n = n;
    it.ns = n;
    return light_point_pdf(sc, sc.lights[light], it, wi, mode);
}

}  // namespace orc
