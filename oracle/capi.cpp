// Oracle (test infrastructure only): C entry points used by tests/ (ctypes)
// and by bench.py's CPU-baseline leg. Not linked into the shipped library.
#include <atomic>
#include <cstring>
#include <thread>
#include <vector>

#include "bvh.h"
#include "camera.h"
#include "integrator.h"

using namespace orc;

namespace {

struct OrcScene {
    Scene scene;
    RefBVH bvh;
    ShadeScene shade;
};

// Packed node layout shared with the GPU library's export (32 bytes):
// bmin[3], bmax[3], offset, meta (count << 3 | has_sphere << 2 | axis).
struct PackedNode {
    float bmin[3], bmax[3];
    uint32_t offset, meta;
};
static_assert(sizeof(PackedNode) == 32, "packed node is 32 bytes");

template <class F>
void parallel_for(size_t n, int threads, F fn) {
    if (threads <= 1 || n < 1024) {
        for (size_t i = 0; i < n; ++i) fn(i);
        return;
    }
    std::atomic<size_t> next(0);
    const size_t chunk = 256;
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t) {
        pool.emplace_back([&] {
            for (;;) {
                size_t b = next.fetch_add(chunk);
                if (b >= n) break;
                size_t e = b + chunk < n ? b + chunk : n;
                for (size_t i = b; i < e; ++i) fn(i);
            }
        });
    }
    for (auto& th : pool) th.join();
}

inline Ray load_ray(const float* r) {
    return Ray{V3(r[0], r[1], r[2]), V3(r[4], r[5], r[6]), r[3], r[7]};
}

}  // namespace

extern "C" {

void* orc_scene_new(const float* tri_pos, uint64_t n_tris, const uint8_t* tri_flip, const float* sph_o2w,
                    const float* sph_w2o, const float* sph_params, const uint8_t* sph_rev, uint64_t n_sph) {
    OrcScene* s = new OrcScene();
    s->scene.tris.resize(n_tris);
    s->scene.tri_flip.assign(n_tris, 0);
    for (uint64_t i = 0; i < n_tris; ++i) {
        const float* v = tri_pos + 9 * i;
        s->scene.tris[i] = TriVerts{V3(v[0], v[1], v[2]), V3(v[3], v[4], v[5]), V3(v[6], v[7], v[8])};
        if (tri_flip) s->scene.tri_flip[i] = tri_flip[i];
    }
    for (uint64_t i = 0; i < n_sph; ++i) {
        Transform o2w{mat4_from(sph_o2w + 16 * i), mat4_from(sph_w2o + 16 * i)};
        const float* p = sph_params + 4 * i;
        s->scene.spheres.push_back(make_sphere(o2w, sph_rev ? sph_rev[i] != 0 : false, p[0], p[1], p[2], p[3]));
    }
    return s;
}

void orc_scene_free(void* h) { delete static_cast<OrcScene*>(h); }

// Builds the reference BVH; returns the node count.
uint64_t orc_bvh_build(void* h) {
    OrcScene* s = static_cast<OrcScene*>(h);
    s->bvh = build_bvh(s->scene);
    return s->bvh.nodes.size();
}

uint32_t orc_bvh_height(void* h) {
    OrcScene* s = static_cast<OrcScene*>(h);
    // height = max number of interior ancestors of any node
    std::vector<uint32_t> depth(s->bvh.nodes.size(), 0);
    uint32_t best = 0;
    for (size_t i = 0; i < s->bvh.nodes.size(); ++i) {
        const RefNode& n = s->bvh.nodes[i];
        if (depth[i] > best) best = depth[i];
        if (n.count == 0) {
            depth[i + 1] = depth[i] + 1;
            depth[n.second_child_offset] = depth[i] + 1;
        }
    }
    return best;
}

void orc_bvh_export(void* h, void* nodes_out, uint32_t* ordered_out) {
    OrcScene* s = static_cast<OrcScene*>(h);
    PackedNode* out = static_cast<PackedNode*>(nodes_out);
    for (size_t i = 0; i < s->bvh.nodes.size(); ++i) {
        const RefNode& n = s->bvh.nodes[i];
        PackedNode p;
        p.bmin[0] = n.bounds.mn.x; p.bmin[1] = n.bounds.mn.y; p.bmin[2] = n.bounds.mn.z;
        p.bmax[0] = n.bounds.mx.x; p.bmax[1] = n.bounds.mx.y; p.bmax[2] = n.bounds.mx.z;
        if (n.count > 0) {
            bool sph = false;
            for (uint32_t k = 0; k < n.count; ++k) sph = sph || s->scene.is_sphere(s->bvh.ordered[n.primitive_offset + k]);
            p.offset = n.primitive_offset;
            p.meta = (n.count << 3) | (sph ? 4u : 0u);
        } else {
            p.offset = n.second_child_offset;
            p.meta = n.axis;
        }
        out[i] = p;
    }
    if (ordered_out) std::memcpy(ordered_out, s->bvh.ordered.data(), s->bvh.ordered.size() * 4);
}

// Replaces the oracle's BVH with an externally supplied flattened one.
void orc_bvh_import(void* h, const void* nodes_in, uint64_t n_nodes, const uint32_t* ordered, uint64_t n_ordered) {
    OrcScene* s = static_cast<OrcScene*>(h);
    const PackedNode* in = static_cast<const PackedNode*>(nodes_in);
    s->bvh.nodes.resize(n_nodes);
    for (uint64_t i = 0; i < n_nodes; ++i) {
        RefNode& n = s->bvh.nodes[i];
        n.bounds.mn = V3(in[i].bmin[0], in[i].bmin[1], in[i].bmin[2]);
        n.bounds.mx = V3(in[i].bmax[0], in[i].bmax[1], in[i].bmax[2]);
        n.count = in[i].meta >> 3;
        if (n.count > 0) {
            n.primitive_offset = in[i].offset;
            n.second_child_offset = 0;
            n.axis = 0;
        } else {
            n.second_child_offset = in[i].offset;
            n.primitive_offset = 0;
            n.axis = in[i].meta & 3u;
        }
    }
    s->bvh.ordered.assign(ordered, ordered + n_ordered);
}

// rays: 8 floats each (o.xyz, t_max, d.xyz, time). hits: 4 x 32-bit each
// (prim, t, b1, b2). counters (optional): 3 x u32 per ray (nodes, tris, spheres).
void orc_intersect(void* h, const float* rays, uint64_t n, uint32_t any, uint32_t* hits, uint32_t* counters,
                   int threads) {
    OrcScene* s = static_cast<OrcScene*>(h);
    parallel_for(n, threads, [&](size_t i) {
        Ray r = load_ray(rays + 8 * i);
        Hit hit;
        Counters c;
        if (any) bvh_intersect_test(s->scene, s->bvh, r, &hit, counters ? &c : nullptr);
        else bvh_intersect(s->scene, s->bvh, &r, &hit, nullptr, counters ? &c : nullptr);
        hits[4 * i + 0] = hit.prim;
        std::memcpy(&hits[4 * i + 1], &hit.t, 4);
        std::memcpy(&hits[4 * i + 2], &hit.b1, 4);
        std::memcpy(&hits[4 * i + 3], &hit.b2, 4);
        if (counters) {
            counters[3 * i + 0] = c.nodes;
            counters[3 * i + 1] = c.tris;
            counters[3 * i + 2] = c.spheres;
        }
    });
}

// Brute force over every primitive in the given order with the closest-hit
// acceptance rule (a ground truth for small scenes, independent of the BVH).
void orc_intersect_brute(void* h, const float* rays, uint64_t n, const uint32_t* order, uint32_t* hits) {
    OrcScene* s = static_cast<OrcScene*>(h);
    const size_t np = s->scene.n_prims();
    for (uint64_t i = 0; i < n; ++i) {
        Ray r = load_ray(rays + 8 * i);
        Hit hit;
        for (size_t k = 0; k < np; ++k) {
            uint32_t id = order[k];
            Float t, b[3];
            bool ok;
            if (id < s->scene.tris.size()) {
                ok = tri_intersect(s->scene.tris[id], r, false, &t, b, nullptr);
                if (ok) { hit.b1 = b[1]; hit.b2 = b[2]; }
            } else {
                ok = sphere_intersect(s->scene.spheres[id - s->scene.tris.size()], r, &t, b, nullptr);
                if (ok) { hit.b1 = b[0]; hit.b2 = b[1]; }
            }
            if (ok) { r.t_max = t; hit.prim = id; hit.t = t; }
        }
        hits[4 * i + 0] = hit.prim;
        std::memcpy(&hits[4 * i + 1], &hit.t, 4);
        std::memcpy(&hits[4 * i + 2], &hit.b1, 4);
        std::memcpy(&hits[4 * i + 3], &hit.b2, 4);
    }
}

// Single-primitive tests, for closed-form unit checks. kind 0 = closest
// (intersect), 1 = test. Returns 1 on hit and fills out[0..3] = t, b0|u, b1|v, b2.
int orc_prim_intersect(void* h, uint32_t id, const float* ray, uint32_t kind, float* out) {
    OrcScene* s = static_cast<OrcScene*>(h);
    Ray r = load_ray(ray);
    Float t = 0, b[3] = {0, 0, 0};
    bool ok;
    if (id < s->scene.tris.size()) {
        ok = kind ? tri_intersect_test(s->scene.tris[id], r, &t) : tri_intersect(s->scene.tris[id], r, false, &t, b, nullptr);
    } else {
        const Sphere& sp = s->scene.spheres[id - s->scene.tris.size()];
        ok = kind ? sphere_intersect_test(sp, r, &t) : sphere_intersect(sp, r, &t, b, nullptr);
    }
    out[0] = t; out[1] = b[0]; out[2] = b[1]; out[3] = b[2];
    return ok ? 1 : 0;
}

// The interaction GeometricPrimitive::intersect (geometric.rs:27-38) fills for
// primitive id on a ray (triangle.rs:135-388 / sphere.rs through Isect).
// Returns 1 on hit and fills out[0..17] = t, p, p_error, n, ns, shading dpdu, u, v.
int orc_surface(void* h, uint32_t id, const float* ray, float* out) {
    OrcScene* s = static_cast<OrcScene*>(h);
    Ray r = load_ray(ray);
    Float t = 0, b[3] = {0, 0, 0};
    Isect it;
    bool ok;
    if (id < s->scene.tris.size()) {
        ok = tri_intersect(s->scene.tris[id], r, s->scene.tri_flip[id] != 0, &t, b, &it);
    } else {
        ok = sphere_intersect(s->scene.spheres[id - s->scene.tris.size()], r, &t, b, &it);
    }
    if (!ok) return 0;
    const Float vals[18] = {t, it.p.x, it.p.y, it.p.z, it.perr.x, it.perr.y, it.perr.z, it.n.x, it.n.y, it.n.z,
                            it.ns.x, it.ns.y, it.ns.z, it.dpdu_s.x, it.dpdu_s.y, it.dpdu_s.z, it.u, it.v};
    for (int k = 0; k < 18; ++k) out[k] = vals[k];
    return 1;
}

// Camera rays for a pinhole/thin-lens camera (perspective.rs:49-175); see camera.h.
void orc_camera_rays(const float* c2w, const float* c2w_inv, float fov, float near_clip, float far_clip,
                     uint32_t width, uint32_t height, float fx, float fy, float* rays_out) {
    PerspCamera cam = make_camera(mat4_from(c2w), mat4_from(c2w_inv), fov, near_clip, far_clip, 0.0f, 1e6f, 0.0f,
                                  0.0f, (Float)width, (Float)height);
    for (uint32_t y = 0; y < height; ++y)
        for (uint32_t x = 0; x < width; ++x) {
            Ray r = camera_generate_ray(cam, (Float)x + fx, (Float)y + fy, 0.5f, 0.5f, 0.0f);
            float* o = rays_out + 8 * ((uint64_t)y * width + x);
            o[0] = r.o.x; o[1] = r.o.y; o[2] = r.o.z; o[3] = r.t_max;
            o[4] = r.d.x; o[5] = r.d.y; o[6] = r.d.z; o[7] = r.time;
        }
}

// Materials: mat_type[n], mat_params[n * 10] = (r[3], alpha, eta[3], k[3]); prim_material[n_prims];
// lights: light_prim[nl], light_L[nl * 3], light_double_sided[nl].
void orc_scene_set_shading(void* h, uint32_t n_mat, const uint32_t* mat_type, const float* mat_params,
                           const uint32_t* prim_material, uint32_t n_lights, const uint32_t* light_prim,
                           const float* light_L, const uint8_t* light_ds) {
    OrcScene* s = static_cast<OrcScene*>(h);
    s->shade.materials.resize(n_mat);
    for (uint32_t i = 0; i < n_mat; ++i) {
        Material& m = s->shade.materials[i];
        const float* q = mat_params + 10 * i;
        m.type = mat_type[i];
        m.r = RGB(q[0], q[1], q[2]);
        m.alpha = q[3];
        m.eta = RGB(q[4], q[5], q[6]);
        m.k = RGB(q[7], q[8], q[9]);
    }
    const size_t np = s->scene.n_prims();
    s->shade.prim_material.assign(prim_material, prim_material + np);
    s->shade.prim_light.assign(np, -1);
    s->shade.lights.clear();
    for (uint32_t i = 0; i < n_lights; ++i) {
        s->shade.lights.push_back(AreaLight{light_prim[i], RGB(light_L[3 * i], light_L[3 * i + 1], light_L[3 * i + 2]),
                                            light_ds[i] != 0});
        s->shade.prim_light[light_prim[i]] = (int32_t)i;
    }
}

// Renders window [x0,x1) x [y0,y1) of a width x height film; writes per-pixel
// double RGB sums (3 per pixel) and filter-weight sums.
void orc_render(void* h, const float* c2w, const float* c2w_inv, float fov, float near_clip, float far_clip,
                float lens_radius, float focal_distance, float shutter_open, float shutter_close, uint32_t width,
                uint32_t height, uint32_t spp, uint32_t max_depth, float rr, uint64_t seed, uint32_t mode,
                uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1, uint32_t integrator, uint32_t strategy, int threads,
                double* rgb_sum, double* weight_sum) {
    OrcScene* s = static_cast<OrcScene*>(h);
    s->shade.geom = &s->scene;
    s->shade.bvh = &s->bvh;
    PerspCamera cam = make_camera(mat4_from(c2w), mat4_from(c2w_inv), fov, near_clip, far_clip, shutter_open,
                                  shutter_close, lens_radius, focal_distance, (Float)width, (Float)height);
    RenderParams p{width, height, spp, max_depth, rr, seed, mode == 0 ? MODE_REF : MODE_CORRECTED, x0, y0, x1, y1,
                   integrator, strategy};
    render(s->shade, cam, p, threads, rgb_sum, weight_sum);
}

// Helpers for closed-form tests of the numerics.
float orc_gamma(float n) { return gamma(n); }
float orc_next_float_up(float v) { return next_float_up(v); }
float orc_next_float_down(float v) { return next_float_down(v); }
// look_at (transform.rs:152-189): writes m and m_inverse (16 floats each).
void orc_look_at(const float* pos, const float* look, const float* up, float* m, float* minv) {
    Transform t = look_at(V3(pos[0], pos[1], pos[2]), V3(look[0], look[1], look[2]), V3(up[0], up[1], up[2]));
    for (int i = 0; i < 16; ++i) { m[i] = t.m.m[i / 4][i % 4]; minv[i] = t.minv.m[i / 4][i % 4]; }
}
// Transform::translate then ::scale products for sphere placement: T(c) * S(r,r,r)
void orc_translate_scale(const float* c, float sx, float sy, float sz, float* m, float* minv) {
    Transform t = tmul(translate(V3(c[0], c[1], c[2])), scale(sx, sy, sz));
    for (int i = 0; i < 16; ++i) { m[i] = t.m.m[i / 4][i % 4]; minv[i] = t.minv.m[i / 4][i % 4]; }
}
void orc_mat4_inverse(const float* in, float* out) {
    Mat4 r = inverse(mat4_from(in));
    for (int i = 0; i < 16; ++i) out[i] = r.m[i / 4][i % 4];
}

// One BSDF probe (integrator.h bsdf_probe). params: reflectance[3], alpha, eta[3], k[3].
// out: f[3], pdf, sampled wi[3], sampled f[3], sampled pdf, sampled type (as float), component count.
void orc_bsdf_probe(uint32_t type, const float* params, int mode, const float* wo, const float* wi, float u0, float u1,
                    float* out) {
    Material m;
    m.type = type;
    m.r = RGB(params[0], params[1], params[2]);
    m.alpha = params[3];
    m.eta = RGB(params[4], params[5], params[6]);
    m.k = RGB(params[7], params[8], params[9]);
    BsdfProbe r = bsdf_probe(m, (Mode)mode, V3(wo[0], wo[1], wo[2]), V3(wi[0], wi[1], wi[2]), u0, u1);
    out[0] = r.f[0]; out[1] = r.f[1]; out[2] = r.f[2]; out[3] = r.pdf;
    out[4] = r.swi.x; out[5] = r.swi.y; out[6] = r.swi.z;
    out[7] = r.sf[0]; out[8] = r.sf[1]; out[9] = r.sf[2]; out[10] = r.spdf;
    out[11] = (float)r.stype; out[12] = (float)r.ncomp;
}

// Light sampling of light `light` of a scene with shading set, from n points (p, perr, n: float
// triples): out 10 floats per point (probe_light_sample); light_pdf: one pdf per (point, wi).
void orc_light_sample(void* h, uint32_t light, int mode, uint64_t n, const float* p, const float* perr,
                      const float* nrm, const float* u0, const float* u1, float* out) {
    OrcScene* s = static_cast<OrcScene*>(h);
    s->shade.geom = &s->scene;
    s->shade.bvh = &s->bvh;
    for (uint64_t i = 0; i < n; ++i)
        probe_light_sample(s->shade, light, (Mode)mode, V3(p[3 * i], p[3 * i + 1], p[3 * i + 2]),
                           V3(perr[3 * i], perr[3 * i + 1], perr[3 * i + 2]),
                           V3(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]), u0[i], u1[i], out + 10 * i);
}

void orc_light_pdf(void* h, uint32_t light, int mode, uint64_t n, const float* p, const float* perr, const float* nrm,
                   const float* wi, float* out) {
    OrcScene* s = static_cast<OrcScene*>(h);
    s->shade.geom = &s->scene;
    s->shade.bvh = &s->bvh;
    for (uint64_t i = 0; i < n; ++i)
        out[i] = probe_light_pdf(s->shade, light, (Mode)mode, V3(p[3 * i], p[3 * i + 1], p[3 * i + 2]),
                                 V3(perr[3 * i], perr[3 * i + 1], perr[3 * i + 2]),
                                 V3(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]),
                                 V3(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]));
}

// ------------------------------------------------ single-function probes
// n inputs each; arrays are packed float triples.
void orc_fresnel_conductor(uint64_t n, const float* cos_i, const float* eta, const float* k, float* out) {
    for (uint64_t i = 0; i < n; ++i) {
        RGB r = probe_fresnel_conductor(cos_i[i], RGB(eta[3 * i], eta[3 * i + 1], eta[3 * i + 2]),
                                        RGB(k[3 * i], k[3 * i + 1], k[3 * i + 2]));
        out[3 * i] = r[0]; out[3 * i + 1] = r[1]; out[3 * i + 2] = r[2];
    }
}

// out: 4 floats per (wo, wh) pair: D(wh), Lambda(wo), G1(wo), pdf(wo, wh).
void orc_tr_eval(int mode, float alpha, uint64_t n, const float* wo, const float* wh, float* out) {
    for (uint64_t i = 0; i < n; ++i)
        probe_tr((Mode)mode, alpha, V3(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]),
                 V3(wh[3 * i], wh[3 * i + 1], wh[3 * i + 2]), out + 4 * i);
}

void orc_sample11(int mode, uint64_t n, const float* cos_theta, const float* u1, const float* u2, float* out) {
    for (uint64_t i = 0; i < n; ++i) probe_sample11((Mode)mode, cos_theta[i], u1[i], u2[i], out + 2 * i);
}

// One outgoing direction wo, n uniform pairs (u1, u2): n sampled microfacet normals.
void orc_tr_sample(int mode, float alpha, const float* wo, uint64_t n, const float* u1, const float* u2, float* out) {
    const V3 w(wo[0], wo[1], wo[2]);
    for (uint64_t i = 0; i < n; ++i) {
        V3 h = probe_tr_sample((Mode)mode, alpha, w, u1[i], u2[i]);
        out[3 * i] = h.x; out[3 * i + 1] = h.y; out[3 * i + 2] = h.z;
    }
}

// out: 8 floats per ray (o, t_max, d, time).
void orc_spawn_ray(int mode, uint64_t n, const float* p, const float* perr, const float* nrm, const float* d,
                   float* out) {
    for (uint64_t i = 0; i < n; ++i) {
        Ray r = probe_spawn_ray(V3(p[3 * i], p[3 * i + 1], p[3 * i + 2]),
                                V3(perr[3 * i], perr[3 * i + 1], perr[3 * i + 2]),
                                V3(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]), 0.25f,
                                V3(d[3 * i], d[3 * i + 1], d[3 * i + 2]), (Mode)mode);
        float* o = out + 8 * i;
        o[0] = r.o.x; o[1] = r.o.y; o[2] = r.o.z; o[3] = r.t_max;
        o[4] = r.d.x; o[5] = r.d.y; o[6] = r.d.z; o[7] = r.time;
    }
}

// out: 11 floats per ray (o, t_max, d, time, then the offset target point).
void orc_spawn_ray_to(uint64_t n, const float* p0, const float* e0, const float* n0, const float* p1, const float* e1,
                      const float* n1, float* out) {
    auto v = [](const float* a, uint64_t i) { return V3(a[3 * i], a[3 * i + 1], a[3 * i + 2]); };
    for (uint64_t i = 0; i < n; ++i) {
        V3 t;
        Ray r = probe_spawn_ray_to(v(p0, i), v(e0, i), v(n0, i), 0.25f, v(p1, i), v(e1, i), v(n1, i), &t);
        float* o = out + 11 * i;
        o[0] = r.o.x; o[1] = r.o.y; o[2] = r.o.z; o[3] = r.t_max;
        o[4] = r.d.x; o[5] = r.d.y; o[6] = r.d.z; o[7] = r.time;
        o[8] = t.x; o[9] = t.y; o[10] = t.z;
    }
}

}  // extern "C"
