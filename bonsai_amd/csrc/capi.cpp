// C ABI implementation (include/bonsai_gpu.h): context, scene upload and the
// batched query entry points.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <thread>

#include "bg_internal.h"
#include "tri_degenerate.h"

using bg::HostBVH;

namespace {

int fail(bg_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}

#define HIP_TRY(ctx, expr)                                                                      \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return fail((ctx), BG_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

float rmin(float a, float b) { return a != a ? b : (b != b ? a : (a < b ? a : b)); }
float rmax(float a, float b) { return a != a ? b : (b != b ? a : (a > b ? a : b)); }

// Point3::transform (point3.rs:51-69)
void xform_point(const float* m, const float p[3], float out[3]) {
    const float xp = m[0] * p[0] + m[1] * p[1] + m[2] * p[2] + m[3];
    const float yp = m[4] * p[0] + m[5] * p[1] + m[6] * p[2] + m[7];
    const float zp = m[8] * p[0] + m[9] * p[1] + m[10] * p[2] + m[11];
    const float wp = m[12] * p[0] + m[13] * p[1] + m[14] * p[2] + m[15];
    if (wp == 1.0f) { out[0] = xp; out[1] = yp; out[2] = zp; return; }
    const float iw = 1.0f / wp;
    out[0] = xp * iw; out[1] = yp * iw; out[2] = zp * iw;
}

float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

bool affine(const float* m) { return m[12] == 0.0f && m[13] == 0.0f && m[14] == 0.0f && m[15] == 1.0f; }

template <class T>
int upload(bg_ctx* ctx, T** dst, const T* src, size_t count, uint64_t* bytes) {
    *dst = nullptr;
    if (count == 0) return BG_OK;
    HIP_TRY(ctx, hipMalloc((void**)dst, count * sizeof(T)));
    HIP_TRY(ctx, hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice));
    *bytes += count * sizeof(T);
    return BG_OK;
}

int ensure(bg_ctx* ctx, void** buf, size_t* have, size_t need) {
    if (*have >= need) return BG_OK;
    if (*buf) hipFree(*buf);
    *buf = nullptr;
    *have = 0;
    HIP_TRY(ctx, hipMalloc(buf, need));
    *have = need;
    return BG_OK;
}

int traverse_grid(const bg_ctx* ctx, size_t n) {
    const uint64_t tiles = (n + ctx->shape.block - 1) / ctx->shape.block;
    uint64_t cap = ctx->shape.blocks_per_cu > 0 ? (uint64_t)ctx->num_cus * ctx->shape.blocks_per_cu : tiles;
    cap = std::min<uint64_t>(cap, 1u << 20);
    return (int)std::max<uint64_t>(1, std::min(tiles, cap));
}

// Device memory a traversal call works in beyond its rays and hits: the spill
// rows of the LDS stacks, the three BG_COUNT counters and the work-queue
// counter. Null members: the context's buffers, grown on demand.
struct TraverseScratch {
    uint32_t* spill = nullptr;
    size_t spill_bytes = 0;
    unsigned long long* counters = nullptr;
    unsigned long long* queue = nullptr;
};

int run_traverse(bg_ctx* ctx, const bg_scene* s, const void* d_rays, size_t n, void* d_hits, uint32_t flags,
                 bg_stats* stats, const TraverseScratch* fixed = nullptr) {
    if (n == 0) {
        if (stats) *stats = bg_stats{};
        return BG_OK;
    }
    int grid = traverse_grid(ctx, n);
    const uint32_t depth = s->bvh.max_depth + 1;
    const size_t row_bytes = (size_t)depth * bg::kSpillWords * sizeof(uint32_t);
    const size_t block_bytes = (size_t)ctx->shape.block * row_bytes;
    if (fixed) {
        // the grid is narrowed to the spill rows the scratch holds; traversal results do not
        // depend on the grid (each ray is traced the same way by whichever lane takes it)
        const size_t fit = fixed->spill_bytes / block_bytes;
        if (fit == 0)
            return fail(ctx, BG_ERR_INVALID,
                        "bg_intersect_device_noalloc: one block of " + std::to_string(ctx->shape.block) +
                            " lanes needs " + std::to_string(block_bytes) + " bytes of stack spill (BVH depth " +
                            std::to_string(depth) + "), the static scratch holds " +
                            std::to_string(fixed->spill_bytes));
        grid = (int)std::min<size_t>((size_t)grid, fit);
    } else {
        int rc = ensure(ctx, (void**)&ctx->spill, &ctx->spill_bytes, (size_t)grid * block_bytes);
        if (rc) return rc;
    }
    bg::TraverseArgs a;
    a.nodes = s->d_nodes;
    a.pairs = s->d_pairs;
    a.n_top = (uint32_t)std::min<uint64_t>(s->n_pairs, bg::kTopRecords);
    a.leaves = s->d_leaves;
    a.tri_verts = s->d_tri_verts;
    a.slot_prim = s->d_slot_prim;
    a.spheres = s->d_spheres;
    a.sphere_aux = s->d_sphere_aux;
    a.n_tris = (uint32_t)s->n_tris;
    a.rays = (const float4*)d_rays;
    a.hits = (uint4*)d_hits;
    a.n_rays = n;
    a.spill = fixed ? fixed->spill : ctx->spill;
    a.spill_depth = depth;
    a.counters = nullptr;
    a.queue = fixed ? fixed->queue : ctx->queue;
    const bool any = (flags & BG_ANY) != 0;
    const bool count = (flags & BG_COUNT) != 0;
    const unsigned long long zeros[3] = {0, 0, 0};
    if (count && fixed) {
        HIP_TRY(ctx, bg::noalloc_zero_counters(fixed->counters, 3, ctx->stream));
        a.counters = fixed->counters;
    } else if (count) {
        if (!ctx->counters) HIP_TRY(ctx, hipMalloc((void**)&ctx->counters, sizeof(zeros)));
        HIP_TRY(ctx, hipMemcpyAsync(ctx->counters, zeros, sizeof(zeros), hipMemcpyHostToDevice, ctx->stream));
        a.counters = ctx->counters;
    }
    auto t0 = std::chrono::steady_clock::now();
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    HIP_TRY(ctx, bg::launch_traverse(a, any, count, ctx->shape, grid, ctx->stream));
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    HIP_TRY(ctx, hipEventSynchronize(ctx->ev1));
    auto t1 = std::chrono::steady_clock::now();
    float ms = 0.0f;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    unsigned long long cnt[3] = {0, 0, 0};
    if (count) HIP_TRY(ctx, hipMemcpy(cnt, a.counters, sizeof(cnt), hipMemcpyDeviceToHost));
    if (stats) {
        *stats = bg_stats{};
        stats->kernel_ms = ms;
        stats->total_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        stats->rays = n;
        stats->launches = 1;
        stats->nodes_visited = cnt[0];
        stats->tris_tested = cnt[1];
        stats->spheres_tested = cnt[2];
    }
    return BG_OK;
}

// Primitive bounds in scene order plus the device sphere records.
void prepare_prims(const bg_scene_desc* d, std::vector<bg::PrimBounds>* bounds_out, std::vector<uint8_t>* is_sphere_out,
                   std::vector<bg::SphereDev>* sph_out, std::vector<bg::SphereAux>* aux_out) {
    const uint64_t nt = d->n_tris, ns = d->n_spheres, np = nt + ns;
    std::vector<bg::PrimBounds>& bounds = *bounds_out;
    std::vector<uint8_t>& is_sphere = *is_sphere_out;
    bounds.assign(np, bg::PrimBounds{});
    is_sphere.assign(np, 0);
    for (uint64_t i = 0; i < nt; ++i) {
        const float* v = d->tri_positions + 9 * i;
        // Bounds3::new(p0, p1).union_point(p2) (triangle.rs:128-133)
        for (int k = 0; k < 3; ++k) {
            bounds[i].mn[k] = rmin(rmin(v[k], v[3 + k]), v[6 + k]);
            bounds[i].mx[k] = rmax(rmax(v[k], v[3 + k]), v[6 + k]);
        }
    }
    std::vector<bg::SphereDev>& sph = *sph_out;
    std::vector<bg::SphereAux>& aux = *aux_out;
    sph.assign(ns, bg::SphereDev{});
    aux.assign(ns, bg::SphereAux{});
    for (uint64_t i = 0; i < ns; ++i) {
        const float* prm = d->sphere_params + 4 * i;
        const float r = prm[0], za = prm[1], zb = prm[2];
        // Sphere::new (sphere.rs:39-65); world_to_object is the transform's m_inverse.
        const float zmin = clampf(rmin(za, zb), -r, r), zmax = clampf(rmax(za, zb), -r, r);
        const float pi = 3.14159265358979323846f;
        sph[i].radius = r;
        sph[i].z_min = zmin;
        sph[i].z_max = zmax;
        sph[i].phi_max = clampf(prm[3], 0.0f, 360.0f) * (pi / 180.0f);
        aux[i].theta_min = std::acos(clampf(rmin(za, zb) / r, -1.0f, 1.0f));
        aux[i].theta_max = std::acos(clampf(rmax(za, zb) / r, -1.0f, 1.0f));
        const float* o2w = d->sphere_object_to_world + 16 * i;
        bool identity = true;
        for (int k = 0; k < 16; ++k) identity = identity && o2w[k] == ((k % 5 == 0) ? 1.0f : 0.0f);
        const float* w2o = identity ? o2w : d->sphere_world_to_object + 16 * i;
        std::memcpy(sph[i].w2o, w2o, 12 * sizeof(float));
        // Transform::transform_bounds of the object bounds (transform.rs:75-85)
        const float lo[3] = {-r, -r, zmin}, hi[3] = {r, r, zmax};
        const float corners[8][3] = {{lo[0], lo[1], lo[2]}, {hi[0], lo[1], lo[2]}, {lo[0], hi[1], lo[2]},
                                     {lo[0], lo[1], hi[2]}, {lo[0], hi[1], hi[2]}, {hi[0], hi[1], lo[2]},
                                     {hi[0], lo[1], hi[2]}, {hi[0], hi[1], hi[2]}};
        bg::PrimBounds& b = bounds[nt + i];
        float p[3];
        xform_point(o2w, corners[0], p);
        for (int k = 0; k < 3; ++k) { b.mn[k] = p[k]; b.mx[k] = p[k]; }
        for (int c = 1; c < 8; ++c) {
            xform_point(o2w, corners[c], p);
            for (int k = 0; k < 3; ++k) { b.mn[k] = rmin(b.mn[k], p[k]); b.mx[k] = rmax(b.mx[k], p[k]); }
        }
        is_sphere[nt + i] = 1;
    }
}

}  // namespace

extern "C" {

int bg_abi_version(void) { return BG_ABI_VERSION; }

int bg_device_count(int* count) {
    if (!count) return BG_ERR_INVALID;
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *count = c;
    return BG_OK;
}

int bg_create(int device, bg_ctx** out) {
    if (!out) return BG_ERR_INVALID;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return BG_ERR_INVALID;
    bg_ctx* ctx = new (std::nothrow) bg_ctx();
    if (!ctx) return BG_ERR_NOMEM;
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess) {
        delete ctx;
        return BG_ERR_HIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
        delete ctx;
        return BG_ERR_HIP;
    }
    ctx->num_cus = prop.multiProcessorCount;
    // the static scratch of bg_intersect_device_noalloc: its address on this device, taken once
    if (bg::noalloc_scratch_init(device) != hipSuccess || hipMalloc((void**)&ctx->queue, 256) != hipSuccess) {
        if (ctx->queue) hipFree(ctx->queue);
        hipEventDestroy(ctx->ev0);
        hipEventDestroy(ctx->ev1);
        hipStreamDestroy(ctx->stream);
        delete ctx;
        return BG_ERR_HIP;
    }
    *out = ctx;
    return BG_OK;
}

void bg_destroy(bg_ctx* ctx) {
    if (!ctx) return;
    hipSetDevice(ctx->device);
    for (hipEvent_t e : ctx->prof_events) hipEventDestroy(e);
    if (ctx->spill) hipFree(ctx->spill);
    if (ctx->counters) hipFree(ctx->counters);
    if (ctx->queue) hipFree(ctx->queue);
    if (ctx->io) hipFree(ctx->io);
    if (ctx->paths.mem) hipFree(ctx->paths.mem);
    if (ctx->paths.guide_mem) hipFree(ctx->paths.guide_mem);
    if (ctx->direct.mem) hipFree(ctx->direct.mem);
    if (ctx->ev0) hipEventDestroy(ctx->ev0);
    if (ctx->ev1) hipEventDestroy(ctx->ev1);
    for (int k = 0; k < 2; ++k) {
        if (ctx->side[k]) hipStreamDestroy(ctx->side[k]);
        if (ctx->join_ev[k]) hipEventDestroy(ctx->join_ev[k]);
    }
    if (ctx->fork_ev) hipEventDestroy(ctx->fork_ev);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* bg_last_error(const bg_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

void* bg_stream(bg_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int bg_set_launch_config(bg_ctx* ctx, const bg_launch_config* cfg) {
    if (!ctx || !cfg) return BG_ERR_INVALID;
    bg::LaunchShape s;
    s.block = (int)cfg->block;
    s.stack = (int)cfg->stack;
    s.blocks_per_cu = (int)cfg->blocks_per_cu;
    s.variant = (int)cfg->variant;
    s.compact = cfg->compact != 0;
    if (!bg::traverse_shape_supported(s)) return fail(ctx, BG_ERR_INVALID, "unsupported launch config");
    ctx->shape = s;
    ctx->render_shape = s;
    return BG_OK;
}

int bg_set_profiling(bg_ctx* ctx, uint32_t flags) {
    if (!ctx) return BG_ERR_INVALID;
    if (flags & ~(BG_PROFILE_COUNT | BG_PROFILE_TRAVERSAL_TIME)) return fail(ctx, BG_ERR_INVALID, "unknown profiling flags");
    ctx->profile = flags;
    return BG_OK;
}

int bg_get_launch_config(const bg_ctx* ctx, bg_launch_config* cfg) {
    if (!ctx || !cfg) return BG_ERR_INVALID;
    cfg->block = (uint32_t)ctx->shape.block;
    cfg->stack = (uint32_t)ctx->shape.stack;
    cfg->blocks_per_cu = (uint32_t)ctx->shape.blocks_per_cu;
    cfg->variant = (uint32_t)ctx->shape.variant;
    cfg->compact = ctx->shape.compact ? 1u : 0u;
    return BG_OK;
}

int bg_scene_create(bg_ctx* ctx, const bg_scene_desc* d, bg_scene** out) {
    if (!ctx || !d || !out) return fail(ctx, BG_ERR_INVALID, "null argument");
    *out = nullptr;
    if (d->n_tris > 0 && !d->tri_positions) return fail(ctx, BG_ERR_INVALID, "tri_positions is null");
    if (d->n_spheres > 0 && (!d->sphere_object_to_world || !d->sphere_world_to_object || !d->sphere_params))
        return fail(ctx, BG_ERR_INVALID, "sphere arrays are null");
    if (d->n_tris + d->n_spheres >= 0x7FFFFFFFull) return fail(ctx, BG_ERR_INVALID, "too many primitives");
    if (d->n_lights > 0 && !d->lights) return fail(ctx, BG_ERR_INVALID, "lights is null");
    if (d->n_materials > 0 && !d->materials) return fail(ctx, BG_ERR_INVALID, "materials is null");
    const uint64_t nt = d->n_tris, ns = d->n_spheres, np = nt + ns;
    for (uint64_t i = 0; i < ns; ++i) {
        if (!affine(d->sphere_object_to_world + 16 * i) || !affine(d->sphere_world_to_object + 16 * i))
            return fail(ctx, BG_ERR_UNSUPPORTED, "sphere transforms must be affine");
        if (!(d->sphere_params[4 * i] > 0.0f)) return fail(ctx, BG_ERR_INVALID, "sphere radius must be > 0");
    }
    for (uint32_t i = 0; i < d->n_lights; ++i)
        if (d->lights[i].prim_id >= np) return fail(ctx, BG_ERR_INVALID, "area light on a primitive the scene does not have");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    auto t0 = std::chrono::steady_clock::now();

    bg_scene* s = new (std::nothrow) bg_scene();
    if (!s) return fail(ctx, BG_ERR_NOMEM, "out of host memory");
    s->ctx = ctx;
    s->n_tris = nt;
    s->n_spheres = ns;
    s->tri_positions.assign(d->tri_positions, d->tri_positions + 9 * nt);
    s->tri_flip.assign(nt, 0);
    if (d->tri_reverse_orientation)
        for (uint64_t i = 0; i < nt; ++i) s->tri_flip[i] = d->tri_reverse_orientation[i] ? 1 : 0;
    s->prim_material.assign(np, 0);
    if (d->tri_material) std::copy(d->tri_material, d->tri_material + nt, s->prim_material.begin());
    if (d->sphere_material) std::copy(d->sphere_material, d->sphere_material + ns, s->prim_material.begin() + nt);
    if (d->n_materials >= bg::kPrimFlip) {
        delete s;
        return fail(ctx, BG_ERR_INVALID, "too many materials");
    }
    if (d->n_materials) s->materials.assign(d->materials, d->materials + d->n_materials);
    for (uint32_t m : s->prim_material)
        if (!s->materials.empty() && m >= s->materials.size()) {
            delete s;
            return fail(ctx, BG_ERR_INVALID, "material index out of range");
        }
    for (const bg_material& m : s->materials)
        if (m.type > BG_MAT_PLASTIC) {
            delete s;
            return fail(ctx, BG_ERR_INVALID, "unknown material type");
        }
    s->prim_light.assign(np, -1);
    if (d->n_lights) s->lights.assign(d->lights, d->lights + d->n_lights);
    for (uint32_t i = 0; i < d->n_lights; ++i) s->prim_light[d->lights[i].prim_id] = (int32_t)i;
    if (ns) {
        s->sphere_o2w.assign(d->sphere_object_to_world, d->sphere_object_to_world + 16 * ns);
        s->sphere_w2o.assign(d->sphere_world_to_object, d->sphere_world_to_object + 16 * ns);
        s->sphere_params.assign(d->sphere_params, d->sphere_params + 4 * ns);
        s->sphere_flags.assign(ns, 0);
        if (d->sphere_reverse_orientation)
            for (uint64_t i = 0; i < ns; ++i) s->sphere_flags[i] = d->sphere_reverse_orientation[i] ? 1 : 0;
    }

    std::vector<bg::PrimBounds> bounds;
    std::vector<uint8_t> is_sphere;
    std::vector<bg::SphereDev> sph;
    std::vector<bg::SphereAux> aux;
    prepare_prims(d, &bounds, &is_sphere, &sph, &aux);

    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    bg::build_bvh(bounds, is_sphere, (int)std::min(hw, 16u), &s->bvh);

    // Slot-ordered triangle vertices for traversal. A degenerate triangle (rejected by every
    // intersection test, whatever the ray) is stored as its first vertex three times: its edge
    // functions are then exactly zero and the test rejects it at det == 0, so the traversal
    // kernels skip the per-test degeneracy check.
    const size_t nslot = s->bvh.slot_prim.size();
    std::vector<float> verts(9 * nslot, 0.0f);
    for (size_t k = 0; k < nslot; ++k) {
        const uint32_t id = s->bvh.slot_prim[k];
        if (id >= nt) continue;
        const float* v = d->tri_positions + 9 * (uint64_t)id;
        float* o = &verts[9 * k];
        std::memcpy(o, v, 9 * sizeof(float));
        if (bg::tri_degenerate(v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], v[8]))
            for (int c = 3; c < 9; ++c) o[c] = v[c % 3];
    }
    uint64_t bytes = 0;
    int rc = BG_OK;
    if (!s->bvh.nodes.empty())
        rc = upload(ctx, (bg_bvh_node**)&s->d_nodes, s->bvh.nodes.data(), s->bvh.nodes.size(), &bytes);
    if (!s->bvh.nodes.empty()) {
        std::vector<bg::PairRecord> pairs;
        std::vector<uint32_t> leaves;
        bg::build_pairs(s->bvh, &pairs, &leaves);
        if (!rc) rc = upload(ctx, (bg::PairRecord**)&s->d_pairs, pairs.data(), pairs.size(), &bytes);
        s->n_pairs = pairs.size();
        if (!rc) rc = upload(ctx, (uint32_t**)&s->d_leaves, leaves.data(), leaves.size(), &bytes);
    }
    if (!rc) rc = upload(ctx, &s->d_tri_verts, verts.data(), verts.size(), &bytes);
    if (!rc) rc = upload(ctx, &s->d_slot_prim, s->bvh.slot_prim.data(), nslot, &bytes);
    if (!rc) rc = upload(ctx, &s->d_spheres, sph.data(), sph.size(), &bytes);
    if (!rc) rc = upload(ctx, &s->d_sphere_aux, aux.data(), aux.size(), &bytes);
    // shading records
    if (!rc) rc = upload(ctx, &s->d_tri_orig, d->tri_positions, 9 * nt, &bytes);
    {
        std::vector<uint2> info(np);
        for (uint64_t p = 0; p < np; ++p)
            info[p] = bg::prim_info_of(s->prim_light[p], s->prim_material[p], p < nt && s->tri_flip[p] != 0);
        if (!rc) rc = upload(ctx, &s->d_prim_info, info.data(), np, &bytes);
    }
    std::vector<bg::MaterialDev> mats(s->materials.size());
    for (size_t i = 0; i < mats.size(); ++i) {
        const bg_material& m = s->materials[i];
        mats[i] = bg::MaterialDev{m.type, {m.reflectance[0], m.reflectance[1], m.reflectance[2]}, m.alpha,
                                  {m.eta[0], m.eta[1], m.eta[2]}, {m.k[0], m.k[1], m.k[2]}, 0u};
    }
    if (!rc) rc = upload(ctx, &s->d_mats, mats.data(), mats.size(), &bytes);
    std::vector<bg::LightDev> lts(s->lights.size());
    for (size_t i = 0; i < lts.size(); ++i) {
        const bg_area_light& l = s->lights[i];
        if (l.prim_id >= nt) {
            // sphere.rs:375-377: phi_max * radius * (z_max - z_min)
            const bg::SphereDev& sp = sph[l.prim_id - nt];
            lts[i] = bg::LightDev{l.prim_id, {l.radiance[0], l.radiance[1], l.radiance[2]}, l.double_sided ? 1u : 0u,
                                  sp.phi_max * sp.radius * (sp.z_max - sp.z_min), {0u, 0u}};
            continue;
        }
        const float* v = d->tri_positions + 9 * (uint64_t)l.prim_id;
        // triangle.rs:561-566: 0.5 * |(p1 - p0) x (p2 - p0)| with the f64 cross of vec3.rs:108-122
        const double ax = v[3] - v[0], ay = v[4] - v[1], az = v[5] - v[2];
        const double bx = v[6] - v[0], by = v[7] - v[1], bz = v[8] - v[2];
        const float ax_ = (float)ax, ay_ = (float)ay, az_ = (float)az, bx_ = (float)bx, by_ = (float)by, bz_ = (float)bz;
        const float cx = (float)((double)ay_ * bz_ - (double)az_ * by_);
        const float cy = (float)((double)az_ * bx_ - (double)ax_ * bz_);
        const float cz = (float)((double)ax_ * by_ - (double)ay_ * bx_);
        lts[i] = bg::LightDev{l.prim_id, {l.radiance[0], l.radiance[1], l.radiance[2]}, l.double_sided ? 1u : 0u,
                              0.5f * std::sqrt(cx * cx + cy * cy + cz * cz), {0u, 0u}};
    }
    if (!rc) rc = upload(ctx, &s->d_lights, lts.data(), lts.size(), &bytes);
    std::vector<bg::SphereShadeDev> ssh(ns);
    for (uint64_t i = 0; i < ns; ++i) {
        const float* m = d->sphere_object_to_world + 16 * i;
        std::memcpy(ssh[i].o2w, m, 12 * sizeof(float));
        std::memcpy(ssh[i].w2o, sph[i].w2o, 12 * sizeof(float));
        ssh[i].radius = sph[i].radius;
        ssh[i].z_min = sph[i].z_min;
        ssh[i].z_max = sph[i].z_max;
        ssh[i].phi_max = sph[i].phi_max;
        ssh[i].theta_min = aux[i].theta_min;
        ssh[i].theta_max = aux[i].theta_max;
        // transform.rs:218-224
        const float det = m[0] * (m[5] * m[10] - m[6] * m[9]) - m[1] * (m[4] * m[10] - m[6] * m[8]) +
                          m[2] * (m[4] * m[9] - m[5] * m[8]);
        ssh[i].flip = ((s->sphere_flags[i] != 0) != (det < 0.0f)) ? 1u : 0u;
        ssh[i].reverse = s->sphere_flags[i] != 0 ? 1u : 0u;
    }
    if (!rc) rc = upload(ctx, &s->d_sph_shade, ssh.data(), ssh.size(), &bytes);
    if (rc) {
        bg_scene_destroy(s);
        return rc;
    }
    auto t1 = std::chrono::steady_clock::now();
    s->info.n_prims = np;
    s->info.n_nodes = s->bvh.nodes.size();
    s->info.max_depth = s->bvh.max_depth;
    s->info.max_leaf = s->bvh.max_leaf;
    s->info.build_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    s->info.device_bytes = bytes;
    *out = s;
    return BG_OK;
}

int bg_build_bvh(const bg_scene_desc* d, bg_bvh_node* nodes, uint64_t* n_nodes, uint32_t* slot_prim,
                 uint64_t* n_prims, uint32_t* max_depth) {
    if (!d || !n_nodes || !n_prims) return BG_ERR_INVALID;
    if (d->n_tris > 0 && !d->tri_positions) return BG_ERR_INVALID;
    if (d->n_spheres > 0 && (!d->sphere_object_to_world || !d->sphere_world_to_object || !d->sphere_params))
        return BG_ERR_INVALID;
    for (uint64_t i = 0; i < d->n_spheres; ++i)
        if (!affine(d->sphere_object_to_world + 16 * i) || !affine(d->sphere_world_to_object + 16 * i))
            return BG_ERR_UNSUPPORTED;
    std::vector<bg::PrimBounds> bounds;
    std::vector<uint8_t> is_sphere;
    std::vector<bg::SphereDev> sph;
    std::vector<bg::SphereAux> aux;
    prepare_prims(d, &bounds, &is_sphere, &sph, &aux);
    HostBVH bvh;
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    bg::build_bvh(bounds, is_sphere, (int)std::min(hw, 16u), &bvh);
    if (nodes) {
        if (*n_nodes < bvh.nodes.size() || *n_prims < bvh.slot_prim.size() || !slot_prim) return BG_ERR_INVALID;
        std::memcpy(nodes, bvh.nodes.data(), bvh.nodes.size() * sizeof(bg_bvh_node));
        std::memcpy(slot_prim, bvh.slot_prim.data(), bvh.slot_prim.size() * sizeof(uint32_t));
    }
    *n_nodes = bvh.nodes.size();
    *n_prims = bvh.slot_prim.size();
    if (max_depth) *max_depth = bvh.max_depth;
    return BG_OK;
}

void bg_scene_destroy(bg_scene* s) {
    if (!s) return;
    if (s->ctx) hipSetDevice(s->ctx->device);
    if (s->d_nodes) hipFree(s->d_nodes);
    if (s->d_pairs) hipFree(s->d_pairs);
    if (s->d_leaves) hipFree(s->d_leaves);
    if (s->d_tri_verts) hipFree(s->d_tri_verts);
    if (s->d_slot_prim) hipFree(s->d_slot_prim);
    if (s->d_spheres) hipFree(s->d_spheres);
    if (s->d_sphere_aux) hipFree(s->d_sphere_aux);
    void* extra[] = {s->d_tri_orig, s->d_prim_info, s->d_mats, s->d_lights, s->d_sph_shade};
    for (void* p : extra)
        if (p) hipFree(p);
    delete s;
}

int bg_scene_info_get(const bg_scene* s, bg_scene_info* info) {
    if (!s || !info) return BG_ERR_INVALID;
    *info = s->info;
    return BG_OK;
}

int bg_scene_bvh(const bg_scene* s, bg_bvh_node* nodes, uint32_t* slot_prim) {
    if (!s) return BG_ERR_INVALID;
    if (nodes && !s->bvh.nodes.empty())
        std::memcpy(nodes, s->bvh.nodes.data(), s->bvh.nodes.size() * sizeof(bg_bvh_node));
    if (slot_prim && !s->bvh.slot_prim.empty())
        std::memcpy(slot_prim, s->bvh.slot_prim.data(), s->bvh.slot_prim.size() * sizeof(uint32_t));
    return BG_OK;
}

int bg_intersect_device(bg_ctx* ctx, const bg_scene* s, const void* d_rays, size_t n, void* d_hits, uint32_t flags,
                        bg_stats* stats) {
    if (!ctx || !s || s->ctx != ctx) return fail(ctx, BG_ERR_INVALID, "scene does not belong to this context");
    if (n > 0 && (!d_rays || !d_hits)) return fail(ctx, BG_ERR_INVALID, "null ray or hit buffer");
    if (flags & ~3u) return fail(ctx, BG_ERR_INVALID, "unknown flags");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (s->bvh.nodes.empty()) {
        // Empty scene: every query misses (bvh.rs:286, 359).
        std::vector<bg_hit> miss(n, bg_hit{BG_MISS, 0.0f, 0.0f, 0.0f});
        if (n) HIP_TRY(ctx, hipMemcpyAsync(d_hits, miss.data(), n * sizeof(bg_hit), hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        if (stats) *stats = bg_stats{0.0, 0.0, n, 0};
        return BG_OK;
    }
    return run_traverse(ctx, s, d_rays, n, d_hits, flags, stats);
}

int bg_intersect_device_noalloc(bg_ctx* ctx, const bg_scene* s, const void* d_rays, size_t n, void* d_hits,
                                uint32_t flags, bg_stats* stats) {
    if (!ctx || !s || s->ctx != ctx) return fail(ctx, BG_ERR_INVALID, "scene does not belong to this context");
    if (n > 0 && (!d_rays || !d_hits)) return fail(ctx, BG_ERR_INVALID, "null ray or hit buffer");
    if (flags & ~3u) return fail(ctx, BG_ERR_INVALID, "unknown flags");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    char* base = static_cast<char*>(bg::noalloc_scratch(ctx->device));
    if (!base) return fail(ctx, BG_ERR_HIP, "bg_intersect_device_noalloc: static scratch has no address on this device");
    if (s->bvh.nodes.empty()) {
        // Empty scene: every query misses (bvh.rs:286, 359).
        HIP_TRY(ctx, bg::noalloc_fill_miss(d_hits, n, ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        if (stats) *stats = bg_stats{0.0, 0.0, n, 0};
        return BG_OK;
    }
    // layout: counters and the queue counter in the first 256 bytes, spill rows after them
    TraverseScratch sc;
    sc.counters = reinterpret_cast<unsigned long long*>(base);
    sc.queue = reinterpret_cast<unsigned long long*>(base + 64);
    sc.spill = reinterpret_cast<uint32_t*>(base + 256);
    sc.spill_bytes = (size_t)BG_NOALLOC_SCRATCH_BYTES - 256;
    return run_traverse(ctx, s, d_rays, n, d_hits, flags, stats, &sc);
}

int bg_intersect(bg_ctx* ctx, const bg_scene* s, const bg_ray* rays, size_t n, bg_hit* hits, uint32_t flags,
                 bg_stats* stats) {
    if (!ctx || !s) return fail(ctx, BG_ERR_INVALID, "null argument");
    if (n > 0 && (!rays || !hits)) return fail(ctx, BG_ERR_INVALID, "null ray or hit buffer");
    if (n == 0) {
        if (stats) *stats = bg_stats{};
        return BG_OK;
    }
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const size_t rb = n * sizeof(bg_ray), hb = n * sizeof(bg_hit);
    int rc = ensure(ctx, &ctx->io, &ctx->io_bytes, rb + hb);
    if (rc) return rc;
    char* base = (char*)ctx->io;
    HIP_TRY(ctx, hipMemcpyAsync(base, rays, rb, hipMemcpyHostToDevice, ctx->stream));
    rc = bg_intersect_device(ctx, s, base, n, base + rb, flags, stats);
    if (rc) return rc;
    HIP_TRY(ctx, hipMemcpyAsync(hits, base + rb, hb, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return BG_OK;
}

int bg_surface_interactions(bg_ctx* ctx, const bg_scene* s, const bg_ray* rays, const bg_hit* hits, size_t n,
                            bg_surface* out) {
    if (!ctx || !s) return fail(ctx, BG_ERR_INVALID, "null argument");
    if (n > 0 && (!rays || !hits || !out)) return fail(ctx, BG_ERR_INVALID, "null ray, hit or output buffer");
    if (n == 0) return BG_OK;
    for (size_t i = 0; i < n; ++i)
        if (hits[i].prim_id != BG_MISS && hits[i].prim_id >= s->n_tris + s->n_spheres)
            return fail(ctx, BG_ERR_INVALID, "hit names a primitive the scene does not have");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const size_t rb = n * sizeof(bg_ray), hb = n * sizeof(bg_hit), ob = n * sizeof(bg_surface);
    int rc = ensure(ctx, &ctx->io, &ctx->io_bytes, rb + hb + ob);
    if (rc) return rc;
    char* base = (char*)ctx->io;
    HIP_TRY(ctx, hipMemcpyAsync(base, rays, rb, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(base + rb, hits, hb, hipMemcpyHostToDevice, ctx->stream));
    bg::ShadeSceneDev sc;
    sc.tri_orig = s->d_tri_orig;
    sc.prim_info = s->d_prim_info;
    sc.spheres = s->d_sph_shade;
    sc.n_tris = (uint32_t)s->n_tris;
    HIP_TRY(ctx, bg::launch_surface(sc, (const float4*)base, (const uint4*)(base + rb), n, (float4*)(base + rb + hb),
                                    ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(out, base + rb + hb, ob, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return BG_OK;
}

int bg_camera_rays_device(bg_ctx* ctx, const bg_camera_desc* cam, float fx, float fy, void* d_rays) {
    if (!ctx || !cam || !d_rays) return fail(ctx, BG_ERR_INVALID, "null argument");
    if (cam->width == 0 || cam->height == 0) return fail(ctx, BG_ERR_INVALID, "empty film");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    bg::CameraDev c;
    bg::camera_setup(*cam, &c);
    HIP_TRY(ctx, bg::launch_camera_rays(c, fx, fy, (float4*)d_rays, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return BG_OK;
}

int bg_camera_rays(bg_ctx* ctx, const bg_camera_desc* cam, float fx, float fy, bg_ray* rays) {
    if (!ctx || !cam || !rays) return fail(ctx, BG_ERR_INVALID, "null argument");
    const size_t n = (size_t)cam->width * cam->height;
    int rc = ensure(ctx, &ctx->io, &ctx->io_bytes, n * sizeof(bg_ray));
    if (rc) return rc;
    rc = bg_camera_rays_device(ctx, cam, fx, fy, ctx->io);
    if (rc) return rc;
    HIP_TRY(ctx, hipMemcpy(rays, ctx->io, n * sizeof(bg_ray), hipMemcpyDeviceToHost));
    return BG_OK;
}

}  // extern "C"
