// Host driver of the wavefront path integrator and the film
// (SamplerIntegrator::render, base/integrator.rs:35-133; Film, base/film.rs).
#include <algorithm>
#include <chrono>
#include <cstring>
#include <new>
#include <string>

#include "bg_internal.h"

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

}  // namespace

// Path-state arrays carved out of one allocation, 256-byte aligned each.
int bg::ensure_paths(bg_ctx* ctx, size_t cap) {
    bg::PathBuffers& b = ctx->paths;
    if (b.cap >= cap) return BG_OK;
    if (b.mem) hipFree(b.mem);
    // only the path-state allocation is replaced: the guiding buffers (a separate allocation,
    // sized by ensure_guide_bufs) keep their memory and pointers
    bg::PathBuffers kept = b;
    b = bg::PathBuffers{};
    b.guide_cap = kept.guide_cap;
    b.rec_cap = kept.rec_cap;
    b.guide_mem = kept.guide_mem;
    b.gin = kept.gin;
    b.gout = kept.gout;
    b.gslot = kept.gslot;
    b.gcount = kept.gcount;
    b.gblk = kept.gblk;
    b.rec_in = kept.rec_in;
    b.rec_tgt = kept.rec_tgt;
    b.rec_aux = kept.rec_aux;
    b.rec_blk = kept.rec_blk;
    b.last_records = kept.last_records;
    const size_t per_path[] = {32, 16, 16, 16, 8, 8, 8, 32, 16, 32, 16, 16, 16};
    size_t total = 256;                                                            // counters
    total += 3 * (((4 * (cap / 256 + 1) + 255) / 256) * 256 + 256);              // rblk[3], rcount[3]
    total += 6 * (((32 * (cap / 256 + 1) + 255) / 256) * 256);                   // rmask[3], wmask[3]
    total += 3 * (((4 * cap + 255) / 256) * 256);                                // ridx[3]
    for (size_t s : per_path) total += ((s * cap + 255) / 256) * 256;
    void* mem = nullptr;
    HIP_TRY(ctx, hipMalloc(&mem, total));
    char* p = (char*)mem;
    auto take = [&](size_t bytes) {
        void* r = p;
        p += ((bytes + 255) / 256) * 256;
        return r;
    };
    b.mem = mem;
    b.ray = (float4*)take(32 * cap);
    b.hit = (uint4*)take(16 * cap);
    b.beta = (float4*)take(16 * cap);
    b.L = (float4*)take(16 * cap);
    b.key = (uint2*)take(8 * cap);
    b.state = (uint2*)take(8 * cap);
    b.film = (float2*)take(8 * cap);
    b.sray = (float4*)take(32 * cap);
    b.shit = (uint4*)take(16 * cap);
    b.mray = (float4*)take(32 * cap);
    b.mhit = (uint4*)take(16 * cap);
    b.nee_a = (float4*)take(16 * cap);
    b.nee_b = (float4*)take(16 * cap);
    for (int k = 0; k < 3; ++k) {
        b.ridx[k] = (uint32_t*)take(4 * cap);
        b.rblk[k] = (uint32_t*)take(4 * (cap / 256 + 1));
        b.rmask[k] = (uint64_t*)take(32 * (cap / 256 + 1));
        b.wmask[k] = (uint64_t*)take(32 * (cap / 256 + 1));
        b.rcount[k] = (uint32_t*)take(256);
    }
    b.counters = (unsigned long long*)take(256);
    b.cap = cap;
    return BG_OK;
}

// The two side streams (default priority) and the events that fork and join them, made once.
int bg::ensure_side_streams(bg_ctx* ctx) {
    if (ctx->side[0]) return BG_OK;
    for (int k = 0; k < 2; ++k) {
        HIP_TRY(ctx, hipStreamCreateWithFlags(&ctx->side[k], hipStreamNonBlocking));
        HIP_TRY(ctx, hipEventCreateWithFlags(&ctx->join_ev[k], hipEventDisableTiming));
    }
    HIP_TRY(ctx, hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming));
    return BG_OK;
}

int bg::ensure_spill(bg_ctx* ctx, size_t need) {
    if (ctx->spill_bytes >= need) return BG_OK;
    if (ctx->spill) hipFree(ctx->spill);
    ctx->spill = nullptr;
    ctx->spill_bytes = 0;
    HIP_TRY(ctx, hipMalloc((void**)&ctx->spill, need));
    ctx->spill_bytes = need;
    return BG_OK;
}

namespace {

// Guiding buffers: network input/output per path and training records.
int ensure_guide_bufs(bg_ctx* ctx, size_t cap, size_t rec_cap) {
    bg::PathBuffers& b = ctx->paths;
    if (b.guide_cap >= cap && b.rec_cap >= rec_cap) return BG_OK;
    if (b.guide_mem) hipFree(b.guide_mem);
    b.guide_mem = nullptr;
    b.guide_cap = b.rec_cap = 0;
    cap = std::max(cap, (size_t)1);
    rec_cap = std::max(rec_cap, (size_t)1);
    const size_t sizes[] = {48 * cap, 160 * cap, 48 * rec_cap, 16 * rec_cap, 32 * rec_cap, 4 * (rec_cap / 256 + 1),
                            4 * cap, 256, 4 * (cap / 256 + 1)};
    size_t total = 0;
    for (size_t z : sizes) total += ((z + 255) / 256) * 256;
    HIP_TRY(ctx, hipMalloc(&b.guide_mem, total));
    char* q = (char*)b.guide_mem;
    void* parts[9];
    for (int k = 0; k < 9; ++k) {
        parts[k] = q;
        q += ((sizes[k] + 255) / 256) * 256;
    }
    b.gin = (float4*)parts[0];
    b.gout = (float*)parts[1];
    b.rec_in = (float4*)parts[2];
    b.rec_tgt = (float4*)parts[3];
    b.rec_aux = (float4*)parts[4];
    b.rec_blk = (uint32_t*)parts[5];
    b.gslot = (uint32_t*)parts[6];
    b.gcount = (uint32_t*)parts[7];
    b.gblk = (uint32_t*)parts[8];
    b.guide_cap = cap;
    b.rec_cap = rec_cap;
    return BG_OK;
}

// The guide's replay ring (allocated at its first trained render).
int ensure_replay(bg_guide* g, uint32_t cap) {
    bg_ctx* ctx = g->ctx;
    if (g->replay.cap == cap) return BG_OK;
    if (g->replay_mem) hipFree(g->replay_mem);
    g->replay_mem = nullptr;
    g->replay = bg::Replay{};
    const size_t bytes = 256 + ((48ull * cap + 255) / 256) * 256 + 16ull * cap;
    HIP_TRY(ctx, hipMalloc(&g->replay_mem, bytes));
    char* q = (char*)g->replay_mem;
    g->replay.state = (uint32_t*)q;
    g->replay.in = (float4*)(q + 256);
    g->replay.tgt = (float4*)(q + 256 + ((48ull * cap + 255) / 256) * 256);
    g->replay.cap = cap;
    HIP_TRY(ctx, hipMemsetAsync(g->replay.state, 0, 16, ctx->stream));
    return BG_OK;
}

// set: the ray set (0 extension, 1 shadow, 2 MIS branch) whose spill rows and queue counter the
// launch uses, so that launches of different sets can run at once.
bg::TraverseArgs traverse_args(const bg_scene* s, bg_ctx* ctx, const float4* rays, uint4* hits, uint64_t n,
                               int set = 0) {
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
    a.rays = rays;
    a.hits = hits;
    a.n_rays = n;
    a.spill = ctx->spill + (set ? set * (ctx->spill_set_bytes / sizeof(uint32_t)) : 0);
    a.spill_depth = s->bvh.max_depth + 1;
    a.counters = nullptr;
    a.queue = ctx->queue + 8 * set;
    return a;
}

// One traversal launch of a path-integrator pass over n path slots (the path buffers hold at least
// n) of ray set `set`. With the launch shape's compaction (launches of 1M slots or more) only the
// rays that can hit are traced, through an index list built on the device from the masks the ray
// writer stored; the kernels that read hits read only those of rays that can hit.
hipError_t trace_pass(const bg_scene* s, bg_ctx* ctx, const float4* rays, uint4* hits, uint64_t n, bool any, int grid,
                      int set, const uint64_t* mask, hipStream_t st) {
    bg::TraverseArgs a = traverse_args(s, ctx, rays, hits, n, set);
    const bg::LaunchShape& shape = ctx->render_shape_for(s->n_tris + s->n_spheres);
    if (bg::traverse_compacts(a, shape, grid)) {
        const bg::PathBuffers& b = ctx->paths;
        const hipError_t e =
            bg::launch_ray_compact(rays, n, b.rblk[set], b.rmask[set], b.ridx[set], b.rcount[set], hits, st, mask);
        if (e != hipSuccess) return e;
        a.idx = b.ridx[set];
        a.n_idx = b.rcount[set];
    }
    return bg::profiled_traverse(ctx, a, any, shape, grid, st);
}

}  // namespace

// A render's traversal launch with the context's measurement (bg_set_profiling): test counting
// into ctx->counters, and a HIP event pair around the kernel.
hipError_t bg::profiled_traverse(bg_ctx* ctx, bg::TraverseArgs a, bool any, const bg::LaunchShape& shape, int grid,
                                  hipStream_t st) {
    const bool count = (ctx->profile & BG_PROFILE_COUNT) != 0;
    if (count) a.counters = ctx->counters;
    const bool timed = (ctx->profile & BG_PROFILE_TRAVERSAL_TIME) != 0;
    if (timed) {
        while (ctx->prof_events.size() < 2 * (ctx->prof_used + 1)) {
            hipEvent_t e;
            const hipError_t r = hipEventCreate(&e);
            if (r != hipSuccess) return r;
            ctx->prof_events.push_back(e);
        }
        const hipError_t r = hipEventRecord(ctx->prof_events[2 * ctx->prof_used], st);
        if (r != hipSuccess) return r;
    }
    const hipError_t e = bg::launch_traverse(a, any, count, shape, grid, st);
    if (e != hipSuccess || !timed) return e;
    return hipEventRecord(ctx->prof_events[2 * ctx->prof_used++ + 1], st);
}

// mask: the ray masks the kernel that wrote the rays stored (PathBuffers::wmask).
// Start and end of a render call's measurement.
int bg::profile_begin(bg_ctx* ctx) {
    ctx->prof_used = 0;
    if (ctx->profile & BG_PROFILE_COUNT) {
        if (!ctx->counters) HIP_TRY(ctx, hipMalloc((void**)&ctx->counters, 3 * sizeof(unsigned long long)));
        HIP_TRY(ctx, hipMemsetAsync(ctx->counters, 0, 3 * sizeof(unsigned long long), ctx->stream));
    }
    return BG_OK;
}

int bg::profile_end(bg_ctx* ctx, bg_stats* stats) {
    if (!stats) return BG_OK;
    if (ctx->profile & BG_PROFILE_COUNT) {
        unsigned long long c[3] = {0, 0, 0};
        HIP_TRY(ctx, hipMemcpy(c, ctx->counters, sizeof(c), hipMemcpyDeviceToHost));
        stats->nodes_visited = c[0];
        stats->tris_tested = c[1];
        stats->spheres_tested = c[2];
    }
    if (ctx->profile & BG_PROFILE_TRAVERSAL_TIME) {
        // the time during which at least one traversal kernel ran: launches of different ray sets
        // overlap on the side streams, so the union of the [start, end] intervals, each taken from
        // the first launch's start event
        std::vector<std::pair<double, double>> iv(ctx->prof_used);
        for (size_t k = 0; k < ctx->prof_used; ++k) {
            float a = 0.0f, e = 0.0f;
            HIP_TRY(ctx, hipEventElapsedTime(&a, ctx->prof_events[0], ctx->prof_events[2 * k]));
            HIP_TRY(ctx, hipEventElapsedTime(&e, ctx->prof_events[0], ctx->prof_events[2 * k + 1]));
            iv[k] = {a, e};
        }
        std::sort(iv.begin(), iv.end());
        double sum = 0.0, lo = 0.0, hi = -1.0;
        for (const auto& x : iv) {
            if (x.first > hi) {
                if (hi > lo) sum += hi - lo;
                lo = x.first;
                hi = x.second;
            } else if (x.second > hi) {
                hi = x.second;
            }
        }
        if (hi > lo) sum += hi - lo;
        stats->traverse_ms = sum;
    }
    return BG_OK;
}

// The guide's replay ring of `cap` samples (kept when the capacity is unchanged).
int bg::guide_ensure_replay(bg_guide* g, uint32_t cap) { return ensure_replay(g, cap); }

extern "C" {

int bg_film_create(bg_ctx* ctx, uint32_t width, uint32_t height, uint32_t flags, bg_film** out) {
    if (!ctx || !out || width == 0 || height == 0 || (flags & ~BG_FILM_F64)) return fail(ctx, BG_ERR_INVALID, "bad film");
    *out = nullptr;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    bg_film* f = new (std::nothrow) bg_film();
    if (!f) return BG_ERR_NOMEM;
    f->ctx = ctx;
    f->width = width;
    f->height = height;
    f->flags = flags;
    const size_t n = (size_t)width * height;
    if (hipMalloc((void**)&f->sums, n * sizeof(float4)) != hipSuccess ||
        ((flags & BG_FILM_F64) && hipMalloc((void**)&f->sums64, n * sizeof(double4)) != hipSuccess)) {
        bg_film_destroy(f);
        return fail(ctx, BG_ERR_NOMEM, "film allocation failed");
    }
    *out = f;
    return bg_film_clear(f);
}

void bg_film_destroy(bg_film* f) {
    if (!f) return;
    if (f->sums) hipFree(f->sums);
    if (f->sums64) hipFree(f->sums64);
    delete f;
}

int bg_film_clear(bg_film* f) {
    if (!f) return BG_ERR_INVALID;
    bg_ctx* ctx = f->ctx;
    const size_t n = (size_t)f->width * f->height;
    HIP_TRY(ctx, hipMemsetAsync(f->sums, 0, n * sizeof(float4), ctx->stream));
    if (f->sums64) HIP_TRY(ctx, hipMemsetAsync(f->sums64, 0, n * sizeof(double4), ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return BG_OK;
}

int bg_film_read(bg_film* f, float* rgbw, double* rgbw64) {
    if (!f) return BG_ERR_INVALID;
    bg_ctx* ctx = f->ctx;
    const size_t n = (size_t)f->width * f->height;
    if (rgbw64 && !f->sums64) return fail(ctx, BG_ERR_INVALID, "film has no FP64 sums");
    if (rgbw) HIP_TRY(ctx, hipMemcpy(rgbw, f->sums, n * sizeof(float4), hipMemcpyDeviceToHost));
    if (rgbw64) HIP_TRY(ctx, hipMemcpy(rgbw64, f->sums64, n * sizeof(double4), hipMemcpyDeviceToHost));
    return BG_OK;
}

int bg_film_load(bg_film* f, const float* rgbw) {
    if (!f || !rgbw) return fail(f ? f->ctx : nullptr, BG_ERR_INVALID, "null argument");
    bg_ctx* ctx = f->ctx;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const size_t n = (size_t)f->width * f->height;
    HIP_TRY(ctx, hipMemcpy(f->sums, rgbw, n * sizeof(float4), hipMemcpyHostToDevice));
    if (f->sums64) {
        std::vector<double> w64(rgbw, rgbw + 4 * n);
        HIP_TRY(ctx, hipMemcpy(f->sums64, w64.data(), n * sizeof(double4), hipMemcpyHostToDevice));
    }
    return BG_OK;
}

int bg_film_resolve_device(bg_film* f, float splat_scale, float scale, void* d_rgb, double* ms) {
    if (!f || !d_rgb) return fail(f ? f->ctx : nullptr, BG_ERR_INVALID, "null argument");
    bg_ctx* ctx = f->ctx;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const uint64_t n = (uint64_t)f->width * f->height;
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    HIP_TRY(ctx, bg::launch_film_resolve(f->sums, n, splat_scale, scale, (float*)d_rgb, ctx->stream));
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    HIP_TRY(ctx, hipEventSynchronize(ctx->ev1));
    if (ms) {
        float t = 0.0f;
        HIP_TRY(ctx, hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
        *ms = t;
    }
    return BG_OK;
}

int bg_film_write_image(bg_film* f, float splat_scale, float scale, float* rgb_out, const char* path) {
    if (!f) return BG_ERR_INVALID;
    bg_ctx* ctx = f->ctx;
    if (!rgb_out && !path) return fail(ctx, BG_ERR_INVALID, "neither an output buffer nor a path");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const uint64_t n = (uint64_t)f->width * f->height;
    float* d_rgb = nullptr;
    if (hipMalloc((void**)&d_rgb, 3 * n * sizeof(float)) != hipSuccess) return fail(ctx, BG_ERR_NOMEM, "resolve buffer");
    std::vector<float> host;
    float* dst = rgb_out;
    if (!dst) {
        host.resize(3 * n);
        dst = host.data();
    }
    hipError_t e = bg::launch_film_resolve(f->sums, n, splat_scale, scale, d_rgb, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dst, d_rgb, 3 * n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    hipFree(d_rgb);
    if (e != hipSuccess) return fail(ctx, BG_ERR_HIP, hipGetErrorString(e));
    if (path && bg::exr_write_file(path, f->width, f->height, dst) != BG_OK)
        return fail(ctx, BG_ERR_INVALID, "cannot write the image file");
    return BG_OK;
}

}  // extern "C"

namespace {

// Checks a render call's arguments and resolves its pixel window.
int check_render(bg_ctx* ctx, const bg_scene* s, const bg_render_desc* d, const bg_film* film, uint32_t win[4]) {
    if (!ctx || !s || !d || !film || s->ctx != ctx || film->ctx != ctx) return fail(ctx, BG_ERR_INVALID, "null or foreign argument");
    const bg_camera_desc& cam = d->camera;
    if (cam.width != film->width || cam.height != film->height) return fail(ctx, BG_ERR_INVALID, "film size differs from camera");
    if (cam.lens_radius > 0.0f) return fail(ctx, BG_ERR_UNSUPPORTED, "thin-lens cameras are not supported on the GPU path");
    if (d->mode > BG_MODE_CORRECTED) return fail(ctx, BG_ERR_INVALID, "unknown mode");
    if (s->materials.empty()) return fail(ctx, BG_ERR_INVALID, "scene has no materials");
    uint32_t x0 = d->x0, y0 = d->y0, x1 = d->x1, y1 = d->y1;
    if (x1 == 0 && y1 == 0) { x0 = 0; y0 = 0; x1 = cam.width; y1 = cam.height; }
    if (x0 >= x1 || y0 >= y1 || x1 > cam.width || y1 > cam.height) return fail(ctx, BG_ERR_INVALID, "bad window");
    win[0] = x0; win[1] = y0; win[2] = x1; win[3] = y1;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    return BG_OK;
}

bg::ShadeSceneDev shade_scene(const bg_scene* s) {
    bg::ShadeSceneDev sc;
    sc.tri_orig = s->d_tri_orig;
    sc.prim_info = s->d_prim_info;
    sc.mats = s->d_mats;
    sc.lights = s->d_lights;
    sc.spheres = s->d_sph_shade;
    sc.n_tris = (uint32_t)s->n_tris;
    sc.n_lights = (uint32_t)s->lights.size();
    sc.multi_lobe = bg::has_multi_lobe(s->materials);
    sc.sphere_lights = bg::has_sphere_lights(s->lights, s->n_tris);
    return sc;
}

int render_impl(bg_ctx* ctx, const bg_scene* s, const bg_render_desc* d, bg_film* film, bg_stats* stats,
                bg_guide* guide, const bg_guiding_desc* gd) {
    uint32_t win[4];
    int rc = check_render(ctx, s, d, film, win);
    if (rc) return rc;
    const bg_camera_desc& cam = d->camera;
    const uint32_t x0 = win[0], y0 = win[1], x1 = win[2], y1 = win[3];
    if (stats) *stats = bg_stats{};
    if (d->spp == 0 || s->bvh.nodes.empty()) return BG_OK;

    const uint64_t npix = (uint64_t)(x1 - x0) * (y1 - y0);
    // default: about 64M paths per pass (264 bytes each, 476 guided: 17 / 31 GB of the 288 GB), so
    // that the deep bounces, whose launches carry few live rays, still fill the chip (1080p at 64
    // spp, depth 8: 0.162 s in 16M-path passes, 0.150 s in 32M, 0.143 s in 64M, 0.142 s in one
    // pass of 133M); while a guide trains, 16M, as one training round follows each pass
    const bool training = guide != nullptr && gd != nullptr && gd->train;
    uint64_t target = d->paths_per_pass ? d->paths_per_pass : (uint64_t)(training ? 16 : 64) << 20;
    uint32_t spp_pass = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(d->spp, target / npix));
    const uint64_t cap = npix * spp_pass;
    rc = bg::ensure_paths(ctx, cap);
    if (rc) return rc;
    const bg::LaunchShape& shape = ctx->render_shape_for(s->n_tris + s->n_spheres);
    const int block = shape.block;
    const uint64_t grid_cap = shape.blocks_per_cu > 0 ? (uint64_t)ctx->num_cus * shape.blocks_per_cu : (1u << 20);
    const uint64_t grid = std::max<uint64_t>(1, std::min<uint64_t>((cap + block - 1) / block, std::min<uint64_t>(grid_cap, 1u << 20)));
    // spill rows for each of the three ray sets, which trace at once
    const size_t per_set = (((size_t)grid * block * (s->bvh.max_depth + 1) * bg::kSpillWords * sizeof(uint32_t) + 255) / 256) * 256;
    rc = bg::ensure_spill(ctx, 3 * per_set);
    if (rc) return rc;
    ctx->spill_set_bytes = per_set;
    rc = bg::ensure_side_streams(ctx);
    if (rc) return rc;

    bg::RenderParamsDev p;
    bg::camera_setup(cam, &p.cam);
    p.seed = d->seed;
    p.x0 = x0; p.y0 = y0; p.x1 = x1; p.y1 = y1;
    p.max_depth = d->max_depth;
    p.rr = d->rr_threshold;
    p.mode = d->mode;
    const bool guided = guide != nullptr;
    uint64_t train_paths = 0, batch = 0;
    uint32_t steps = 0;
    float lr = 1e-2f;
    if (guided) {
        if (guide->ctx != ctx) return fail(ctx, BG_ERR_INVALID, "guide belongs to another context");
        if (d->mode != BG_MODE_CORRECTED) return fail(ctx, BG_ERR_INVALID, "guiding needs BG_MODE_CORRECTED");
        if (!(gd->select_prob >= 0.0f && gd->select_prob <= 1.0f)) return fail(ctx, BG_ERR_INVALID, "select_prob outside [0, 1]");
        p.guided = 1;
        p.guide_alpha = gd->select_prob;
        p.guide_lobes = gd->lobes ? gd->lobes : 3u;
        p.guide_bounces = gd->bounces ? gd->bounces : 0xFFFFFFFFu;
        if (gd->rr_start) p.rr_start = gd->rr_start;
        if (gd->train) {
            train_paths = gd->train_paths ? gd->train_paths : (1u << 20);
            batch = std::min<uint64_t>(gd->batch ? gd->batch : 16384, guide->scratch.cap);
            steps = gd->steps ? gd->steps : 8;
            lr = gd->lr > 0.0f ? gd->lr : 1e-2f;
            rc = ensure_replay(guide, gd->replay ? gd->replay : (1u << 20));
            if (rc) return rc;
        }
        rc = ensure_guide_bufs(ctx, cap, train_paths * d->max_depth);
        if (rc) return rc;
        // positions are normalised to the scene's bounds
        const bg_bvh_node& root = s->bvh.nodes[0];
        for (int k = 0; k < 3; ++k) {
            const float ext = std::max(root.bmax[k] - root.bmin[k], 1e-6f);
            guide->box.lo[k] = root.bmin[k] - 1e-3f * ext;
            guide->box.hi[k] = root.bmax[k] + 1e-3f * ext;
        }
    }
    const int guide_grid = bg::kTrainBlocksPerCu * std::max(ctx->num_cus, 1);
    const int infer_grid = bg::kInferBlocksPerCu * std::max(ctx->num_cus, 1);
    uint64_t train_steps = 0;
    const bg::ShadeSceneDev sc = shade_scene(s);
    const bg::PathBuffers& b = ctx->paths;
    hipStream_t st = ctx->stream;

    HIP_TRY(ctx, hipMemsetAsync(b.counters, 0, 2 * sizeof(unsigned long long), st));
    rc = bg::profile_begin(ctx);
    if (rc) return rc;
    auto t0 = std::chrono::steady_clock::now();
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, st));
    uint64_t launches = 0;
    for (uint32_t s0 = 0; s0 < d->spp; s0 += spp_pass) {
        const uint32_t ns = std::min(spp_pass, d->spp - s0);
        p.first_sample = d->first_sample + s0;
        p.n_samples = ns;
        p.n_paths = npix * ns;
        const int g = (int)std::min<uint64_t>(grid, (p.n_paths + block - 1) / block);
        if (train_paths) {
            // spread the recording paths over the whole pass (paths are pixel-major within a sample)
            p.rec_stride = (uint32_t)std::max<uint64_t>(1, (p.n_paths + train_paths - 1) / train_paths);
            p.rec_paths = (uint32_t)std::min<uint64_t>(train_paths, (p.n_paths + p.rec_stride - 1) / p.rec_stride);
        }
        HIP_TRY(ctx, bg::launch_init_paths(p, b, st));
        HIP_TRY(ctx, trace_pass(s, ctx, b.ray, b.hit, p.n_paths, false, g, 0, b.wmask[0], st));
        for (uint32_t bounce = 0; bounce <= d->max_depth; ++bounce) {
            if (guided && bounce < d->max_depth && bounce < p.guide_bounces && (p.guide_alpha > 0.0f || train_paths)) {
                HIP_TRY(ctx, bg::launch_guide_inputs(p, sc, b, st));  // the records copy these inputs
                launches += 3;
                if (p.guide_alpha > 0.0f) {  // the live paths only (count on the device)
                    HIP_TRY(ctx, bg::guide_infer(guide->dev, guide->box, b.gin, p.n_paths, b.gout, infer_grid, st,
                                                 b.gcount, true));
                    ++launches;
                }
            }
            HIP_TRY(ctx, bg::launch_shade(p, sc, b, bounce, st));
            // The shadow and MIS-branch rays of this bounce trace on the side streams while the next
            // bounce's extension rays trace here: the three read only what shading wrote and write
            // their own hits, compaction buffers and spill rows. NEE resolves after all of them.
            const bool nee = sc.n_lights > 0 && bounce < d->max_depth;
            if (nee) {
                HIP_TRY(ctx, hipEventRecord(ctx->fork_ev, st));
                HIP_TRY(ctx, hipStreamWaitEvent(ctx->side[0], ctx->fork_ev, 0));
                HIP_TRY(ctx, trace_pass(s, ctx, b.sray, b.shit, p.n_paths, true, g, 1, b.wmask[1], ctx->side[0]));
                HIP_TRY(ctx, hipEventRecord(ctx->join_ev[0], ctx->side[0]));
                // the guided estimator has no separate MIS ray: its continuation plays that part
                if (!guided) {
                    HIP_TRY(ctx, hipStreamWaitEvent(ctx->side[1], ctx->fork_ev, 0));
                    HIP_TRY(ctx, trace_pass(s, ctx, b.mray, b.mhit, p.n_paths, false, g, 2, b.wmask[2], ctx->side[1]));
                    HIP_TRY(ctx, hipEventRecord(ctx->join_ev[1], ctx->side[1]));
                    ++launches;
                }
            }
            if (bounce < d->max_depth) {
                HIP_TRY(ctx, trace_pass(s, ctx, b.ray, b.hit, p.n_paths, false, g, 0, b.wmask[0], st));
                ++launches;
            }
            if (nee) {
                HIP_TRY(ctx, hipStreamWaitEvent(st, ctx->join_ev[0], 0));
                if (!guided) HIP_TRY(ctx, hipStreamWaitEvent(st, ctx->join_ev[1], 0));
                HIP_TRY(ctx, bg::launch_nee_resolve(p, sc, b, bounce, st));
                launches += 2;
            }
            ++launches;
        }
        ++launches;  // the first extension launch
        HIP_TRY(ctx, bg::launch_film(p, b, film->width, film->height, film->sums, film->sums64, st));
        HIP_TRY(ctx, bg::launch_count_rays(p, b, b.counters, st));
        launches += 3;
        if (train_paths && p.rec_paths) {
            // one training round on this pass's records
            HIP_TRY(ctx, bg::launch_guide_targets(p, b, st));
            const uint64_t nrec = (uint64_t)p.rec_paths * d->max_depth;
            ctx->paths.last_records = nrec;
            HIP_TRY(ctx, bg::launch_replay_append(b, nrec, guide->replay, st));
            uint32_t rs[4];
            HIP_TRY(ctx, hipMemcpyAsync(rs, guide->replay.state, sizeof(rs), hipMemcpyDeviceToHost, st));
            HIP_TRY(ctx, hipStreamSynchronize(st));
            const uint64_t avail = rs[1];
            const uint64_t nb = std::min<uint64_t>(batch, avail);
            for (uint32_t k = 0; nb > 0 && k < steps; ++k) {
                HIP_TRY(ctx, bg::launch_replay_gather(guide->replay, (uint64_t)guide->step * nb, nb, guide->batch_in,
                                                      guide->batch_tgt, st));
                ++guide->step;
                HIP_TRY(ctx, bg::guide_train_step(guide->dev, guide->box, guide->batch_in, guide->batch_tgt, nb,
                                                  guide->scratch, lr, guide->step, guide_grid, nullptr, st));
                ++train_steps;
                launches += 3;  // gather, train, Adam
            }
            launches += 4;
        }
    }
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, st));
    HIP_TRY(ctx, hipEventSynchronize(ctx->ev1));
    auto t1 = std::chrono::steady_clock::now();
    float ms = 0.0f;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    unsigned long long nr[2] = {0, 0};
    HIP_TRY(ctx, hipMemcpy(nr, b.counters, sizeof(nr), hipMemcpyDeviceToHost));
    if (stats) {
        stats->rays = nr[0];
        stats->shadow_rays = nr[1];
        stats->kernel_ms = ms;
        stats->total_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        stats->launches = launches;
        stats->samples = npix * d->spp;
        stats->train_steps = train_steps;
    }
    rc = bg::profile_end(ctx, stats);
    if (rc) return rc;
    return BG_OK;
}

// Light-slot arrays of the direct-lighting and Whitted integrators, one allocation.
int ensure_direct(bg_ctx* ctx, size_t slots) {
    bg::DirectSlots& d = ctx->direct;
    if (d.cap >= slots) return BG_OK;
    if (d.mem) hipFree(d.mem);
    d = bg::DirectSlots{};
    const size_t per_slot[] = {32, 16, 32, 16, 16, 16, 8};
    size_t total = 0;
    for (size_t b : per_slot) total += ((b * slots + 255) / 256) * 256;
    void* mem = nullptr;
    HIP_TRY(ctx, hipMalloc(&mem, total));
    char* p = (char*)mem;
    auto take = [&](size_t bytes) {
        void* r = p;
        p += ((bytes + 255) / 256) * 256;
        return r;
    };
    d.mem = mem;
    d.sray = (float4*)take(32 * slots);
    d.shit = (uint4*)take(16 * slots);
    d.mray = (float4*)take(32 * slots);
    d.mhit = (uint4*)take(16 * slots);
    d.a = (float4*)take(16 * slots);
    d.b = (float4*)take(16 * slots);
    d.c = (float2*)take(8 * slots);
    d.cap = slots;
    return BG_OK;
}

// SamplerIntegrator::render (integrator.rs:35-133) with DirectLightingIntegrator
// or WhittedIntegrator radiance: camera rays, their closest hits, the light
// slots, the shadow and MIS queries of every slot, the per-path sums, the film.
int direct_impl(bg_ctx* ctx, const bg_scene* s, const bg_render_desc* d, const bg_direct_desc* dd, bg_film* film,
                bg_stats* stats) {
    uint32_t win[4];
    int rc = check_render(ctx, s, d, film, win);
    if (rc) return rc;
    if (!dd) return fail(ctx, BG_ERR_INVALID, "null integrator descriptor");
    if (dd->integrator != BG_INTEGRATOR_DIRECT && dd->integrator != BG_INTEGRATOR_WHITTED)
        return fail(ctx, BG_ERR_INVALID, "unknown integrator");
    if (dd->integrator == BG_INTEGRATOR_DIRECT && dd->strategy > BG_LIGHT_SAMPLE_ONE)
        return fail(ctx, BG_ERR_INVALID, "unknown light strategy");
    const int kind = dd->integrator == BG_INTEGRATOR_WHITTED ? 2 : (dd->strategy == BG_LIGHT_SAMPLE_ONE ? 1 : 0);
    const uint32_t n_lights = (uint32_t)s->lights.size();
    // per path: the camera query plus up to two per slot, counted in 16 bits
    if (kind != 1 && n_lights > 16383) return fail(ctx, BG_ERR_UNSUPPORTED, "more than 16383 lights with a slot per light");
    if (stats) *stats = bg_stats{};
    if (d->spp == 0 || s->bvh.nodes.empty()) return BG_OK;
    const uint32_t x0 = win[0], y0 = win[1], x1 = win[2], y1 = win[3];
    const uint64_t npix = (uint64_t)(x1 - x0) * (y1 - y0);
    const uint64_t n_slots = kind == 1 ? 1 : std::max<uint32_t>(n_lights, 1u);
    // a pass holds whole samples of the window; its slots stay within a budget when they can:
    // about 64M slots (136 bytes each, 8.7 GB). The 1080p, 16-spp renders of the 8-light scene
    // measured all lights 0.065 / 0.057 s, one light 0.018 / 0.014 s and Whitted 0.055 / 0.046 s
    // with budgets of 4M / 64M slots (256M: within 3% of 64M).
    const uint64_t target = d->paths_per_pass ? d->paths_per_pass : (uint64_t)64 << 20;
    const uint32_t spp_pass =
        (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(d->spp, std::max<uint64_t>(target, 1) / (npix * n_slots)));
    const uint64_t cap = npix * spp_pass;
    rc = bg::ensure_paths(ctx, cap);
    if (!rc) rc = ensure_direct(ctx, cap * n_slots);
    if (rc) return rc;
    // the queries' shape: with the render calls' shape (128 threads, 16 blocks per CU, variant 5)
    // the same renders measured 0.069 / 0.016 / 0.053 s against 0.059 / 0.015 / 0.046 s
    const int block = ctx->shape.block;
    const uint64_t grid_cap = ctx->shape.blocks_per_cu > 0 ? (uint64_t)ctx->num_cus * ctx->shape.blocks_per_cu : (1u << 20);
    const uint64_t grid = std::max<uint64_t>(1, std::min<uint64_t>((cap * n_slots + block - 1) / block,
                                                                    std::min<uint64_t>(grid_cap, 1u << 20)));
    // spill rows for the camera rays and the two light-slot ray sets, which trace at once
    const size_t per_set = (((size_t)grid * block * (s->bvh.max_depth + 1) * bg::kSpillWords * sizeof(uint32_t) + 255) / 256) * 256;
    rc = bg::ensure_spill(ctx, 3 * per_set);
    if (rc) return rc;
    ctx->spill_set_bytes = per_set;
    rc = bg::ensure_side_streams(ctx);
    if (rc) return rc;

    bg::RenderParamsDev p;
    bg::camera_setup(d->camera, &p.cam);
    p.seed = d->seed;
    p.x0 = x0; p.y0 = y0; p.x1 = x1; p.y1 = y1;
    p.max_depth = d->max_depth;
    p.rr = d->rr_threshold;
    p.mode = d->mode;
    const bg::ShadeSceneDev sc = shade_scene(s);
    const bg::PathBuffers& b = ctx->paths;
    const bg::DirectSlots& ds = ctx->direct;
    hipStream_t st = ctx->stream;
    HIP_TRY(ctx, hipMemsetAsync(b.counters, 0, 2 * sizeof(unsigned long long), st));
    rc = bg::profile_begin(ctx);
    if (rc) return rc;
    auto t0 = std::chrono::steady_clock::now();
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, st));
    uint64_t launches = 0;
    for (uint32_t s0 = 0; s0 < d->spp; s0 += spp_pass) {
        const uint32_t ns = std::min(spp_pass, d->spp - s0);
        p.first_sample = d->first_sample + s0;
        p.n_samples = ns;
        p.n_paths = npix * ns;
        const uint64_t nq = p.n_paths * (n_lights ? n_slots : 0);
        const int g = (int)std::min<uint64_t>(grid, (p.n_paths + block - 1) / block);
        const int gq = (int)std::max<uint64_t>(1, std::min<uint64_t>(grid, (nq + block - 1) / block));
        HIP_TRY(ctx, bg::launch_init_paths(p, b, st));
        HIP_TRY(ctx, bg::profiled_traverse(ctx, traverse_args(s, ctx, b.ray, b.hit, p.n_paths), false, ctx->shape, g, st));
        HIP_TRY(ctx, bg::launch_direct_shade(p, sc, b, ds, kind, st));
        launches += 3;
        if (nq) {
            // the shadow and BSDF-sampling slots trace at once on the side streams
            HIP_TRY(ctx, hipEventRecord(ctx->fork_ev, st));
            HIP_TRY(ctx, hipStreamWaitEvent(ctx->side[0], ctx->fork_ev, 0));
            HIP_TRY(ctx, bg::profiled_traverse(ctx, traverse_args(s, ctx, ds.sray, ds.shit, nq, 1), true, ctx->shape, gq,
                                               ctx->side[0]));
            HIP_TRY(ctx, hipEventRecord(ctx->join_ev[0], ctx->side[0]));
            ++launches;
            if (kind != 2) {
                HIP_TRY(ctx, hipStreamWaitEvent(ctx->side[1], ctx->fork_ev, 0));
                HIP_TRY(ctx, bg::profiled_traverse(ctx, traverse_args(s, ctx, ds.mray, ds.mhit, nq, 2), false, ctx->shape,
                                                   gq, ctx->side[1]));
                HIP_TRY(ctx, hipEventRecord(ctx->join_ev[1], ctx->side[1]));
                ++launches;
            }
            HIP_TRY(ctx, hipStreamWaitEvent(st, ctx->join_ev[0], 0));
            if (kind != 2) HIP_TRY(ctx, hipStreamWaitEvent(st, ctx->join_ev[1], 0));
        }
        HIP_TRY(ctx, bg::launch_direct_resolve(p, sc, b, ds, kind, st));
        HIP_TRY(ctx, bg::launch_film(p, b, film->width, film->height, film->sums, film->sums64, st));
        HIP_TRY(ctx, bg::launch_count_rays(p, b, b.counters, st));
        launches += 3;
    }
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, st));
    HIP_TRY(ctx, hipEventSynchronize(ctx->ev1));
    auto t1 = std::chrono::steady_clock::now();
    float ms = 0.0f;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    unsigned long long nr[2] = {0, 0};
    HIP_TRY(ctx, hipMemcpy(nr, b.counters, sizeof(nr), hipMemcpyDeviceToHost));
    if (stats) {
        stats->rays = nr[0];
        stats->shadow_rays = nr[1];
        stats->kernel_ms = ms;
        stats->total_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        stats->launches = launches;
        stats->samples = npix * d->spp;
    }
    return bg::profile_end(ctx, stats);
}

}  // namespace

extern "C" {

int bg_render_film(bg_ctx* ctx, const bg_scene* s, const bg_render_desc* d, bg_film* film, bg_stats* stats) {
    return render_impl(ctx, s, d, film, stats, nullptr, nullptr);
}

int bg_render_film_direct(bg_ctx* ctx, const bg_scene* s, const bg_render_desc* d, const bg_direct_desc* dd,
                          bg_film* film, bg_stats* stats) {
    return direct_impl(ctx, s, d, dd, film, stats);
}

int bg_render_film_guided(bg_ctx* ctx, const bg_scene* s, const bg_render_desc* d, bg_guide* guide,
                          const bg_guiding_desc* gd, bg_film* film, bg_stats* stats) {
    if (!guide || !gd) return fail(ctx, BG_ERR_INVALID, "null guide or guiding descriptor");
    return render_impl(ctx, s, d, film, stats, guide, gd);
}

int bg_guide_records(bg_ctx* ctx, float* inputs, float* targets, uint64_t max_records, uint64_t* n) {
    if (!ctx || !n) return fail(ctx, BG_ERR_INVALID, "null argument");
    const bg::PathBuffers& b = ctx->paths;
    const uint64_t m = std::min<uint64_t>(max_records, b.last_records);
    *n = b.last_records;
    if (m == 0) return BG_OK;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (inputs) HIP_TRY(ctx, hipMemcpy(inputs, b.rec_in, 48 * m, hipMemcpyDeviceToHost));
    if (targets) HIP_TRY(ctx, hipMemcpy(targets, b.rec_tgt, 16 * m, hipMemcpyDeviceToHost));
    return BG_OK;
}

int bg_render(bg_ctx* ctx, const bg_scene* s, const bg_render_desc* d, float* rgb_out, bg_stats* stats) {
    if (!ctx || !d || !rgb_out) return fail(ctx, BG_ERR_INVALID, "null argument");
    bg_film* film = nullptr;
    int rc = bg_film_create(ctx, d->camera.width, d->camera.height, 0, &film);
    if (rc) return rc;
    rc = bg_render_film(ctx, s, d, film, stats);
    // integrator.rs:132: film.write_image(1.0) with the film scale 1
    if (!rc) rc = bg_film_write_image(film, 1.0f, 1.0f, rgb_out, nullptr);
    bg_film_destroy(film);
    return rc;
}

}  // extern "C"
