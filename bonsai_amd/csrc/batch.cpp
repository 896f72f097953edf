// Scene batches: many scenes, cameras and films rendered by one wavefront loop
// (no counterpart in the reference, whose SamplerIntegrator::render,
// base/integrator.rs:35-133, renders one scene per call).
//
// The scenes' device data are concatenated once: pair records and leaf tables
// with their references and slot offsets rebased, leaf slots mapped to batch-wide primitive ids
// (every triangle of every scene, then every sphere), shading records with
// material and light indices rebased. A path then needs only its item (in its
// flags) and its BVH root (per path) to run the single-scene kernels. One pass
// (every item's film times samples-per-pass paths) is captured in a hipGraph
// and replayed; the pass reads its first sample index from device memory and
// its last node advances it. Before the first pass an autotuner times the
// traversal launch shapes on the batch's own camera, secondary and shadow
// rays and keeps the fastest.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <vector>

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

template <class T>
int dalloc(bg_ctx* ctx, T** p, size_t count) {
    *p = nullptr;
    if (count == 0) return BG_OK;
    HIP_TRY(ctx, hipMalloc((void**)p, count * sizeof(T)));
    return BG_OK;
}

}  // namespace

struct bg_batch {
    bg_ctx* ctx = nullptr;
    std::vector<bg_batch_item> items;
    std::vector<uint32_t> item_scene;  // item -> unique scene index
    std::vector<uint64_t> item_pix_base;
    uint64_t total_pix = 0;
    // concatenated scene data
    uint64_t n_tris = 0, n_spheres = 0, n_nodes = 0, n_slots = 0, n_pairs = 0, n_leaves = 0;
    uint32_t n_lights = 0, max_depth = 0;
    std::vector<uint32_t> scene_root, scene_light_base, scene_n_lights;
    float4* pairs = nullptr;
    uint2* leaves = nullptr;
    float* tri_verts = nullptr;
    uint32_t* slot_prim = nullptr;
    bg::SphereDev* spheres = nullptr;
    bg::SphereAux* sphere_aux = nullptr;
    float* tri_orig = nullptr;
    uint2* prim_info = nullptr;
    bg::MaterialDev* mats = nullptr;
    bg::LightDev* lights = nullptr;
    bg::SphereShadeDev* sph_shade = nullptr;
    // per render
    bg::BatchItemDev* d_items = nullptr;
    float4* film = nullptr;
    uint32_t* d_pass = nullptr;
    unsigned long long* d_queue = nullptr;
    uint32_t* d_root = nullptr;
    size_t root_cap = 0;
    bg::LaunchShape shape;
    bool tuned = false;
    uint32_t multi_lobe = 0, sphere_lights = 0;
    double tune_ms = 0.0;
};

namespace {

void free_batch(bg_batch* b) {
    void* ptrs[] = {b->pairs, b->leaves, b->tri_verts, b->slot_prim, b->spheres, b->sphere_aux, b->tri_orig,
                    b->prim_info, b->mats, b->lights, b->sph_shade, b->d_items, b->film,
                    b->d_pass, b->d_queue, b->d_root};
    for (void* p : ptrs)
        if (p) hipFree(p);
}

// set: the ray set (0 extension, 1 shadow, 2 MIS branch): its own spill rows and queue counter, so
// that launches of different sets can run at once.
bg::TraverseArgs batch_args(const bg_batch* bt, const bg_ctx* ctx, const float4* rays, uint4* hits, uint64_t n,
                            const uint32_t* roots, bool queue, int set) {
    bg::TraverseArgs a;
    a.nodes = nullptr;  // batches traverse the pair layout only (variants 0, 2, 4 and 5)
    a.pairs = bt->pairs;
    a.leaves = bt->leaves;
    a.tri_verts = bt->tri_verts;
    a.slot_prim = bt->slot_prim;
    a.spheres = bt->spheres;
    a.sphere_aux = bt->sphere_aux;
    a.n_tris = (uint32_t)bt->n_tris;
    a.rays = rays;
    a.hits = hits;
    a.n_rays = n;
    a.spill = ctx->spill + (set ? set * (ctx->spill_set_bytes / sizeof(uint32_t)) : 0);
    a.spill_depth = bt->max_depth + 1;
    a.counters = nullptr;
    a.roots = roots;
    a.queue = queue ? bt->d_queue + 8 * set : nullptr;
    return a;
}

int grid_for(const bg_ctx* ctx, const bg::LaunchShape& sh, uint64_t n) {
    const uint64_t tiles = (n + sh.block - 1) / sh.block;
    const uint64_t cap = sh.blocks_per_cu > 0 ? (uint64_t)ctx->num_cus * sh.blocks_per_cu : tiles;
    return (int)std::max<uint64_t>(1, std::min<uint64_t>(std::min(tiles, cap), 1u << 20));
}

// mask: the ray masks the kernel that wrote the rays stored (PathBuffers::wmask); set: the ray set
// (its spill rows, queue counter and compaction buffers); measured: with the context's measurement
// (bg_set_profiling), as a render's launches are (the autotuner's are not).
hipError_t traverse(const bg_batch* bt, bg_ctx* ctx, const bg::LaunchShape& sh, const float4* rays, uint4* hits,
                    uint64_t n, const uint32_t* roots, bool any, const uint64_t* mask, hipStream_t st,
                    bool measured = false, int set = 0) {
    const bool queue = sh.variant == 2 || sh.variant == 5;  // (the launch zeroes the counter)
    bg::LaunchShape shape = sh;
    if (shape.variant != 2 && shape.variant != 4 && shape.variant != 5) shape.variant = 0;  // the node-record variants need node roots
    bg::TraverseArgs a = batch_args(bt, ctx, rays, hits, n, roots, queue, set);
    const int grid = grid_for(ctx, shape, n);
    if (bg::traverse_compacts(a, shape, grid)) {
        // only the rays that can hit, through a device index list (the others' hits become misses)
        const bg::PathBuffers& b = ctx->paths;
        const hipError_t e =
            bg::launch_ray_compact(rays, n, b.rblk[set], b.rmask[set], b.ridx[set], b.rcount[set], hits, st, mask);
        if (e != hipSuccess) return e;
        a.idx = b.ridx[set];
        a.n_idx = b.rcount[set];
    }
    if (measured) return bg::profiled_traverse(ctx, a, any, shape, grid, st);
    return bg::launch_traverse(a, any, false, shape, grid, st);
}

}  // namespace

extern "C" {

int bg_batch_create(bg_ctx* ctx, const bg_batch_item* items, uint32_t n, bg_batch** out) {
    if (!ctx || !out || (n && !items)) return fail(ctx, BG_ERR_INVALID, "null argument");
    *out = nullptr;
    if (n == 0 || n > 65535) return fail(ctx, BG_ERR_INVALID, "a batch holds 1 to 65535 items");
    std::vector<const bg_scene*> uniq;
    std::map<const bg_scene*, uint32_t> index;
    for (uint32_t k = 0; k < n; ++k) {
        const bg_scene* s = items[k].scene;
        if (!s || s->ctx != ctx) return fail(ctx, BG_ERR_INVALID, "batch item with a null or foreign scene");
        if (s->materials.empty()) return fail(ctx, BG_ERR_INVALID, "batch scene without materials");
        if (s->bvh.nodes.empty()) return fail(ctx, BG_ERR_INVALID, "batch scene without primitives");
        const bg_camera_desc& c = items[k].camera;
        if (c.width == 0 || c.height == 0 || c.lens_radius > 0.0f)
            return fail(ctx, BG_ERR_INVALID, "batch camera: empty film or thin lens");
        if (!index.count(s)) {
            index[s] = (uint32_t)uniq.size();
            uniq.push_back(s);
        }
    }
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    bg_batch* bt = new (std::nothrow) bg_batch();
    if (!bt) return fail(ctx, BG_ERR_NOMEM, "out of host memory");
    bt->ctx = ctx;
    bt->items.assign(items, items + n);
    for (uint32_t k = 0; k < n; ++k) {
        bt->item_scene.push_back(index[items[k].scene]);
        bt->item_pix_base.push_back(bt->total_pix);
        bt->total_pix += (uint64_t)items[k].camera.width * items[k].camera.height;
    }
    // bases of every unique scene in the concatenated arrays
    const size_t U = uniq.size();
    std::vector<uint64_t> node_base(U), slot_base(U), tri_base(U), sph_base(U), mat_base(U), light_base(U);
    std::vector<uint64_t> pair_base(U), leaf_base(U);
    std::vector<std::vector<bg::PairRecord>> spairs(U);
    std::vector<std::vector<uint32_t>> sleaves(U);
    uint64_t n_mats = 0;
    for (size_t u = 0; u < U; ++u) {
        const bg_scene* s = uniq[u];
        bg::build_pairs(s->bvh, &spairs[u], &sleaves[u]);
        pair_base[u] = bt->n_pairs;
        leaf_base[u] = bt->n_leaves;
        bt->n_pairs += spairs[u].size();
        bt->n_leaves += sleaves[u].size() / 2;
        node_base[u] = bt->n_nodes;
        slot_base[u] = bt->n_slots;
        tri_base[u] = bt->n_tris;
        sph_base[u] = bt->n_spheres;
        mat_base[u] = n_mats;
        light_base[u] = bt->n_lights;
        bt->n_nodes += s->bvh.nodes.size();
        bt->n_slots += s->bvh.slot_prim.size();
        bt->n_tris += s->n_tris;
        bt->n_spheres += s->n_spheres;
        n_mats += s->materials.size();
        bt->multi_lobe |= bg::has_multi_lobe(s->materials);
        bt->sphere_lights |= bg::has_sphere_lights(s->lights, s->n_tris);
        bt->n_lights += (uint32_t)s->lights.size();
        bt->max_depth = std::max(bt->max_depth, s->bvh.max_depth);
        bt->scene_root.push_back((uint32_t)pair_base[u]);  // the scene's virtual root record
        bt->scene_light_base.push_back((uint32_t)light_base[u]);
        bt->scene_n_lights.push_back((uint32_t)s->lights.size());
    }
    if (bt->n_pairs >= 0x7FFFFFFFull || bt->n_leaves >= 0x7FFFFFFFull || bt->n_tris + bt->n_spheres >= 0x7FFFFFFFull) {
        delete bt;
        return fail(ctx, BG_ERR_INVALID, "batch too large for 32-bit node and primitive ids");
    }
    const uint64_t np = bt->n_tris + bt->n_spheres;
    std::vector<bg::PairRecord> pairs(bt->n_pairs);
    std::vector<uint32_t> leaves(2 * bt->n_leaves);
    std::vector<uint32_t> slots(bt->n_slots);
    std::vector<uint2> pinfo(np);
    std::vector<bg::LightDev> lts(bt->n_lights);
    for (size_t u = 0; u < U; ++u) {
        const bg_scene* s = uniq[u];
        auto rebase = [&](uint32_t r) -> uint32_t {
            if (r == bg::kRefNone) return r;
            if (r & bg::kRefLeaf) return bg::kRefLeaf | ((r & ~bg::kRefLeaf) + (uint32_t)leaf_base[u]);
            return r + (uint32_t)pair_base[u];
        };
        for (size_t k = 0; k < spairs[u].size(); ++k) {
            bg::PairRecord r = spairs[u][k];
            r.ref0 = rebase(r.ref0);
            r.ref1 = rebase(r.ref1);
            pairs[pair_base[u] + k] = r;
        }
        for (size_t k = 0; k < sleaves[u].size() / 2; ++k) {
            leaves[2 * (leaf_base[u] + k)] = sleaves[u][2 * k] + (uint32_t)slot_base[u];
            leaves[2 * (leaf_base[u] + k) + 1] = sleaves[u][2 * k + 1];
        }
        const uint64_t nt = s->n_tris;
        auto global = [&](uint64_t p) -> uint64_t {
            return p < nt ? tri_base[u] + p : bt->n_tris + sph_base[u] + (p - nt);
        };
        for (size_t k = 0; k < s->bvh.slot_prim.size(); ++k) slots[slot_base[u] + k] = (uint32_t)global(s->bvh.slot_prim[k]);
        for (uint64_t p = 0; p < nt + s->n_spheres; ++p) {
            const int32_t li = s->prim_light[p] >= 0 ? s->prim_light[p] + (int32_t)light_base[u] : -1;
            pinfo[global(p)] = bg::prim_info_of(li, s->prim_material[p] + (uint32_t)mat_base[u], p < nt && s->tri_flip[p] != 0);
        }
    }
    uint64_t bytes = 0;
    int rc = BG_OK;
    auto up = [&](auto** dst, const auto* src, size_t count) {
        if (rc || count == 0) return;
        rc = dalloc(ctx, dst, count);
        if (!rc && hipMemcpy(*dst, src, count * sizeof(**dst), hipMemcpyHostToDevice) != hipSuccess)
            rc = fail(ctx, BG_ERR_HIP, "batch upload failed");
        bytes += count * sizeof(**dst);
    };
    up((bg::PairRecord**)&bt->pairs, pairs.data(), pairs.size());
    up((uint32_t**)&bt->leaves, leaves.data(), leaves.size());
    up(&bt->slot_prim, slots.data(), slots.size());
    up(&bt->prim_info, pinfo.data(), pinfo.size());
    // device-to-device copies of the per-scene arrays that need no rebasing
    if (!rc) rc = dalloc(ctx, &bt->tri_verts, 9 * bt->n_slots);
    if (!rc) rc = dalloc(ctx, &bt->tri_orig, 9 * bt->n_tris);
    if (!rc) rc = dalloc(ctx, &bt->spheres, bt->n_spheres);
    if (!rc) rc = dalloc(ctx, &bt->sphere_aux, bt->n_spheres);
    if (!rc) rc = dalloc(ctx, &bt->sph_shade, bt->n_spheres);
    if (!rc) rc = dalloc(ctx, &bt->mats, n_mats);
    if (!rc) rc = dalloc(ctx, &bt->lights, bt->n_lights);
    for (size_t u = 0; u < U && !rc; ++u) {
        const bg_scene* s = uniq[u];
        const size_t ns = s->bvh.slot_prim.size();
        auto d2d = [&](void* dst, const void* src, size_t b) {
            if (!rc && b && hipMemcpy(dst, src, b, hipMemcpyDeviceToDevice) != hipSuccess)
                rc = fail(ctx, BG_ERR_HIP, "batch device copy failed");
        };
        d2d(bt->tri_verts + 9 * slot_base[u], s->d_tri_verts, 9 * ns * sizeof(float));
        d2d(bt->tri_orig + 9 * tri_base[u], s->d_tri_orig, 9 * s->n_tris * sizeof(float));
        d2d(bt->spheres + sph_base[u], s->d_spheres, s->n_spheres * sizeof(bg::SphereDev));
        d2d(bt->sphere_aux + sph_base[u], s->d_sphere_aux, s->n_spheres * sizeof(bg::SphereAux));
        d2d(bt->sph_shade + sph_base[u], s->d_sph_shade, s->n_spheres * sizeof(bg::SphereShadeDev));
        d2d(bt->mats + mat_base[u], s->d_mats, s->materials.size() * sizeof(bg::MaterialDev));
        if (!s->lights.empty() && !rc) {
            std::vector<bg::LightDev> l(s->lights.size());
            if (hipMemcpy(l.data(), s->d_lights, l.size() * sizeof(bg::LightDev), hipMemcpyDeviceToHost) != hipSuccess)
                rc = fail(ctx, BG_ERR_HIP, "batch light copy failed");
            const uint64_t ntu = s->n_tris;
            for (size_t k = 0; k < l.size(); ++k) {
                // batch-wide primitive ids: every scene's triangles first, then every scene's spheres
                const uint64_t pr = l[k].prim;
                l[k].prim = (uint32_t)(pr < ntu ? tri_base[u] + pr : bt->n_tris + sph_base[u] + (pr - ntu));
                lts[light_base[u] + k] = l[k];
            }
        }
    }
    if (!rc && !lts.empty() && hipMemcpy(bt->lights, lts.data(), lts.size() * sizeof(bg::LightDev), hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(ctx, BG_ERR_HIP, "batch light upload failed");
    if (!rc) rc = dalloc(ctx, &bt->d_items, n);
    if (!rc) rc = dalloc(ctx, &bt->film, bt->total_pix);
    if (!rc) rc = dalloc(ctx, &bt->d_pass, 1);
    if (!rc) rc = dalloc(ctx, &bt->d_queue, 32);  // one counter per ray set and the persistent kernel's, 64 bytes apart
    if (rc) {
        free_batch(bt);
        delete bt;
        return rc;
    }
    bt->shape = ctx->shape;
    *out = bt;
    return BG_OK;
}

void bg_batch_destroy(bg_batch* bt) {
    if (!bt) return;
    free_batch(bt);
    delete bt;
}

int bg_batch_film_read(bg_batch* bt, uint32_t item, float* rgbw) {
    if (!bt || !rgbw || item >= bt->items.size()) return fail(bt ? bt->ctx : nullptr, BG_ERR_INVALID, "bad argument");
    bg_ctx* ctx = bt->ctx;
    const uint64_t np = (uint64_t)bt->items[item].camera.width * bt->items[item].camera.height;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    HIP_TRY(ctx, hipMemcpy(rgbw, bt->film + bt->item_pix_base[item], np * sizeof(float4), hipMemcpyDeviceToHost));
    return BG_OK;
}

int bg_batch_launch_config(const bg_batch* bt, bg_launch_config* cfg) {
    if (!bt || !cfg) return BG_ERR_INVALID;
    cfg->block = (uint32_t)bt->shape.block;
    cfg->stack = (uint32_t)bt->shape.stack;
    cfg->blocks_per_cu = (uint32_t)bt->shape.blocks_per_cu;
    cfg->variant = (uint32_t)bt->shape.variant;
    cfg->compact = bt->shape.compact ? 1u : 0u;
    return BG_OK;
}

int bg_render_batch(bg_ctx* ctx, bg_batch* bt, const bg_batch_desc* d, bg_stats* stats) {
    if (!ctx || !bt || !d || bt->ctx != ctx) return fail(ctx, BG_ERR_INVALID, "null or foreign argument");
    if (d->mode > BG_MODE_CORRECTED) return fail(ctx, BG_ERR_INVALID, "unknown mode");
    if (d->schedule > BG_SCHEDULE_PERSISTENT) return fail(ctx, BG_ERR_INVALID, "unknown schedule");
    const bool persistent = d->schedule == BG_SCHEDULE_PERSISTENT;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (stats) *stats = bg_stats{};
    const uint32_t n = (uint32_t)bt->items.size();
    HIP_TRY(ctx, hipMemsetAsync(bt->film, 0, bt->total_pix * sizeof(float4), ctx->stream));
    if (d->spp == 0) return BG_OK;
    // samples per pass: a divisor of spp so that every replay of the captured pass is identical.
    // Default about 256M paths (264 bytes each: up to 68 GB of the 288 GB); the 64-scene,
    // 28.3M-pixel, 16-spp batch at depth 6 measured 0.366 s in passes of 1 spp, 0.348 s of 2,
    // 0.337 s of 4 and 0.315 s of 8: larger passes keep the deep bounces' launches full.
    const uint64_t target = d->paths_per_pass ? d->paths_per_pass : (uint64_t)256 << 20;
    uint32_t spp_pass = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(d->spp, target / bt->total_pix));
    while (d->spp % spp_pass) --spp_pass;
    const uint64_t cap = bt->total_pix * spp_pass;
    int rc = bg::ensure_paths(ctx, cap);
    if (rc) return rc;
    if (bt->root_cap < cap) {
        if (bt->d_root) hipFree(bt->d_root);
        bt->d_root = nullptr;
        bt->root_cap = 0;
        HIP_TRY(ctx, hipMalloc((void**)&bt->d_root, cap * sizeof(uint32_t)));
        bt->root_cap = cap;
    }
    std::vector<bg::BatchItemDev> it(n);
    for (uint32_t k = 0; k < n; ++k) {
        bg::camera_setup(bt->items[k].camera, &it[k].cam);
        it[k].seed = bt->items[k].seed;
        it[k].path_base = bt->item_pix_base[k] * spp_pass;
        it[k].pix_base = bt->item_pix_base[k];
        const uint32_t u = bt->item_scene[k];
        it[k].root = bt->scene_root[u];
        it[k].light_base = bt->scene_light_base[u];
        it[k].n_lights = bt->scene_n_lights[u];
        it[k].pad = 0;
    }
    hipStream_t st = ctx->stream;
    HIP_TRY(ctx, hipStreamSynchronize(st));
    HIP_TRY(ctx, hipMemcpy(bt->d_items, it.data(), n * sizeof(bg::BatchItemDev), hipMemcpyHostToDevice));

    bg::PathBuffers b = ctx->paths;
    b.root = bt->d_root;
    bg::RenderParamsDev p;
    p.seed = 0;
    p.x0 = p.y0 = p.x1 = p.y1 = 0;
    p.first_sample = 0;
    p.n_samples = spp_pass;
    p.max_depth = d->max_depth;
    p.rr = d->rr_threshold;
    p.mode = d->mode;
    p.n_paths = cap;
    p.batch = bt->d_items;
    p.n_batch = n;
    p.pass_first = bt->d_pass;
    std::memset(&p.cam, 0, sizeof(p.cam));
    bg::ShadeSceneDev sc;
    sc.tri_orig = bt->tri_orig;
    sc.prim_info = bt->prim_info;
    sc.mats = bt->mats;
    sc.lights = bt->lights;
    sc.spheres = bt->sph_shade;
    sc.n_tris = (uint32_t)bt->n_tris;
    sc.n_lights = bt->n_lights;
    sc.multi_lobe = bt->multi_lobe;
    sc.sphere_lights = bt->sphere_lights;

    // The autotuner times two stages of candidates: the traversal variant and compaction at two
    // shapes (16), then the shape (block x stack x blocks per CU, 16) with each of the best two
    // (variant, compaction) pairs (32): 48 candidates instead of the 128 of the full product, each
    // timed on all the traversal launches of one pass.
    std::vector<bg::LaunchShape> cands;
    const int kBlocks[] = {128, 256}, kStacks[] = {8, 16}, kBpc[] = {8, 16, 32, 64};
    const bool tuning = d->autotune && !bt->tuned && !persistent;  // (the persistent kernel has one shape)
    if (tuning)
        for (int block : kBlocks)
            for (int bpc : kBpc) cands.push_back(bg::LaunchShape{block, 16, bpc, 0, 1});
    // spill rows for the largest grid any launch shape of this call uses
    uint64_t max_threads = (uint64_t)grid_for(ctx, bt->shape, cap) * bt->shape.block;
    for (const bg::LaunchShape& c : cands) max_threads = std::max<uint64_t>(max_threads, (uint64_t)grid_for(ctx, c, cap) * c.block);
    // the persistent path kernel: as many 256-thread blocks as stay resident
    const int pgrid = persistent ? std::max(1, ctx->num_cus) * bg::path_persistent_blocks_per_cu(sc, bt->spheres != nullptr) : 0;
    max_threads = std::max<uint64_t>(max_threads, (uint64_t)pgrid * 256);
    // (three sets of spill rows: a pass traces its ray sets at once)
    const size_t per_set = (((size_t)max_threads * (bt->max_depth + 1) * bg::kSpillWords * sizeof(uint32_t) + 255) / 256) * 256;
    rc = bg::ensure_spill(ctx, 3 * per_set);
    if (rc) return rc;
    ctx->spill_set_bytes = per_set;
    rc = bg::ensure_side_streams(ctx);
    if (rc) return rc;

    if (tuning) {
        // time every candidate of a stage on every traversal launch of one pass (the extension,
        // shadow and MIS-branch rays of each bounce, on the rays the shading kernels wrote: every
        // candidate gives the same hits); the candidate of least total time wins
        auto t0 = std::chrono::steady_clock::now();
        // the candidates of `cs` ordered by their total time, fastest first
        auto tune = [&](const std::vector<bg::LaunchShape>& cs, std::vector<bg::LaunchShape>* ranked) -> int {
            HIP_TRY(ctx, hipMemsetAsync(bt->d_pass, 0, sizeof(uint32_t), st));
            HIP_TRY(ctx, bg::launch_init_paths(p, b, st));
            std::vector<float> score(cs.size(), 0.0f);
            auto time_all = [&](const float4* rays, uint4* hits, bool any, const uint64_t* mask) -> int {
                for (size_t c = 0; c < cs.size(); ++c) {
                    HIP_TRY(ctx, hipEventRecord(ctx->ev0, st));
                    HIP_TRY(ctx, traverse(bt, ctx, cs[c], rays, hits, cap, b.root, any, mask, st));
                    HIP_TRY(ctx, hipEventRecord(ctx->ev1, st));
                    HIP_TRY(ctx, hipEventSynchronize(ctx->ev1));
                    float ms = 0.0f;
                    HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
                    score[c] += ms;
                }
                return BG_OK;
            };
            for (uint32_t bounce = 0; bounce <= d->max_depth; ++bounce) {
                int r = time_all(b.ray, b.hit, false, b.wmask[0]);
                if (r) return r;
                HIP_TRY(ctx, bg::launch_shade(p, sc, b, bounce, st));
                if (sc.n_lights > 0 && bounce < d->max_depth) {
                    if ((r = time_all(b.sray, b.shit, true, b.wmask[1])) != BG_OK) return r;
                    if ((r = time_all(b.mray, b.mhit, false, b.wmask[2])) != BG_OK) return r;
                }
            }
            std::vector<size_t> order(cs.size());
            for (size_t c = 0; c < cs.size(); ++c) order[c] = c;
            std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return score[x] < score[y]; });
            ranked->clear();
            for (size_t c : order) ranked->push_back(cs[c]);
            return BG_OK;
        };
        std::vector<bg::LaunchShape> stage, ranked;
        for (int variant : {0, 2, 4, 5})
            for (int compact : {0, 1}) {
                stage.push_back(bg::LaunchShape{256, 8, 32, variant, compact});
                stage.push_back(bg::LaunchShape{128, 8, 64, variant, compact});
            }
        rc = tune(stage, &ranked);
        if (rc) return rc;
        std::vector<bg::LaunchShape> pairs;  // the (variant, compaction) pairs by their best time
        for (const bg::LaunchShape& c : ranked) {
            bool seen = false;
            for (const bg::LaunchShape& q : pairs) seen = seen || (q.variant == c.variant && q.compact == c.compact);
            if (!seen) pairs.push_back(c);
        }
        stage.clear();
        for (size_t k = 0; k < 2 && k < pairs.size(); ++k)
            for (int block : kBlocks)
                for (int stack : kStacks)
                    for (int bpc : kBpc)
                        stage.push_back(bg::LaunchShape{block, stack, bpc, pairs[k].variant, pairs[k].compact});
        rc = tune(stage, &ranked);
        if (rc) return rc;
        bt->shape = ranked[0];
        bt->tuned = true;
        bt->tune_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    const bg::LaunchShape sh = bt->shape;

    HIP_TRY(ctx, hipMemsetAsync(b.counters, 0, 2 * sizeof(unsigned long long), st));
    HIP_TRY(ctx, hipMemsetAsync(bt->d_pass, 0, sizeof(uint32_t), st));
    uint64_t launches = 0;
    auto enqueue_pass = [&](hipStream_t s) -> hipError_t {
        hipError_t e;
        if (persistent) {
            // one launch runs the pass's paths; its three ray sets share set 0's spill rows (each
            // thread traces them one after another)
            const bg::TraverseArgs ext = batch_args(bt, ctx, b.ray, b.hit, cap, b.root, false, 0);
            const bg::TraverseArgs sha = batch_args(bt, ctx, b.sray, b.shit, cap, b.root, false, 0);
            const bg::TraverseArgs mis = batch_args(bt, ctx, b.mray, b.mhit, cap, b.root, false, 0);
            if ((e = bg::launch_path_persistent(p, sc, b, ext, sha, mis, bt->d_queue + 24, pgrid, s)) != hipSuccess)
                return e;
            launches += 2;
            if ((e = bg::launch_batch_film(p, b, bt->total_pix, bt->film, s)) != hipSuccess) return e;
            if ((e = bg::launch_count_rays(p, b, b.counters, s)) != hipSuccess) return e;
            if ((e = bg::launch_pass_advance(bt->d_pass, spp_pass, s)) != hipSuccess) return e;
            launches += 3;
            return hipSuccess;
        }
        if ((e = bg::launch_init_paths(p, b, s)) != hipSuccess) return e;
        ++launches;
        if ((e = traverse(bt, ctx, sh, b.ray, b.hit, cap, b.root, false, b.wmask[0], s, true, 0)) != hipSuccess) return e;
        ++launches;
        for (uint32_t bounce = 0; bounce <= d->max_depth; ++bounce) {
            if ((e = bg::launch_shade(p, sc, b, bounce, s)) != hipSuccess) return e;
            ++launches;
            // the shadow and MIS-branch rays on the side streams, beside the next extension rays
            // (as in path renders; captured as a fork and a join of the graph)
            const bool nee = sc.n_lights > 0 && bounce < d->max_depth;
            if (nee) {
                if ((e = hipEventRecord(ctx->fork_ev, s)) != hipSuccess) return e;
                for (int k = 0; k < 2; ++k) {
                    hipStream_t side = ctx->side[k];
                    if ((e = hipStreamWaitEvent(side, ctx->fork_ev, 0)) != hipSuccess) return e;
                    if ((e = traverse(bt, ctx, sh, k ? b.mray : b.sray, k ? b.mhit : b.shit, cap, b.root, k == 0,
                                      b.wmask[1 + k], side, true, 1 + k)) != hipSuccess)
                        return e;
                    if ((e = hipEventRecord(ctx->join_ev[k], side)) != hipSuccess) return e;
                }
            }
            if (bounce < d->max_depth) {
                if ((e = traverse(bt, ctx, sh, b.ray, b.hit, cap, b.root, false, b.wmask[0], s, true, 0)) != hipSuccess) return e;
                ++launches;
            }
            if (nee) {
                for (int k = 0; k < 2; ++k)
                    if ((e = hipStreamWaitEvent(s, ctx->join_ev[k], 0)) != hipSuccess) return e;
                if ((e = bg::launch_nee_resolve(p, sc, b, bounce, s)) != hipSuccess) return e;
                launches += 3;
            }
        }
        if ((e = bg::launch_batch_film(p, b, bt->total_pix, bt->film, s)) != hipSuccess) return e;
        if ((e = bg::launch_count_rays(p, b, b.counters, s)) != hipSuccess) return e;
        if ((e = bg::launch_pass_advance(bt->d_pass, spp_pass, s)) != hipSuccess) return e;
        launches += 3;
        return hipSuccess;
    };
    const uint32_t passes = d->spp / spp_pass;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // a measured render (bg_set_profiling) runs its passes as direct launches, the event pairs
    // around its traversal launches outside any graph
    rc = bg::profile_begin(ctx);
    if (rc) return rc;
    auto t0 = std::chrono::steady_clock::now();
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, st));
    if (d->use_graph && !ctx->profile) {
        HIP_TRY(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
        const hipError_t e = enqueue_pass(st);
        hipGraph_t g = nullptr;
        const hipError_t e2 = hipStreamEndCapture(st, &g);
        graph = g;
        if (e != hipSuccess || e2 != hipSuccess) {
            if (graph) hipGraphDestroy(graph);
            return fail(ctx, BG_ERR_HIP, std::string("pass capture failed: ") + hipGetErrorString(e != hipSuccess ? e : e2));
        }
        const hipError_t e3 = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (e3 != hipSuccess) {
            hipGraphDestroy(graph);
            return fail(ctx, BG_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e3));
        }
        for (uint32_t k = 0; k < passes; ++k) {
            const hipError_t e4 = hipGraphLaunch(exec, st);
            if (e4 != hipSuccess) {
                hipGraphExecDestroy(exec);
                hipGraphDestroy(graph);
                return fail(ctx, BG_ERR_HIP, std::string("hipGraphLaunch: ") + hipGetErrorString(e4));
            }
        }
        launches *= passes;  // counted once while capturing; the graph replays them
    } else {
        for (uint32_t k = 0; k < passes; ++k) HIP_TRY(ctx, enqueue_pass(st));
    }
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, st));
    HIP_TRY(ctx, hipEventSynchronize(ctx->ev1));
    auto t1 = std::chrono::steady_clock::now();
    if (exec) hipGraphExecDestroy(exec);
    if (graph) hipGraphDestroy(graph);
    float ms = 0.0f;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    unsigned long long nr[2] = {0, 0};
    HIP_TRY(ctx, hipMemcpy(nr, b.counters, sizeof(nr), hipMemcpyDeviceToHost));
    if (stats) {
        stats->kernel_ms = ms;
        stats->total_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        stats->launches = launches;
        stats->samples = bt->total_pix * d->spp;
        stats->rays = nr[0];
        stats->shadow_rays = nr[1];
    }
    return bg::profile_end(ctx, stats);
}

}  // extern "C"
