// Internal declarations shared by the library's host sources.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/bonsai_gpu.h"

namespace bg {

// Host-side flattened BVH produced by the builder (bvh_build.cpp).
struct HostBVH {
    std::vector<bg_bvh_node> nodes;
    std::vector<uint32_t> slot_prim;  // leaf slot -> primitive id
    uint32_t max_depth = 0;
    uint32_t max_leaf = 0;
};

// Traversal layout with the boxes of both children in their parent (64 bytes,
// four 16-byte loads): b[0..5] child 0 (lo, hi), b[6..11] child 1, then the two
// child references and the split axis (3: no split, child 0 first). A
// reference is a record index, or kRefLeaf | leaf id into the leaf table
// (first slot, count << 1 | has_sphere), or kRefNone. Record 0 of a scene is a
// virtual root whose child 0 is the BVH root (its box is the root's box).
struct PairRecord {
    float b[12];
    uint32_t ref0, ref1, axis, pad;
};
constexpr uint32_t kRefLeaf = 0x80000000u, kRefNone = 0xFFFFFFFFu;
// build_pairs numbers the first kTopRecords records of a scene breadth-first from the virtual root
// (record 0, the BVH root's record 1, its children's records 2 and 3, ...; the rest follow in
// depth-first order), so the traversal kernels can hold a prefix of the record array, the top
// levels every ray passes through, in LDS.
constexpr uint32_t kTopRecords = 64;
// Words of one traversal stack entry (reference, entry distance) in the spill buffer.
constexpr size_t kSpillWords = 2;

// Per-primitive world bounds, in scene order (triangles then spheres).
struct PrimBounds {
    float mn[3], mx[3];
};

// Builds the SAH BVH exactly as bvh.rs:91-269 describes, with the bucket
// index computed as floor(16 * offset) and zero-initialised buckets.
// `threads` > 1 builds independent subtrees concurrently; the output does not
// depend on it.
void build_bvh(const std::vector<PrimBounds>& bounds, const std::vector<uint8_t>& is_sphere, int threads,
               HostBVH* out);
// The pair-record layout of a built BVH (same tree, same DFS order).
void build_pairs(const HostBVH& bvh, std::vector<PairRecord>* pairs, std::vector<uint32_t>* leaves);

// Sphere record in HBM: world_to_object rows 0-2 (the transform is affine),
// radius, z_min, z_max, phi_max (radians). 64 bytes.
struct SphereDev {
    float w2o[12];
    float radius, z_min, z_max, phi_max;
};
// Sphere data used only when an accepted hit's surface (u, v) is formed.
struct SphereAux {
    float theta_min, theta_max;
};

// Launch parameters of the traversal kernels.
struct TraverseArgs {
    const float4* nodes;       // 2 x float4 per node
    const float* tri_verts;    // 9 floats per slot
    const uint32_t* slot_prim; // slot -> primitive id
    const SphereDev* spheres;
    const SphereAux* sphere_aux;
    uint32_t n_tris;
    const float4* rays;        // 2 x float4 per ray
    uint4* hits;               // 1 x uint4 per ray
    uint64_t n_rays;
    uint32_t* spill;           // per-thread overflow of the LDS stack
    uint32_t spill_depth;
    unsigned long long* counters;  // [nodes, tris, spheres] when counting
    const float4* pairs = nullptr;     // PairRecord array (4 float4 each)
    const uint2* leaves = nullptr;     // leaf table of the pair layout
    const uint32_t* roots = nullptr;   // per ray: root record (scene batches); record/node 0 when null
    unsigned long long* queue = nullptr;  // work-queue counter (variants 2 and 5; zeroed by the launch)
    uint32_t chunk_shift = 9;             // variant 4: log2 of the rays per chunk (set at launch)
    uint32_t refill_min = 16;             // variant 4: finished lanes that trigger a refill
    // Compacted launches: logical ray j is ray idx[j] (read from rays, written to hits), for
    // j < *n_idx (a count on the device); n_rays is then only the upper bound for the grid.
    const uint32_t* idx = nullptr;
    const uint32_t* n_idx = nullptr;
    // records 0 .. n_top - 1 are also read from a copy in LDS (single-scene launches; 0 for none)
    uint32_t n_top = 0;
};

// Default = the best of a sweep over block {64,128,256} x stack {8,16} x
// blocks/CU {8,16,32,64,0} x variant (and, for variant 4, refill_min
// {8,16,24,32,48}) on the 2^24-ray intersection workload.
struct LaunchShape {
    int block = 256;
    int stack = 8;
    int blocks_per_cu = 32;
    // 0: pair records, one ray per loop round; 1: node records, per-lane refill; 2: as 0 with rays
    // taken from an atomic work queue; 3: node records, one ray per loop round; 4: pair records,
    // batched refill of finished lanes from per-wavefront ray sequences; 5: pair records, batched
    // refill from runs of rays taken from an atomic work queue (persistent; index lists and spheres)
    int variant = 4;
    // render passes: trace only the rays that can hit (t_max > 0), compacted into an index list
    // (the others' hits are written as misses), instead of launching over every path slot
    int compact = 1;
};

// traverse.hip
hipError_t launch_traverse(const TraverseArgs& a, bool any_hit, bool count, const LaunchShape& shape, int grid,
                           hipStream_t stream);
bool traverse_shape_supported(const LaunchShape& shape);

// Camera parameters precomputed on the host (perspective.rs:49-114).
struct CameraDev {
    float raster_to_camera[16];
    float camera_to_world[16];
    float lens_radius, focal_distance, shutter_open, shutter_close;
    uint32_t width, height;
};
void camera_setup(const bg_camera_desc& d, CameraDev* out);
hipError_t launch_camera_rays(const CameraDev& cam, float fx, float fy, float4* rays, hipStream_t stream);

// Shading records in HBM.
struct MaterialDev {
    uint32_t type;
    float r[3];
    float alpha;
    float eta[3];
    float k[3];
    uint32_t pad;
};
struct LightDev {
    uint32_t prim;
    float L[3];
    uint32_t double_sided;
    float area;
    uint32_t pad[2];
};
struct SphereShadeDev {
    float o2w[12];
    float w2o[12];
    float radius, z_min, z_max, phi_max, theta_min, theta_max;
    uint32_t flip;     // the normal's flip: reverse_orientation xor the transform swaps handedness
    uint32_t reverse;  // reverse_orientation alone (Sphere::sample, sphere.rs:333-335)
};
// ShadeSceneDev::prim_info: the orientation flag in the material word, and the record of a primitive.
constexpr uint32_t kPrimFlip = 0x80000000u;
inline uint2 prim_info_of(int32_t light, uint32_t material, bool flip) {
    uint2 r;
    r.x = (uint32_t)light;
    r.y = material | (flip ? kPrimFlip : 0u);
    return r;
}
struct ShadeSceneDev {
    const float* tri_orig;      // 9 floats per triangle, scene order
    // per primitive, scene order: x its light index (as int32, -1: none), y its material index, with
    // kPrimFlip set for a triangle of reversed orientation; one 8-byte record, so shading a hit reads
    // one scattered line for these where three arrays took three
    const uint2* prim_info;
    const MaterialDev* mats;
    const LightDev* lights;
    const SphereShadeDev* spheres;
    uint32_t n_tris, n_lights;
    uint32_t multi_lobe = 0;  // a material with two lobes (plastic) is present
    uint32_t sphere_lights = 0;  // an area light is on a sphere (selects the kernels that sample spheres)
};

// Wavefront path state (structure of arrays) for up to `cap` paths.
struct PathBuffers {
    size_t cap = 0;
    void* mem = nullptr;
    float4* ray = nullptr;     // 2 per path: (o, t_max), (d, time)
    uint4* hit = nullptr;
    float4* beta = nullptr;    // rgb, bounce (as float bits unused)
    float4* L = nullptr;       // rgb
    uint2* key = nullptr;      // sample-stream key (lo, hi)
    uint2* state = nullptr;    // x: draw index, y: flags | bounce << 8
    float2* film = nullptr;    // film point
    float4* sray = nullptr;    // shadow ray, 2 per path
    uint4* shit = nullptr;
    float4* mray = nullptr;    // MIS-branch ray, 2 per path
    uint4* mhit = nullptr;
    // a vertex's next-event estimate, already scaled by the throughput at the vertex and the light
    // pick (beta / light pdf): the light branch's radiance to add when its shadow ray is unoccluded
    // (rgb, flags), and the BSDF branch's factor of the emission its ray reaches (rgb, light index)
    float4* nee_a = nullptr;
    float4* nee_b = nullptr;
    unsigned long long* counters = nullptr;  // [queries, shadow queries]
    // path guiding (allocated only for guided renders)
    size_t guide_cap = 0, rec_cap = 0;
    void* guide_mem = nullptr;
    float4* gin = nullptr;     // 3 per live path (compacted): network input (p, wo, n) at the current hit
    float* gout = nullptr;     // 40 per live path (compacted): network output
    // compacted launches, one set per ray set (extension, shadow, MIS branch) so that the three
    // can run at once: indices of the rays that can hit, per 256 rays (or 4096) their count then
    // its exclusive prefix, per 64 rays which can hit (k_ray_count), their number (device)
    uint32_t* ridx[3] = {nullptr, nullptr, nullptr};
    uint32_t* rblk[3] = {nullptr, nullptr, nullptr};
    uint64_t* rmask[3] = {nullptr, nullptr, nullptr};
    uint32_t* rcount[3] = {nullptr, nullptr, nullptr};
    // the same masks for the extension, shadow and MIS-branch rays (ray, sray, mray), stored by the
    // kernel that wrote the rays (k_init_paths, k_shade) so a compacted launch need not read them
    uint64_t* wmask[3] = {nullptr, nullptr, nullptr};
    uint32_t* gslot = nullptr; // per path: its slot in gin/gout this bounce, ~0 when not live
    uint32_t* gcount = nullptr; // live-path count of this bounce (device)
    uint32_t* gblk = nullptr;   // per 256 paths: live count, then its exclusive prefix
    uint32_t* root = nullptr;   // per path: BVH root of its scene (scene batches)
    float4* rec_in = nullptr;  // 3 per training record: the input the direction was sampled with
    float4* rec_tgt = nullptr; // per record: sampled direction, then (after the pass) its weight in .w
    float4* rec_aux = nullptr; // 2 per record: throughput before scattering, radiance after the vertex
    uint32_t* rec_blk = nullptr; // per 256 records: count of nonzero weights, then its exclusive prefix
    uint64_t last_records = 0; // records of the last trained pass (their weights are final)
};

// Light slots of the direct-lighting and Whitted integrators, slot-major (slot
// j of path i at j * n_paths + i); one allocation.
struct DirectSlots {
    size_t cap = 0;
    void* mem = nullptr;
    float4* sray = nullptr;  // shadow ray, 2 per slot
    uint4* shit = nullptr;
    float4* mray = nullptr;  // BSDF-sampling (MIS) ray, 2 per slot
    uint4* mhit = nullptr;
    float4* a = nullptr;     // light-branch term rgb, flags
    float4* b = nullptr;     // BSDF-branch f rgb, MIS weight
    float2* c = nullptr;     // BSDF-branch pdf, light index
};

// Replay ring of training samples with nonzero weight, kept by the guide
// across passes; state = [head, count, base of the last append, last append size].
struct Replay {
    float4* in = nullptr;   // 3 per sample
    float4* tgt = nullptr;  // direction, weight
    uint32_t* state = nullptr;
    uint32_t cap = 0;
};

// One scene of a batch rendered in a single wavefront loop.
struct BatchItemDev {
    CameraDev cam;        // its camera; its film is cam.width x cam.height
    uint64_t seed;        // sample-stream seed
    uint64_t path_base;   // first path of this item in a pass
    uint64_t pix_base;    // first pixel of its film in the batch film
    uint32_t root;        // its BVH root in the batch node array
    uint32_t light_base, n_lights;
    uint32_t pad;
};

struct RenderParamsDev {
    CameraDev cam;
    uint64_t seed;
    uint32_t x0, y0, x1, y1;  // window
    uint32_t first_sample, n_samples;  // samples of this pass
    uint32_t max_depth;
    float rr;
    uint32_t mode;
    uint64_t n_paths;
    // path guiding: one-sample MIS between the network's mixture (probability guide_alpha) and
    // the BSDF; every rec_stride-th path (up to rec_paths of them) records its vertices
    uint32_t guided = 0;
    float guide_alpha = 0.0f;
    uint32_t guide_lobes = 3;  // bit 0: guide at diffuse vertices, bit 1: at glossy vertices
    uint32_t guide_bounces = 0xFFFFFFFFu;  // vertices (bounces) below this are guided
    uint32_t rr_start = 4;                 // Russian roulette from this bounce on (path.rs:133: bounces > 3)
    uint32_t rec_stride = 1, rec_paths = 0;
    // scene batches: items sorted by path_base (and pix_base); the first sample of the pass
    // is read from pass_first (device) so a captured pass can be replayed
    const BatchItemDev* batch = nullptr;
    uint32_t n_batch = 0;
    const uint32_t* pass_first = nullptr;
};

// path.hip
hipError_t launch_init_paths(const RenderParamsDev& p, const PathBuffers& b, hipStream_t s);
hipError_t launch_shade(const RenderParamsDev& p, const ShadeSceneDev& sc, const PathBuffers& b, uint32_t bounce,
                        hipStream_t s);
hipError_t launch_nee_resolve(const RenderParamsDev& p, const ShadeSceneDev& sc, const PathBuffers& b,
                              uint32_t bounce, hipStream_t s);
hipError_t launch_count_rays(const RenderParamsDev& p, const PathBuffers& b, unsigned long long* counters,
                             hipStream_t s);
// A render's traversal launch with the context's measurement (bg_set_profiling): test counting
// into ctx->counters, and a HIP event pair around the kernel; profile_begin / profile_end open a
// render call's measurement and write it into its stats.
hipError_t profiled_traverse(bg_ctx* ctx, TraverseArgs a, bool any, const LaunchShape& shape, int grid,
                             hipStream_t st);
int profile_begin(bg_ctx* ctx);
int profile_end(bg_ctx* ctx, bg_stats* stats);
// Whether a render pass compacts this launch's rays first: the shape asks for it and the launch
// covers at least 1M ray slots.
bool traverse_compacts(const TraverseArgs& a, const LaunchShape& shape, int grid);
// Writes idx (ascending indices of the rays with t_max > 0), their number to *count and misses
// for the other rays; blk holds ceil(n / 256) words and mask 4 * ceil(n / 256). With `produced`
// (the masks the kernel that wrote the rays stored, one word per 64 rays, the first ceil(n / 64)
// of them read) the rays are not read at all and no misses are written (the path kernels read
// only hits of rays that can hit); otherwise their t_max values make the masks in mask.
hipError_t launch_ray_compact(const float4* rays, uint64_t n, uint32_t* blk, uint64_t* mask, uint32_t* idx,
                              uint32_t* count, uint4* hits, hipStream_t s, const uint64_t* produced = nullptr);
hipError_t launch_guide_inputs(const RenderParamsDev& p, const ShadeSceneDev& sc, const PathBuffers& b, hipStream_t s);
hipError_t launch_guide_targets(const RenderParamsDev& p, const PathBuffers& b, hipStream_t s);
// Appends the pass's nonzero-weight records to the ring (order-preserving,
// deterministic: block counts, one-block exclusive scan, scatter).
hipError_t launch_replay_append(const PathBuffers& b, uint64_t n_records, const Replay& r, hipStream_t s);
// Minibatch k = 0..n-1 takes ring sample ((first + k) * 2654435761) mod count:
// a stride-permutation of the ring (the prime exceeds any count).
hipError_t launch_replay_gather(const Replay& r, uint64_t first, uint64_t n, float4* in, float4* tgt, hipStream_t s);
hipError_t launch_batch_film(const RenderParamsDev& p, const PathBuffers& b, uint64_t total_pix, float4* film,
                             hipStream_t s);
// Adds `spp` to *pass_first (the last node of a captured pass).
hipError_t launch_pass_advance(uint32_t* pass_first, uint32_t spp, hipStream_t s);
hipError_t launch_zero_u64(unsigned long long* p, hipStream_t s);
// A plain pass as one persistent launch (k_path_persistent): zeroes *queue, then `grid` blocks of
// 256 threads run whole paths, each lane taking a new path slot (from runs of 64 its wavefront
// claims off the counter) when its path ends. The three argument sets are the
// extension, shadow and MIS-branch queries (their spill rows: ext's, one row per thread).
hipError_t launch_path_persistent(const RenderParamsDev& p, const ShadeSceneDev& sc, const PathBuffers& b,
                                  const TraverseArgs& ext, const TraverseArgs& sha, const TraverseArgs& mis,
                                  unsigned long long* queue, int grid, hipStream_t s);
// Resident 256-thread blocks per CU of the persistent path kernel for this scene's instantiation.
int path_persistent_blocks_per_cu(const ShadeSceneDev& sc, bool spheres);
hipError_t launch_film(const RenderParamsDev& p, const PathBuffers& b, uint32_t width, uint32_t height, float4* film,
                       double4* film64, hipStream_t s);

// noalloc.hip: the static scratch of bg_intersect_device_noalloc. init takes the
// array's address on `device` (which must be current) once; scratch returns it
// (nullptr before init).
hipError_t noalloc_scratch_init(int device);
void* noalloc_scratch(int device);
hipError_t noalloc_zero_counters(unsigned long long* c, uint32_t n, hipStream_t s);
hipError_t noalloc_fill_miss(void* hits, uint64_t n, hipStream_t s);

// render.cpp: (re)allocate the context's path state and traversal spill buffer
int ensure_paths(bg_ctx* ctx, size_t cap);
int ensure_spill(bg_ctx* ctx, size_t bytes);
int ensure_side_streams(bg_ctx* ctx);
// render.cpp: the guide's replay ring of `cap` samples (emptied when the capacity changes)
int guide_ensure_replay(bg_guide* g, uint32_t cap);

// ---------------------------------------------------------------- guiding MLP
constexpr int kParamsTotal = 4 * (64 * 64 + 64) + (40 * 64 + 40);  // 19,240

struct GuideDevice {
    uint16_t* wbf;   // BF16 weights [5][64 out][64 in], rows >= 40 of layer 5 zero
    uint16_t* wbft;  // the same transposed [5][64 in][64 out]
    float* bias;     // [5][64]
    float* master;   // FP32 parameters in flat order (per layer W row-major, then b)
    float* m;        // Adam moments
    float* v;
};
struct GuideBox {
    float lo[3], hi[3];
};
struct GuideScratch {
    float* loss = nullptr;     // [cap]
    float* partial = nullptr;  // [blocks][5][64][64] per-block weight gradients
    float* bpartial = nullptr; // [blocks][5][64] per-block bias gradients
    uint64_t cap = 0;
    uint32_t blocks = 0;       // most blocks a training step launches
};
hipError_t guide_refresh(const GuideDevice& g, hipStream_t s);
// n_dev (optional) holds the sample count on the device; n then only bounds the grid.
// mixture: write 10-float4 mixture records (dev_guide.h) instead of the 40 raw outputs.
// Threads per block and blocks per CU of the inference kernel (the blocks fill
// every SIMD to the wave count its registers are held to): 3 blocks of 8 waves,
// 6 waves per SIMD at 80 VGPRs, LDS 3 x 44 KB. Measured on 2^22 samples:
// 0.334 ms against 0.345 ms with 2 blocks per CU (4 waves per SIMD).
constexpr int kInferBlock = 512;
constexpr int kInferBlocksPerCu = 3;
// Blocks per CU of the fused training kernel (its LDS and registers allow one).
constexpr int kTrainBlocksPerCu = 1;
hipError_t guide_infer(const GuideDevice& g, const GuideBox& box, const float4* in, uint64_t n, float* out, int grid,
                       hipStream_t s, const uint32_t* n_dev = nullptr, bool mixture = false);
hipError_t guide_train_step(const GuideDevice& g, const GuideBox& box, const float4* in, const float4* target,
                            uint64_t n, const GuideScratch& sc, float lr, int t, int grid, float* grad_out,
                            hipStream_t s);

// Direct-lighting / Whitted passes (direct.hip); kind 0: all lights, 1: one light, 2: Whitted.
hipError_t launch_direct_shade(const RenderParamsDev& p, const ShadeSceneDev& sc, const PathBuffers& b,
                               const DirectSlots& d, int kind, hipStream_t s);
hipError_t launch_direct_resolve(const RenderParamsDev& p, const ShadeSceneDev& sc, const PathBuffers& b,
                                 const DirectSlots& d, int kind, hipStream_t s);

// Film::write_image over n pixel sums (film.hip)
hipError_t launch_surface(const ShadeSceneDev& sc, const float4* rays, const uint4* hits, uint64_t n, float4* out,
                          hipStream_t s);
hipError_t launch_film_resolve(const float4* sums, uint64_t n, float splat_scale, float scale, float* rgb_out,
                               hipStream_t s);
// OpenEXR image of width x height RGB f32 pixels (exr.cpp)
int exr_write_file(const char* path, uint32_t width, uint32_t height, const float* rgb);

}  // namespace bg

struct bg_guide {
    bg_ctx* ctx = nullptr;
    bg::GuideDevice dev{};
    bg::GuideScratch scratch;
    bg::GuideBox box{};
    void* mem = nullptr;  // one allocation holding parameters, Adam state and batch buffers
    float4* batch_in = nullptr;   // minibatch buffers in `mem` (max_batch samples)
    float4* batch_tgt = nullptr;
    int step = 0;         // Adam steps taken (1-based inside the step)
    bg::Replay replay;    // training samples gathered by guided renders
    void* replay_mem = nullptr;
};

struct bg_film {
    bg_ctx* ctx = nullptr;
    uint32_t width = 0, height = 0, flags = 0;
    float4* sums = nullptr;
    double4* sums64 = nullptr;
};

struct bg_scene {
    bg_ctx* ctx = nullptr;
    bg_scene_info info{};
    bg::HostBVH bvh;
    uint64_t n_tris = 0, n_spheres = 0;
    // device
    float4* d_nodes = nullptr;
    float4* d_pairs = nullptr;
    uint64_t n_pairs = 0;
    uint2* d_leaves = nullptr;
    float* d_tri_verts = nullptr;
    uint32_t* d_slot_prim = nullptr;
    bg::SphereDev* d_spheres = nullptr;
    bg::SphereAux* d_sphere_aux = nullptr;
    // shading data
    float* d_tri_orig = nullptr;
    uint2* d_prim_info = nullptr;  // ShadeSceneDev::prim_info
    bg::MaterialDev* d_mats = nullptr;
    bg::LightDev* d_lights = nullptr;
    bg::SphereShadeDev* d_sph_shade = nullptr;
    // host copies of scene data the integrator needs
    std::vector<float> tri_positions;
    std::vector<uint8_t> tri_flip;
    std::vector<uint32_t> prim_material;
    std::vector<int32_t> prim_light;
    std::vector<bg_material> materials;
    std::vector<bg_area_light> lights;
    std::vector<float> sphere_o2w, sphere_w2o, sphere_params;
    std::vector<uint8_t> sphere_flags;
};

namespace bg {
// a material whose BSDF holds two lobes (plastic) selects the two-lobe shading kernel
inline uint32_t has_sphere_lights(const std::vector<bg_area_light>& lights, uint64_t n_tris) {
    for (const bg_area_light& l : lights)
        if (l.prim_id >= n_tris) return 1u;
    return 0u;
}
inline uint32_t has_multi_lobe(const std::vector<bg_material>& mats) {
    for (const bg_material& m : mats)
        if (m.type == BG_MAT_PLASTIC) return 1u;
    return 0u;
}
}  // namespace bg

// Scenes from this many primitives up (triangles and spheres) are rendered at the context's render
// shape, smaller ones at its query shape (measured on both sides: 500k and 1.5M-2M primitives).
constexpr uint64_t kRenderShapeMinPrims = 1ull << 20;
struct bg_ctx {
    int device = 0;
    int num_cus = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    bg::LaunchShape shape;  // bg_intersect* queries, direct-lighting and Whitted renders, small scenes
    // path and guided renders of scenes of at least kRenderShapeMinPrims primitives: measured on the 1080p path-tracer scene (64 spp, 2.05M triangles + 256
    // spheres): 0.130 s against 0.138 s with the queries' shape (256, 8, 32, variant 4), the best
    // of 72 shapes; on the door-gap room (triangles only) 0.192 s against 0.209 s (plain, 32 spp)
    bg::LaunchShape render_shape{128, 8, 16, 5, true};
    // the path and guided renders' shape for a scene: the atomic-queue refill pays on large trees
    // and loses on small ones (the 64 batch scenes of 1k-500k triangles rendered one call each:
    // 0.49 s against 0.38-0.39 s with the queries' shape; every film size lost)
    const bg::LaunchShape& render_shape_for(uint64_t n_prims) const {
        return n_prims >= kRenderShapeMinPrims ? render_shape : shape;
    }
    // reusable device buffers
    uint32_t* spill = nullptr;
    unsigned long long* counters = nullptr;
    // work-queue counters of the traversal launches (variant 5): one per ray set, 64 bytes apart
    unsigned long long* queue = nullptr;
    size_t spill_bytes = 0;
    // path renders trace a bounce's shadow and MIS-branch rays on two more streams (default
    // priority), beside the next bounce's extension rays; each ray set has its own spill rows
    // (spill_set_bytes apart) and compaction buffers
    hipStream_t side[2] = {nullptr, nullptr};
    hipEvent_t fork_ev = nullptr, join_ev[2] = {nullptr, nullptr};
    size_t spill_set_bytes = 0;
    void* io = nullptr;
    size_t io_bytes = 0;
    bg::PathBuffers paths;
    bg::DirectSlots direct;
    // render-call measurement (bg_set_profiling): flags, and the event pairs timing traversal launches
    uint32_t profile = 0;
    std::vector<hipEvent_t> prof_events;
    size_t prof_used = 0;
};
