/*
 * bonsai_gpu.h — C ABI of the MI355X (gfx950) ray-traversal and
 * path-integration library.
 *
 * The entry points replace, for batches, the reference crate's hot-path
 * trait methods:
 *   bg_scene_create      ~ accelerators::BVH::new            (src/accelerators/bvh.rs:53-89)
 *                          over GeometricPrimitive{Triangle,Sphere}
 *                          (src/primitives/geometric.rs:15-20,
 *                           src/shapes/triangle.rs:93-105, src/shapes/sphere.rs:38-65)
 *   bg_intersect(CLOSEST) ~ Primitive::intersect for BVH      (src/accelerators/bvh.rs:281-352)
 *   bg_intersect(ANY)     ~ Primitive::intersect_test for BVH (src/accelerators/bvh.rs:354-423)
 *   bg_camera_rays       ~ PerspectiveCamera::generate_ray    (src/cameras/perspective.rs:117-175)
 *   bg_render            ~ SamplerIntegrator::render with PathIntegrator::radiance
 *                          (src/base/integrator.rs:35-133, src/integrators/path.rs:51-146)
 *
 * Conventions: every function returns 0 (BG_OK) or a negative error code;
 * the message is available from bg_last_error(ctx). Host buffers belong to
 * the caller and are copied; device memory made by the library is freed by
 * the matching destroy call. A context is used by one host thread at a time.
 */
#ifndef BONSAI_GPU_H
#define BONSAI_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BG_OK 0
#define BG_ERR_INVALID (-1)
#define BG_ERR_HIP (-2)
#define BG_ERR_NOMEM (-3)
#define BG_ERR_UNSUPPORTED (-4)

#define BG_ABI_VERSION 4

/* Query flags for bg_intersect / bg_intersect_device. */
#define BG_CLOSEST 0u
#define BG_ANY 1u
/* Also count box and primitive tests into bg_stats (slower; for measuring
 * the algorithmic bytes of a ray set). */
#define BG_COUNT 2u

#define BG_MISS 0xFFFFFFFFu

typedef struct bg_ctx bg_ctx;
typedef struct bg_scene bg_scene;

/* A ray: origin, t_max, direction, time (src/geometries/ray.rs:10-17 minus
 * the differentials, which no hot-path result depends on). 32 bytes. */
typedef struct bg_ray {
    float o[3];
    float t_max;
    float d[3];
    float time;
} bg_ray;

/* A query result. prim_id is the index of the primitive in scene order
 * (triangles [0, n_tris), then spheres), BG_MISS on a miss. For a closest
 * hit on a triangle (b1, b2) are the barycentrics of p1 and p2
 * (triangle.rs:213-217) and t the hit distance; on a sphere (b1, b2) are the
 * surface (u, v) (sphere.rs:166-168). For an any-hit query prim_id is the
 * first primitive, in traversal order, that reported a hit, and t its hit
 * distance. 16 bytes. */
typedef struct bg_hit {
    uint32_t prim_id;
    float t;
    float b1;
    float b2;
} bg_hit;

/* Flattened BVH node as stored in HBM (32 bytes). For a leaf, `offset` is the
 * first slot of its primitive range and meta = count << 3 | has_sphere << 2.
 * For an interior node, `offset` is the second child (the first child is the
 * next node) and meta = axis (count 0). Same DFS order as bvh.rs:249-269. */
typedef struct bg_bvh_node {
    float bmin[3];
    float bmax[3];
    uint32_t offset;
    uint32_t meta;
} bg_bvh_node;

/* Materials (src/materials/matte.rs:20-41 and a conductor assembled from
 * src/bxdfs/microfacet.rs:68-151, src/base/microfacet.rs:32-154,
 * src/base/fresnel.rs:13-46). */
#define BG_MAT_MATTE 0u
#define BG_MAT_CONDUCTOR 1u
/* PlasticMaterial (src/materials/plastic.rs:24-52): Lambertian Kd plus a
 * Trowbridge-Reitz microfacet lobe Ks with FresnelDielectric(1.5, 1.0). Fields:
 * reflectance = Kd, eta = Ks, alpha = roughness, k[0] != 0: remap_roughness. */
#define BG_MAT_PLASTIC 2u
typedef struct bg_material {
    uint32_t type;
    float reflectance[3]; /* matte Kd / conductor R */
    float alpha;          /* Trowbridge-Reitz alpha (conductor) */
    float eta[3];         /* conductor eta */
    float k[3];           /* conductor absorption */
} bg_material;

/* A diffuse area light on one primitive (src/lights/diffuse.rs:98-123). */
typedef struct bg_area_light {
    uint32_t prim_id;
    float radiance[3];
    uint32_t double_sided;
} bg_area_light;

typedef struct bg_scene_desc {
    uint64_t n_tris;
    const float* tri_positions;           /* 9 floats per triangle, world space (non-indexed mesh, triangle.rs:96) */
    const uint8_t* tri_reverse_orientation; /* optional, 1 byte per triangle */
    const uint32_t* tri_material;         /* optional, material index per triangle */
    uint64_t n_spheres;
    const float* sphere_object_to_world;  /* 16 floats per sphere: Transform::m, row-major, affine */
    const float* sphere_world_to_object;  /* 16 floats per sphere: the same Transform's m_inverse */
    const float* sphere_params;           /* 4 floats per sphere: radius, z_min, z_max, phi_max (degrees) */
    const uint8_t* sphere_reverse_orientation; /* optional */
    const uint32_t* sphere_material;      /* optional */
    uint32_t n_materials;
    const bg_material* materials;         /* optional */
    uint32_t n_lights;
    const bg_area_light* lights;          /* optional */
} bg_scene_desc;

typedef struct bg_scene_info {
    uint64_t n_prims;
    uint64_t n_nodes;
    uint32_t max_depth;   /* largest number of pending far children any traversal can hold */
    uint32_t max_leaf;
    double build_ms;
    uint64_t device_bytes;
} bg_scene_info;

typedef struct bg_stats {
    double kernel_ms;     /* summed duration of the hot kernel launches (HIP events) */
    double total_ms;      /* wall time of the call on the device stream */
    uint64_t rays;        /* BVH queries issued */
    uint64_t launches;
    /* Filled only with BG_COUNT: box tests (bvh.rs:308-313) and primitive
     * tests (bvh.rs:316-321, 388-392) summed over all rays. */
    uint64_t nodes_visited;
    uint64_t tris_tested;
    uint64_t spheres_tested;
    /* Filled by the render calls. */
    uint64_t samples;     /* camera samples completed */
    uint64_t shadow_rays; /* any-hit queries among `rays` */
    uint64_t train_steps; /* guiding-network Adam steps taken by the call */
    /* Filled by the path and direct-lighting renders when the context times
     * traversal launches (BG_PROFILE_TRAVERSAL_TIME): summed HIP-event duration
     * of their traversal kernels alone. */
    double traverse_ms;
} bg_stats;

/* Library */
int bg_abi_version(void);
int bg_device_count(int* count);

/* Context: binds a device and owns a HIP stream. */
int bg_create(int device, bg_ctx** out);
void bg_destroy(bg_ctx* ctx);
const char* bg_last_error(const bg_ctx* ctx);
/* Stream the context launches on (hipStream_t), for callers sharing it. */
void* bg_stream(bg_ctx* ctx);

/* Scene: builds the SAH BVH on the host (bvh.rs:91-247 with the bucket index
 * and bucket initialisation fixed), flattens it and uploads it. */
int bg_scene_create(bg_ctx* ctx, const bg_scene_desc* desc, bg_scene** out);
void bg_scene_destroy(bg_scene* scene);
int bg_scene_info_get(const bg_scene* scene, bg_scene_info* info);
/* Copies the flattened BVH out: nodes (n_nodes) and slot -> primitive id
 * (n_prims). Either pointer may be NULL. */
int bg_scene_bvh(const bg_scene* scene, bg_bvh_node* nodes, uint32_t* slot_prim);

/* Host-only BVH build (no device needed): the same builder bg_scene_create
 * runs. Call with nodes == NULL to learn the sizes (*n_nodes, *n_prims), then
 * again with buffers of those sizes. */
int bg_build_bvh(const bg_scene_desc* desc, bg_bvh_node* nodes, uint64_t* n_nodes, uint32_t* slot_prim,
                 uint64_t* n_prims, uint32_t* max_depth);

/* OBJ ingestion: OBJ::read (src/io/obj.rs:16-124) with TriangleMesh::new
 * (src/shapes/triangle.rs:48-90). Host only. object_to_world and its inverse
 * are the Transform's m and m_inverse (row-major). Call with positions == NULL
 * to learn *n_tris and *has (BG_OBJ_NORMALS | BG_OBJ_UVS), then with buffers:
 * positions 9 floats per triangle (world space), normals 9 (transformed as
 * normals), uvs 6; normals / uvs may be NULL. Triangles are consecutive
 * triples of face vertices, as in the reference. Errors: BG_ERR_INVALID for an
 * unreadable file or a token the reference would fail on, BG_ERR_UNSUPPORTED
 * when only some face vertices carry normals or uvs. */
#define BG_OBJ_NORMALS 1u
#define BG_OBJ_UVS 2u
int bg_obj_read(const char* path, const float object_to_world[16], const float world_to_object[16], uint64_t* n_tris,
                uint32_t* has, float* positions, float* normals, float* uvs);

/* Batched queries, host buffers (copied in and out; includes PCIe). */
int bg_intersect(bg_ctx* ctx, const bg_scene* scene, const bg_ray* rays, size_t n, bg_hit* hits,
                 uint32_t flags, bg_stats* stats);
/* Batched queries on device-resident buffers (rays/hits in HBM), launched on
 * the context stream; returns after the kernel completes. */
int bg_intersect_device(bg_ctx* ctx, const bg_scene* scene, const void* d_rays, size_t n, void* d_hits,
                        uint32_t flags, bg_stats* stats);

/* bg_intersect_device without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). No counterpart in the reference. The traversal's stack
 * spill rows and the BG_COUNT counters live in a static device array of
 * BG_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends) instead of
 * context buffers; the grid is narrowed so that its spill rows fit, which
 * changes no result. Returns BG_ERR_INVALID when one block of the context's
 * launch shape needs more spill than the array holds (a BVH deeper than
 * BG_NOALLOC_SCRATCH_BYTES / (8 * block) - 1 pending entries). Every call on a
 * GPU shares that GPU's array: calls that overlap on one GPU (two streams, two
 * contexts or two threads at once) race and must not be issued. Results are
 * those of bg_intersect_device with the same arguments. */
#define BG_NOALLOC_SCRATCH_BYTES (128u << 20)
int bg_intersect_device_noalloc(bg_ctx* ctx, const bg_scene* scene, const void* d_rays, size_t n, void* d_hits,
                                uint32_t flags, bg_stats* stats);

/* The surface interaction of each hit (the geometry GeometricPrimitive::intersect
 * fills into an Interaction, src/primitives/geometric.rs:27-38 over
 * shapes/triangle.rs:135-388 and shapes/sphere.rs): for ray i and hits[i]
 * (from bg_intersect on the same rays) the point, its error bounds, the hit t,
 * the geometric and shading normals, the shading dpdu and the uv. All fields
 * are NaN for a miss. BG_ERR_INVALID if a hit names no primitive of the scene. */
typedef struct bg_surface {
    float p[3];
    float t;
    float p_error[3];
    float n[3];
    float ns[3];
    float dpdu[3];
    float uv[2];
    float pad[2];
} bg_surface;
int bg_surface_interactions(bg_ctx* ctx, const bg_scene* scene, const bg_ray* rays, const bg_hit* hits, size_t n,
                            bg_surface* out);

/* Traversal launch shape (chosen by the autotuner or the caller). A context
 * starts with two: bg_intersect* queries, the direct-lighting and Whitted
 * renders and the renders of scenes under 2^20 primitives use 256 threads, an
 * 8-entry stack, 32 blocks per CU and variant 4; the path and guided renders
 * of larger scenes use 128 threads, an 8-entry stack, 16 blocks per CU and
 * variant 5. bg_set_launch_config sets the one shape both then use;
 * bg_get_launch_config reads the first. Scene batches tune their own
 * (bg_batch_launch_config). */
typedef struct bg_launch_config {
    uint32_t block;       /* threads per block: 64, 128, 256 */
    uint32_t stack;       /* LDS stack entries per lane: 8, 16, 32 */
    uint32_t blocks_per_cu; /* persistent blocks per CU; 0 = one block per ray tile */
    uint32_t variant;     /* 0: a lane finishes its ray before the wavefront takes new rays;
                             1: each lane starts its next ray as soon as its current one ends
                                (node records; scene batches run it as 0);
                             2: as 0, with rays taken in chunks from an atomic work queue
                                (scene batches; other calls run it as 0);
                             3: as 0 on node records (scene batches run it as 0);
                             4 (default): once 16 lanes of a wavefront have finished their
                                rays, they take the next rays of the wavefront's sequence
                                (runs of consecutive rays) together, also over the compacted
                                rays of path-integrator launches; uncompacted launches of
                                fewer than four rays per lane, and scenes with spheres, run
                                it as 0;
                             5: as 4, the wavefronts taking runs of 256 rays from an atomic
                                work queue as they need them (persistent; also with spheres
                                and on compacted path-integrator launches) */
    uint32_t compact;     /* path-integrator passes: nonzero (default) traces only the rays that
                             can hit (t_max > 0), compacted on the device into an index list,
                             instead of every path slot; the hits are the same */
} bg_launch_config;
int bg_set_launch_config(bg_ctx* ctx, const bg_launch_config* cfg);
/* Measurement of the render calls' traversal launches (bg_render_film,
 * bg_render_film_guided, bg_render_film_direct, bg_render, bg_render_batch;
 * a measured batch render runs its passes as direct launches instead of a
 * graph replay): BG_PROFILE_COUNT counts their box and primitive tests into
 * bg_stats as BG_COUNT does for bg_intersect (slower: for the algorithmic
 * bytes of a render, in an untimed call); BG_PROFILE_TRAVERSAL_TIME records a
 * HIP event pair around each traversal kernel and writes the time during
 * which at least one of them ran (path renders trace the shadow and
 * MIS-branch rays beside the next extension rays) into bg_stats.traverse_ms.
 * 0 turns both off (the default). The images do not change. */
#define BG_PROFILE_COUNT 1u
#define BG_PROFILE_TRAVERSAL_TIME 2u
int bg_set_profiling(bg_ctx* ctx, uint32_t flags);
int bg_get_launch_config(const bg_ctx* ctx, bg_launch_config* cfg);

/* Camera (src/cameras/perspective.rs:49-114 with src/base/transform.rs). */
typedef struct bg_camera_desc {
    float camera_to_world[16];      /* row-major m of the (static) camera transform */
    float camera_to_world_inv[16];  /* its m_inverse */
    float fov;                      /* degrees */
    float near_clip, far_clip;
    float lens_radius, focal_distance;
    float shutter_open, shutter_close;
    uint32_t width, height;         /* film full resolution */
} bg_camera_desc;

/* Transform::look_at (src/base/transform.rs:152-189): camera_to_world m and
 * m_inverse, row-major. */
int bg_look_at(const float pos[3], const float look[3], const float up[3], float m[16], float m_inv[16]);

/* One camera ray per pixel through film point (x + fx, y + fy), pixels in
 * row-major order, written to device memory. */
int bg_camera_rays_device(bg_ctx* ctx, const bg_camera_desc* cam, float fx, float fy, void* d_rays);
int bg_camera_rays(bg_ctx* ctx, const bg_camera_desc* cam, float fx, float fy, bg_ray* rays);

/* ------------------------------------------------------------------ render */

/* Integrator semantics. REF reproduces src/integrators/path.rs and
 * src/base/integrator.rs as written (the camera ray is re-intersected with a
 * shrinking t_max instead of spawning continuation rays, the light pdf lacks
 * 1/area, light samples take emission toward +wi, the MIS branch traces the
 * light-sample direction, G1 = Lambda, sample11 keeps slope_x_1). CORRECTED
 * is the pbrt-v3 estimator those lines intend. */
#define BG_MODE_REF 0u
#define BG_MODE_CORRECTED 1u

typedef struct bg_render_desc {
    bg_camera_desc camera;
    uint32_t spp;           /* samples per pixel of this call */
    uint32_t first_sample;  /* sample index of the first one (progressive rendering) */
    uint32_t max_depth;
    float rr_threshold;     /* PathIntegrator rr_threshold (path.rs:133) */
    uint32_t mode;          /* BG_MODE_REF / BG_MODE_CORRECTED */
    uint64_t seed;
    uint32_t x0, y0, x1, y1; /* pixel window; x1 == 0 and y1 == 0 means the full film */
    uint32_t paths_per_pass; /* 0: about 64M for the path integrators (16M while a guide trains:
                                one training round follows each pass; whole samples per pixel, at
                                least one), about 64M slots (paths times light slots) for
                                direct lighting and Whitted */
} bg_render_desc;

/* Film: per-pixel RGB sums and filter-weight sums (src/base/film.rs). */
typedef struct bg_film bg_film;
#define BG_FILM_F64 1u  /* also keep FP64 sums (ground-truth renders) */
int bg_film_create(bg_ctx* ctx, uint32_t width, uint32_t height, uint32_t flags, bg_film** out);
void bg_film_destroy(bg_film* film);
int bg_film_clear(bg_film* film);
/* rgbw: width*height*4 floats (r, g, b, weight sums); rgbw64 likewise in
 * double (requires BG_FILM_F64). Either may be NULL. */
int bg_film_read(bg_film* film, float* rgbw, double* rgbw64);
/* Replaces the film's FP32 sums with rgbw (width*height*4 floats, as
 * bg_film_read returns them), e.g. to resume a render from saved sums; FP64
 * sums, if kept, are set to the same values. No counterpart in the reference,
 * whose film lives only for one render call. */
int bg_film_load(bg_film* film, const float* rgbw);

/* Renders desc->spp samples per pixel of the window into the film
 * (SamplerIntegrator::render, src/base/integrator.rs:35-133). */
int bg_render_film(bg_ctx* ctx, const bg_scene* scene, const bg_render_desc* desc, bg_film* film, bg_stats* stats);

/* The reference's other SamplerIntegrators on the GPU path:
 * DirectLightingIntegrator (src/integrators/direct.rs:72-113, with
 * uniform_sample_all_lights / uniform_sample_one_light and estimate_direct,
 * src/base/integrator.rs:274-437) and WhittedIntegrator
 * (src/integrators/whitted.rs:41-93). desc->mode selects the reference or
 * corrected semantics as for the path integrator; max_depth bounds only the
 * specular recursion, which adds nothing for the supported materials (none
 * has a specular lobe); rr_threshold is unused. Light slots (one per light
 * for BG_LIGHT_SAMPLE_ALL and Whitted, one for BG_LIGHT_SAMPLE_ONE) take 136
 * bytes each per camera sample of a pass; with a slot per light at most 16383
 * lights are supported. */
#define BG_INTEGRATOR_DIRECT 1u
#define BG_INTEGRATOR_WHITTED 2u
#define BG_LIGHT_SAMPLE_ALL 0u /* LightStrategy::UniformSampleAll */
#define BG_LIGHT_SAMPLE_ONE 1u /* LightStrategy::UniformSampleOne */
typedef struct bg_direct_desc {
    uint32_t integrator; /* BG_INTEGRATOR_DIRECT / BG_INTEGRATOR_WHITTED */
    uint32_t strategy;   /* direct lighting only */
} bg_direct_desc;
int bg_render_film_direct(bg_ctx* ctx, const bg_scene* scene, const bg_render_desc* desc, const bg_direct_desc* direct,
                          bg_film* film, bg_stats* stats);

/* Convenience: a fresh film, bg_render_film, then bg_film_write_image with
 * splat_scale 1 and scale 1 (integrator.rs:132) into rgb_out (width*height*3
 * floats of the full film). */
int bg_render(bg_ctx* ctx, const bg_scene* scene, const bg_render_desc* desc, float* rgb_out, bg_stats* stats);

/* Film::write_image (src/base/film.rs:156-202), resolved on the GPU: per pixel
 * the RGB sum goes to XYZ (merge_film_tile, film.rs:138-154) and back to RGB
 * (spectrum.rs:34-44), is divided by the weight sum and clamped at 0 when
 * that is non-zero (f32::max: a NaN becomes 0), gets splat_scale times the
 * (zero) splat added and is multiplied by scale (FilmOptions::scale). rgb_out:
 * width*height*3 floats on the host (top row first), or NULL. path: the
 * OpenEXR file Image::write (src/io/image.rs:47-53) would write, or NULL; at
 * least one of the two must be given. The file is written to path + ".part"
 * and renamed. */
int bg_film_write_image(bg_film* film, float splat_scale, float scale, float* rgb_out, const char* path);
/* The same resolve into device memory (width*height*3 floats) on the
 * context's stream, synchronised; ms (optional) receives the kernel time. */
int bg_film_resolve_device(bg_film* film, float splat_scale, float scale, void* d_rgb, double* ms);
/* Image::write (src/io/image.rs:47-53): width x height RGB f32 pixels (top
 * row first) as a single-part scanline OpenEXR file with FLOAT channels B, G,
 * R and no compression. */
int bg_exr_write(const char* path, uint32_t width, uint32_t height, const float* rgb);
/* The bytes bg_exr_write would write. With out NULL, *size receives the byte
 * count; otherwise *size is out's capacity on entry and the byte count on
 * return (BG_ERR_INVALID if the capacity is short). */
int bg_exr_encode(uint32_t width, uint32_t height, const float* rgb, unsigned char* out, uint64_t* size);

/* ----------------------------------------------------------- scene batches */

/* Many scenes, cameras and films in one wavefront loop (the reference renders
 * one scene per SamplerIntegrator::render call, src/base/integrator.rs:35-133;
 * a batch replaces a loop of such calls). Items may share a scene. Each item's
 * film receives exactly what bg_render_film with the same camera, seed, spp,
 * depth, rr_threshold and mode gives. */
typedef struct bg_batch bg_batch;
typedef struct bg_batch_item {
    const bg_scene* scene;
    bg_camera_desc camera;
    uint64_t seed;
} bg_batch_item;
/* Concatenates the items' scene data on the device (the scenes stay valid and
 * owned by the caller). */
int bg_batch_create(bg_ctx* ctx, const bg_batch_item* items, uint32_t n, bg_batch** out);
void bg_batch_destroy(bg_batch* batch);
typedef struct bg_batch_desc {
    uint32_t spp;            /* samples per pixel of every item */
    uint32_t max_depth;
    float rr_threshold;
    uint32_t mode;           /* BG_MODE_REF / BG_MODE_CORRECTED */
    uint32_t paths_per_pass; /* 0: about 256M (whole samples per pixel dividing spp) */
    uint32_t use_graph;      /* nonzero: capture one pass in a hipGraph and replay it */
    uint32_t autotune;       /* nonzero: time the traversal launch shapes on the first call */
    uint32_t schedule;       /* BG_SCHEDULE_WAVEFRONT: a kernel per stage of a pass (shade, each
                                traversal, resolve); BG_SCHEDULE_PERSISTENT: each pass as one
                                persistent path kernel: each lane runs one path at a time, and a
                                lane whose path has ended takes the next path slot from runs of 64
                                its wavefront claims from an atomic counter. Both give the same
                                films. */
} bg_batch_desc;
#define BG_SCHEDULE_WAVEFRONT 0u
#define BG_SCHEDULE_PERSISTENT 1u
/* Clears the batch film and renders spp samples of every item into it. */
int bg_render_batch(bg_ctx* ctx, bg_batch* batch, const bg_batch_desc* desc, bg_stats* stats);
/* Film sums of one item: width*height*4 floats (r, g, b, weight). */
int bg_batch_film_read(bg_batch* batch, uint32_t item, float* rgbw);
/* The traversal launch shape the batch renders with (tuned or the context's). */
int bg_batch_launch_config(const bg_batch* batch, bg_launch_config* cfg);

/* ------------------------------------------------------------ path guiding */

/* Learned directional proposal. The reference has no path guiding; these
 * entry points have no counterpart there. Network: 64-dim frequency encoding
 * of (position normalised to a box, outgoing direction, normal) -> 4 x 64 ReLU
 * -> 40 outputs = 8 von Mises-Fisher lobes x (mixture logit, direction
 * 3-vector, concentration parameter s: kappa = exp(1.5 + 5.5 tanh((s - 1.5) /
 * 5.5))). BF16 matrix-core products, FP32 accumulation,
 * FP32 master parameters and Adam state. Parameter order: per layer the weight
 * matrix (out x in, row-major) then the bias. */
#define BG_GUIDE_PARAMS 19240u
#define BG_GUIDE_OUTPUTS 40u
typedef struct bg_guide bg_guide;

/* params: BG_GUIDE_PARAMS floats, copied. max_batch bounds n of eval/train. */
int bg_guide_create(bg_ctx* ctx, const float* params, uint64_t max_batch, bg_guide** out);
void bg_guide_destroy(bg_guide* guide);
/* Replaces the parameters, restarts Adam (moments 0, step 0) and empties the
 * replay ring of training samples. */
int bg_guide_set_params(bg_guide* guide, const float* params);
/* Any output pointer may be NULL. */
int bg_guide_get_params(bg_guide* guide, float* params, float* adam_m, float* adam_v);
/* Box the positions are normalised to (each axis maps to [0, 1]). */
int bg_guide_set_bounds(bg_guide* guide, const float lo[3], const float hi[3]);
/* Initial parameters for bg_guide_create / bg_guide_set_params (host only, no
 * device needed): He-uniform weights, each U(-sqrt(6 / fan_in), sqrt(6 /
 * fan_in)) from the counter-based uniform bgx_uniform(seed, parameter index)
 * (24 random bits), and zero biases. params: BG_GUIDE_PARAMS floats. */
int bg_guide_init_params(uint64_t seed, float* params);
/* Snapshot of everything a guided render's training carries from one call to
 * the next: the FP32 parameters, Adam's moments and step count, the position
 * box and the replay ring of training samples (its capacity, contents and
 * cursor). bg_guide_restore of a snapshot into any guide of the same context
 * (its max_batch at least the snapshot's minibatch use) makes the following
 * calls give bit-identical results to those the snapshotted guide would have
 * given. With buf NULL *size receives the byte count; otherwise *size is
 * buf's capacity on entry and the byte count on return (BG_ERR_INVALID if the
 * capacity is short). No counterpart in the reference (it has no guiding). */
int bg_guide_snapshot(bg_guide* guide, void* buf, uint64_t* size);
int bg_guide_restore(bg_guide* guide, const void* buf, uint64_t size);
/* samples: n x 12 floats (p.xyz, -, wo.xyz, -, n.xyz, -); out: n x 40. */
int bg_guide_eval(bg_guide* guide, const float* samples, uint64_t n, float* out);
/* One Adam step on the batch mean of -weight * log p(direction).
 * targets: n x 4 floats (direction.xyz, weight). mean_loss (before the step)
 * and grad (BG_GUIDE_PARAMS floats) may be NULL. */
int bg_guide_train(bg_guide* guide, const float* samples, const float* targets, uint64_t n, float lr,
                   double* mean_loss, float* grad);

/* The same on device-resident buffers (same layouts), on the context stream;
 * kernel_ms (optional) receives the HIP-event time of the kernels. */
int bg_guide_eval_device(bg_guide* guide, const void* d_samples, uint64_t n, void* d_out, double* kernel_ms);
int bg_guide_train_device(bg_guide* guide, const void* d_samples, const void* d_targets, uint64_t n, float lr,
                          double* kernel_ms);

/* Guided rendering (BG_MODE_CORRECTED only). At every path vertex the
 * continuation direction comes from the network's mixture with probability
 * select_prob, else from the BSDF, and is weighted by f |cos| / (select_prob *
 * p_guide + (1 - select_prob) * p_bsdf) (one-sample MIS: the estimator stays
 * unbiased for any network). Light sampling is unchanged. With train != 0,
 * every pass records the vertices of up to train_paths paths, weighs each
 * sampled direction by w = luminance(f |cos| L_i / pdf), appends those with
 * w > 0 to the guide's replay ring, and takes `steps` Adam steps on
 * minibatches of the ring (loss: batch mean of -w log p_guide(direction)).
 * The guide's position box is set to the scene bounds. */
typedef struct bg_guiding_desc {
    float select_prob;     /* in [0, 1] */
    uint32_t train;        /* nonzero: train online after every pass */
    uint32_t train_paths;  /* paths per pass that record, spread over the pass (0: 2^20) */
    uint32_t batch;        /* Adam minibatch (0: 16384; at most the guide's max_batch) */
    uint32_t steps;        /* Adam steps after each pass (0: 8) */
    float lr;              /* Adam learning rate (0: 1e-2) */
    uint32_t replay;       /* replay ring capacity in samples (0: 2^20) */
    uint32_t lobes;        /* where the network is sampled: BG_GUIDE_DIFFUSE | BG_GUIDE_GLOSSY (0: both) */
    uint32_t bounces;      /* guide (and evaluate the network) at the first `bounces` vertices of a
                              path only; later vertices use the same estimator with select_prob 0
                              and are not recorded for training (0: every vertex) */
    uint32_t rr_start;     /* Russian roulette (rr_threshold) from this bounce on (0: bounce 4, as
                              the plain path, path.rs:133) */
} bg_guiding_desc;
#define BG_GUIDE_DIFFUSE 1u
#define BG_GUIDE_GLOSSY 2u
int bg_render_film_guided(bg_ctx* ctx, const bg_scene* scene, const bg_render_desc* desc, bg_guide* guide,
                          const bg_guiding_desc* guiding, bg_film* film, bg_stats* stats);
/* The training records of the last trained pass, in the layouts bg_guide_train
 * takes: inputs n x 12 floats, targets n x 4 (direction, weight; weight 0 for
 * vertices that did not occur). *n receives the record count; at most
 * max_records are copied. */
int bg_guide_records(bg_ctx* ctx, float* inputs, float* targets, uint64_t max_records, uint64_t* n);

#ifdef __cplusplus
}
#endif

#endif /* BONSAI_GPU_H */
