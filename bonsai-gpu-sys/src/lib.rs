//! Raw FFI to libbonsai_gpu (include/bonsai_gpu.h, include/bonsai_scenes.h): the MI355X (gfx950)
//! BVH traversal and path integration library. Every struct is `#[repr(C)]` with the header's
//! field order; tests/test_rust_abi.py checks sizes, field offsets and the function list against
//! the C header.
#![allow(non_camel_case_types)]

use std::os::raw::{c_char, c_int, c_uchar, c_void};

pub const BG_OK: c_int = 0;
pub const BG_ERR_INVALID: c_int = -1;
pub const BG_ERR_HIP: c_int = -2;
pub const BG_ERR_NOMEM: c_int = -3;
pub const BG_ERR_UNSUPPORTED: c_int = -4;
pub const BG_ABI_VERSION: c_int = 4;

pub const BG_CLOSEST: u32 = 0;
pub const BG_ANY: u32 = 1;
pub const BG_COUNT: u32 = 2;
pub const BG_MISS: u32 = 0xFFFF_FFFF;

pub const BG_MAT_MATTE: u32 = 0;
pub const BG_MAT_CONDUCTOR: u32 = 1;
pub const BG_MAT_PLASTIC: u32 = 2;

pub const BG_OBJ_NORMALS: u32 = 1;
pub const BG_OBJ_UVS: u32 = 2;
pub const BG_NOALLOC_SCRATCH_BYTES: u32 = 128 << 20;
pub const BG_PROFILE_COUNT: u32 = 1;
pub const BG_PROFILE_TRAVERSAL_TIME: u32 = 2;

pub const BG_MODE_REF: u32 = 0;
pub const BG_MODE_CORRECTED: u32 = 1;
pub const BG_FILM_F64: u32 = 1;
pub const BG_INTEGRATOR_DIRECT: u32 = 1;
pub const BG_INTEGRATOR_WHITTED: u32 = 2;
pub const BG_LIGHT_SAMPLE_ALL: u32 = 0;
pub const BG_LIGHT_SAMPLE_ONE: u32 = 1;

pub const BG_GUIDE_PARAMS: u32 = 19240;
pub const BG_GUIDE_OUTPUTS: u32 = 40;
pub const BG_GUIDE_DIFFUSE: u32 = 1;
pub const BG_GUIDE_GLOSSY: u32 = 2;

// Opaque handles.
#[repr(C)]
pub struct bg_ctx {
    _p: [u8; 0],
}
#[repr(C)]
pub struct bg_scene {
    _p: [u8; 0],
}
#[repr(C)]
pub struct bg_film {
    _p: [u8; 0],
}
#[repr(C)]
pub struct bg_batch {
    _p: [u8; 0],
}
#[repr(C)]
pub struct bg_guide {
    _p: [u8; 0],
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct bg_ray {
    pub o: [f32; 3],
    pub t_max: f32,
    pub d: [f32; 3],
    pub time: f32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct bg_hit {
    pub prim_id: u32,
    pub t: f32,
    pub b1: f32,
    pub b2: f32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct bg_bvh_node {
    pub bmin: [f32; 3],
    pub bmax: [f32; 3],
    pub offset: u32,
    pub meta: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct bg_material {
    pub type_: u32,
    pub reflectance: [f32; 3],
    pub alpha: f32,
    pub eta: [f32; 3],
    pub k: [f32; 3],
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct bg_area_light {
    pub prim_id: u32,
    pub radiance: [f32; 3],
    pub double_sided: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct bg_scene_desc {
    pub n_tris: u64,
    pub tri_positions: *const f32,
    pub tri_reverse_orientation: *const u8,
    pub tri_material: *const u32,
    pub n_spheres: u64,
    pub sphere_object_to_world: *const f32,
    pub sphere_world_to_object: *const f32,
    pub sphere_params: *const f32,
    pub sphere_reverse_orientation: *const u8,
    pub sphere_material: *const u32,
    pub n_materials: u32,
    pub materials: *const bg_material,
    pub n_lights: u32,
    pub lights: *const bg_area_light,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct bg_scene_info {
    pub n_prims: u64,
    pub n_nodes: u64,
    pub max_depth: u32,
    pub max_leaf: u32,
    pub build_ms: f64,
    pub device_bytes: u64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct bg_stats {
    pub kernel_ms: f64,
    pub total_ms: f64,
    pub rays: u64,
    pub launches: u64,
    pub nodes_visited: u64,
    pub tris_tested: u64,
    pub spheres_tested: u64,
    pub samples: u64,
    pub shadow_rays: u64,
    pub train_steps: u64,
    pub traverse_ms: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct bg_surface {
    pub p: [f32; 3],
    pub t: f32,
    pub p_error: [f32; 3],
    pub n: [f32; 3],
    pub ns: [f32; 3],
    pub dpdu: [f32; 3],
    pub uv: [f32; 2],
    pub pad: [f32; 2],
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct bg_launch_config {
    pub block: u32,
    pub stack: u32,
    pub blocks_per_cu: u32,
    pub variant: u32,
    pub compact: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct bg_camera_desc {
    pub camera_to_world: [f32; 16],
    pub camera_to_world_inv: [f32; 16],
    pub fov: f32,
    pub near_clip: f32,
    pub far_clip: f32,
    pub lens_radius: f32,
    pub focal_distance: f32,
    pub shutter_open: f32,
    pub shutter_close: f32,
    pub width: u32,
    pub height: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct bg_render_desc {
    pub camera: bg_camera_desc,
    pub spp: u32,
    pub first_sample: u32,
    pub max_depth: u32,
    pub rr_threshold: f32,
    pub mode: u32,
    pub seed: u64,
    pub x0: u32,
    pub y0: u32,
    pub x1: u32,
    pub y1: u32,
    pub paths_per_pass: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct bg_direct_desc {
    pub integrator: u32,
    pub strategy: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct bg_batch_item {
    pub scene: *const bg_scene,
    pub camera: bg_camera_desc,
    pub seed: u64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct bg_batch_desc {
    pub spp: u32,
    pub max_depth: u32,
    pub rr_threshold: f32,
    pub mode: u32,
    pub paths_per_pass: u32,
    pub use_graph: u32,
    pub autotune: u32,
    pub schedule: u32,
}

pub const BG_SCHEDULE_WAVEFRONT: u32 = 0;
pub const BG_SCHEDULE_PERSISTENT: u32 = 1;

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct bg_guiding_desc {
    pub select_prob: f32,
    pub train: u32,
    pub train_paths: u32,
    pub batch: u32,
    pub steps: u32,
    pub lr: f32,
    pub replay: u32,
    pub lobes: u32,
    pub bounces: u32,
    pub rr_start: u32,
}

extern "C" {
    // library and context
    pub fn bg_abi_version() -> c_int;
    pub fn bg_device_count(count: *mut c_int) -> c_int;
    pub fn bg_create(device: c_int, out: *mut *mut bg_ctx) -> c_int;
    pub fn bg_destroy(ctx: *mut bg_ctx);
    pub fn bg_last_error(ctx: *const bg_ctx) -> *const c_char;
    pub fn bg_stream(ctx: *mut bg_ctx) -> *mut c_void;

    // scenes and the BVH (accelerators::BVH::new, bvh.rs:53-89)
    pub fn bg_scene_create(ctx: *mut bg_ctx, desc: *const bg_scene_desc, out: *mut *mut bg_scene) -> c_int;
    pub fn bg_scene_destroy(scene: *mut bg_scene);
    pub fn bg_scene_info_get(scene: *const bg_scene, info: *mut bg_scene_info) -> c_int;
    pub fn bg_scene_bvh(scene: *const bg_scene, nodes: *mut bg_bvh_node, slot_prim: *mut u32) -> c_int;
    pub fn bg_build_bvh(desc: *const bg_scene_desc, nodes: *mut bg_bvh_node, n_nodes: *mut u64, slot_prim: *mut u32,
                        n_prims: *mut u64, max_depth: *mut u32) -> c_int;
    pub fn bg_obj_read(path: *const c_char, object_to_world: *const f32, world_to_object: *const f32,
                       n_tris: *mut u64, has: *mut u32, positions: *mut f32, normals: *mut f32, uvs: *mut f32)
                       -> c_int;

    // queries (Primitive::intersect / intersect_test of the BVH, bvh.rs:281-423)
    pub fn bg_intersect(ctx: *mut bg_ctx, scene: *const bg_scene, rays: *const bg_ray, n: usize, hits: *mut bg_hit,
                        flags: u32, stats: *mut bg_stats) -> c_int;
    pub fn bg_intersect_device(ctx: *mut bg_ctx, scene: *const bg_scene, d_rays: *const c_void, n: usize,
                               d_hits: *mut c_void, flags: u32, stats: *mut bg_stats) -> c_int;
    pub fn bg_intersect_device_noalloc(ctx: *mut bg_ctx, scene: *const bg_scene, d_rays: *const c_void, n: usize,
                                       d_hits: *mut c_void, flags: u32, stats: *mut bg_stats) -> c_int;
    pub fn bg_surface_interactions(ctx: *mut bg_ctx, scene: *const bg_scene, rays: *const bg_ray,
                                   hits: *const bg_hit, n: usize, out: *mut bg_surface) -> c_int;
    pub fn bg_set_launch_config(ctx: *mut bg_ctx, cfg: *const bg_launch_config) -> c_int;
    pub fn bg_get_launch_config(ctx: *const bg_ctx, cfg: *mut bg_launch_config) -> c_int;
    pub fn bg_set_profiling(ctx: *mut bg_ctx, flags: u32) -> c_int;

    // camera (perspective.rs:49-175, transform.rs:152-189)
    pub fn bg_look_at(pos: *const f32, look: *const f32, up: *const f32, m: *mut f32, m_inv: *mut f32) -> c_int;
    pub fn bg_camera_rays_device(ctx: *mut bg_ctx, cam: *const bg_camera_desc, fx: f32, fy: f32,
                                 d_rays: *mut c_void) -> c_int;
    pub fn bg_camera_rays(ctx: *mut bg_ctx, cam: *const bg_camera_desc, fx: f32, fy: f32, rays: *mut bg_ray) -> c_int;

    // film and rendering (SamplerIntegrator::render, integrator.rs:35-133)
    pub fn bg_film_create(ctx: *mut bg_ctx, width: u32, height: u32, flags: u32, out: *mut *mut bg_film) -> c_int;
    pub fn bg_film_destroy(film: *mut bg_film);
    pub fn bg_film_clear(film: *mut bg_film) -> c_int;
    pub fn bg_film_read(film: *mut bg_film, rgbw: *mut f32, rgbw64: *mut f64) -> c_int;
    pub fn bg_film_load(film: *mut bg_film, rgbw: *const f32) -> c_int;
    pub fn bg_render_film(ctx: *mut bg_ctx, scene: *const bg_scene, desc: *const bg_render_desc,
                          film: *mut bg_film, stats: *mut bg_stats) -> c_int;
    pub fn bg_render_film_direct(ctx: *mut bg_ctx, scene: *const bg_scene, desc: *const bg_render_desc,
                                 direct: *const bg_direct_desc, film: *mut bg_film, stats: *mut bg_stats) -> c_int;
    pub fn bg_render(ctx: *mut bg_ctx, scene: *const bg_scene, desc: *const bg_render_desc, rgb_out: *mut f32,
                     stats: *mut bg_stats) -> c_int;
    pub fn bg_film_write_image(film: *mut bg_film, splat_scale: f32, scale: f32, rgb_out: *mut f32,
                               path: *const c_char) -> c_int;
    pub fn bg_film_resolve_device(film: *mut bg_film, splat_scale: f32, scale: f32, d_rgb: *mut c_void,
                                  ms: *mut f64) -> c_int;
    pub fn bg_exr_write(path: *const c_char, width: u32, height: u32, rgb: *const f32) -> c_int;
    pub fn bg_exr_encode(width: u32, height: u32, rgb: *const f32, out: *mut c_uchar, size: *mut u64) -> c_int;

    // scene batches
    pub fn bg_batch_create(ctx: *mut bg_ctx, items: *const bg_batch_item, n: u32, out: *mut *mut bg_batch) -> c_int;
    pub fn bg_batch_destroy(batch: *mut bg_batch);
    pub fn bg_render_batch(ctx: *mut bg_ctx, batch: *mut bg_batch, desc: *const bg_batch_desc,
                           stats: *mut bg_stats) -> c_int;
    pub fn bg_batch_film_read(batch: *mut bg_batch, item: u32, rgbw: *mut f32) -> c_int;
    pub fn bg_batch_launch_config(batch: *const bg_batch, cfg: *mut bg_launch_config) -> c_int;

    // path guiding (no counterpart in the reference)
    pub fn bg_guide_create(ctx: *mut bg_ctx, params: *const f32, max_batch: u64, out: *mut *mut bg_guide) -> c_int;
    pub fn bg_guide_destroy(guide: *mut bg_guide);
    pub fn bg_guide_set_params(guide: *mut bg_guide, params: *const f32) -> c_int;
    pub fn bg_guide_get_params(guide: *mut bg_guide, params: *mut f32, adam_m: *mut f32, adam_v: *mut f32) -> c_int;
    pub fn bg_guide_set_bounds(guide: *mut bg_guide, lo: *const f32, hi: *const f32) -> c_int;
    pub fn bg_guide_init_params(seed: u64, params: *mut f32) -> c_int;
    pub fn bg_guide_snapshot(guide: *mut bg_guide, buf: *mut c_void, size: *mut u64) -> c_int;
    pub fn bg_guide_restore(guide: *mut bg_guide, buf: *const c_void, size: u64) -> c_int;
    pub fn bg_guide_eval(guide: *mut bg_guide, samples: *const f32, n: u64, out: *mut f32) -> c_int;
    pub fn bg_guide_train(guide: *mut bg_guide, samples: *const f32, targets: *const f32, n: u64, lr: f32,
                          mean_loss: *mut f64, grad: *mut f32) -> c_int;
    pub fn bg_guide_eval_device(guide: *mut bg_guide, d_samples: *const c_void, n: u64, d_out: *mut c_void,
                                kernel_ms: *mut f64) -> c_int;
    pub fn bg_guide_train_device(guide: *mut bg_guide, d_samples: *const c_void, d_targets: *const c_void, n: u64,
                                 lr: f32, kernel_ms: *mut f64) -> c_int;
    pub fn bg_render_film_guided(ctx: *mut bg_ctx, scene: *const bg_scene, desc: *const bg_render_desc,
                                 guide: *mut bg_guide, guiding: *const bg_guiding_desc, film: *mut bg_film,
                                 stats: *mut bg_stats) -> c_int;
    pub fn bg_guide_records(ctx: *mut bg_ctx, inputs: *mut f32, targets: *mut f32, max_records: u64, n: *mut u64)
                            -> c_int;

    // seeded workload generators (include/bonsai_scenes.h)
    pub fn bgx_icosphere_tri_count(level: c_int) -> u64;
    pub fn bgx_icosphere(level: c_int, center: *const f32, radius: f32, amplitude: f32, frequency: f32, seed: u64,
                         out_tris: *mut f32) -> c_int;
    pub fn bgx_rays_incoherent(n: u64, bmin: *const f32, bmax: *const f32, seed: u64, out: *mut bg_ray) -> c_int;
    pub fn bgx_uniform(seed: u64, index: u64) -> f32;
}

/// The library's message for the last failed call on `ctx`.
pub fn last_error(ctx: *const bg_ctx) -> String {
    unsafe {
        let p = bg_last_error(ctx);
        if p.is_null() {
            return String::new();
        }
        std::ffi::CStr::from_ptr(p).to_string_lossy().into_owned()
    }
}
