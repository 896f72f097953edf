//! `GpuPathIntegrator`: a drop-in for `PathIntegrator` (src/integrators/path.rs:19-40) whose
//! `render` (SamplerIntegrator::render, src/base/integrator.rs:35-133, then Film::write_image,
//! film.rs:156-202) runs the wavefront path tracer on the MI355X. Added to the bonsai crate as
//! `src/integrators/gpu_path.rs` beside `accelerators/gpu_bvh.rs`.
//!
//! The scene it renders is the `GpuBvh` it was given: the reference `Scene` holds its aggregate as
//! an opaque `Box<dyn Primitive>`, so `render(&Scene)` only checks that the scene's bounds are that
//! BVH's (the caller passes `Scene::new(Box::new(...), lights)` over the same GpuBvh, shared here
//! by an Arc). Lights are the area lights the GpuSceneBuilder received.
use std::ffi::CString;
use std::sync::Arc;

use bonsai_gpu_sys as sys;

use crate::{
    accelerators::gpu_bvh::GpuBvh,
    base::{constants::Float, integrator::Integrator, primitive::Primitive, scene::Scene, transform::Transform},
};

/// Which estimator: the reference files as written (camera ray re-intersected instead of a
/// continuation ray, light pdf without 1/area, ...) or the pbrt-v3 estimator they intend.
#[derive(Clone, Copy, Debug, PartialEq, Eq)]
pub enum Semantics {
    Reference,
    Corrected,
}

/// PerspectiveCameraOptions (perspective.rs:36-46) with a static transform, the film's full
/// resolution, scale and output file (FilmOptions, film.rs:53-60; box filter of radius 0.5),
/// StratifiedSamplerOptions with sampled_dimensions 0 (samples per pixel, seed) and
/// PathIntegrator's max_depth and rr_threshold.
#[derive(Clone, Debug)]
pub struct GpuPathIntegratorOptions {
    pub camera_to_world: Transform,
    pub fov: Float,
    pub near: Float,
    pub far: Float,
    pub lens_radius: Float,
    pub focal_distance: Float,
    pub shutter_open: Float,
    pub shutter_close: Float,
    pub width: u32,
    pub height: u32,
    pub scale: Float,
    pub filename: String,
    pub spp: u32,
    pub seed: u64,
    pub max_depth: u32,
    pub rr_threshold: Float,
    pub semantics: Semantics,
}

pub struct GpuPathIntegrator {
    opts: GpuPathIntegratorOptions,
    bvh: Arc<GpuBvh>,
}

fn mat16(m: &crate::geometries::mat4::Mat4) -> [f32; 16] {
    let mut out = [0.0f32; 16];
    for r in 0..4 {
        for c in 0..4 {
            out[4 * r + c] = m.m[r][c] as f32;
        }
    }
    out
}

impl GpuPathIntegrator {
    pub fn new(opts: GpuPathIntegratorOptions, bvh: Arc<GpuBvh>) -> Self {
        Self { opts, bvh }
    }

    fn desc(&self) -> sys::bg_render_desc {
        let o = &self.opts;
        sys::bg_render_desc {
            camera: sys::bg_camera_desc {
                camera_to_world: mat16(&o.camera_to_world.m),
                camera_to_world_inv: mat16(&o.camera_to_world.m_inverse),
                fov: o.fov as f32,
                near_clip: o.near as f32,
                far_clip: o.far as f32,
                lens_radius: o.lens_radius as f32,
                focal_distance: o.focal_distance as f32,
                shutter_open: o.shutter_open as f32,
                shutter_close: o.shutter_close as f32,
                width: o.width,
                height: o.height,
            },
            spp: o.spp,
            first_sample: 0,
            max_depth: o.max_depth,
            rr_threshold: o.rr_threshold as f32,
            mode: match o.semantics {
                Semantics::Reference => sys::BG_MODE_REF,
                Semantics::Corrected => sys::BG_MODE_CORRECTED,
            },
            seed: o.seed,
            x0: 0,
            y0: 0,
            x1: 0,
            y1: 0,
            paths_per_pass: 0,
        }
    }

    /// Renders into a fresh film and returns the resolved image (width x height RGB, top row first)
    /// without writing a file.
    pub fn render_image(&self) -> Vec<f32> {
        let (film, ctx) = self.render_film();
        let mut rgb = vec![0.0f32; 3 * (self.opts.width * self.opts.height) as usize];
        let rc = unsafe { sys::bg_film_write_image(film, 1.0, self.opts.scale as f32, rgb.as_mut_ptr(), std::ptr::null()) };
        unsafe { sys::bg_film_destroy(film) };
        assert_eq!(rc, sys::BG_OK, "{}", sys::last_error(ctx));
        rgb
    }

    fn render_film(&self) -> (*mut sys::bg_film, *const sys::bg_ctx) {
        let d = self.bvh.device();
        let mut film: *mut sys::bg_film = std::ptr::null_mut();
        let rc = unsafe { sys::bg_film_create(d.ctx, self.opts.width, self.opts.height, 0, &mut film) };
        assert_eq!(rc, sys::BG_OK, "{}", sys::last_error(d.ctx));
        let desc = self.desc();
        let rc = unsafe { sys::bg_render_film(d.ctx, d.scene, &desc, film, std::ptr::null_mut()) };
        if rc != sys::BG_OK {
            unsafe { sys::bg_film_destroy(film) };
            panic!("bg_render_film: {}", sys::last_error(d.ctx));
        }
        (film, d.ctx)
    }
}

impl Integrator for GpuPathIntegrator {
    fn render(&self, scene: &Scene) {
        assert_eq!(scene.bounds, self.bvh.bounds(), "GpuPathIntegrator renders the GpuBvh it was built with");
        let (film, ctx) = self.render_film();
        // Film::write_image (splat scale 1, as integrator.rs:132) and Image::write of the EXR
        let path = CString::new(self.opts.filename.as_str()).expect("file name");
        let rc = unsafe { sys::bg_film_write_image(film, 1.0, self.opts.scale as f32, std::ptr::null_mut(), path.as_ptr()) };
        unsafe { sys::bg_film_destroy(film) };
        assert_eq!(rc, sys::BG_OK, "{}", sys::last_error(ctx));
    }
}
