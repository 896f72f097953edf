//! `GpuBvh`: a drop-in for `accelerators::BVH` (src/accelerators/bvh.rs:20-23, 271-445) whose
//! queries run on the MI355X through bonsai-gpu-sys. This file is added to the bonsai crate as
//! `src/accelerators/gpu_bvh.rs` (with `pub mod gpu_bvh;` in `src/accelerators/mod.rs` and
//! `bonsai-gpu-sys = { path = ".../bonsai-gpu-sys" }` in Cargo.toml): the `Primitive` trait takes
//! `&mut Interaction`, a type the crate does not export, so the implementation has to live inside it.
//!
//! The reference `Scene` holds its aggregate as an opaque `Box<dyn Primitive>`, from which no
//! geometry can be read back, so a `GpuBvh` is built from a `GpuSceneBuilder` that receives each
//! primitive twice: as the crate's own `Arc<dyn Primitive>` (a `GeometricPrimitive` over a
//! `Triangle` or `Sphere`) and as the arrays the GPU scene is made of. Primitive ids are scene order:
//! triangles first, then spheres.
//!
//! Single-ray trait calls (`intersect`, `intersect_test`, as `Scene::intersect` makes them,
//! scene.rs:36-42) launch one query; the `Interaction` of a hit is then filled by the hit
//! primitive's own `intersect`, which gives the same t because the GPU's arithmetic is the
//! reference's (the same t_max shrink of geometric.rs:32 applies). Batched queries, which is what
//! the GPU is for, go through `intersect_batch` and `surfaces`.
use std::sync::{Arc, Mutex, MutexGuard};

use bonsai_gpu_sys as sys;

use crate::{
    base::{
        constants::Float,
        interaction::Interaction,
        light::AreaLight,
        material::{Material, TransportMode},
        primitive::Primitive,
        transform::Transform,
    },
    geometries::{bounds3::Bounds3, point3::Point3, ray::Ray},
};

#[derive(Debug)]
pub struct GpuError {
    pub code: i32,
    pub message: String,
}

fn check(ctx: *const sys::bg_ctx, rc: i32) -> Result<(), GpuError> {
    if rc == sys::BG_OK {
        Ok(())
    } else {
        Err(GpuError { code: rc, message: sys::last_error(ctx) })
    }
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

/// Emission of an area light on a primitive (DiffuseAreaLightOptions: radiance, double_sided).
#[derive(Clone, Copy, Debug)]
pub struct Emission {
    pub radiance: [f32; 3],
    pub double_sided: bool,
}

#[derive(Default)]
pub struct GpuSceneBuilder {
    tris: Vec<Arc<dyn Primitive>>,
    spheres: Vec<Arc<dyn Primitive>>,
    tri_positions: Vec<f32>,
    tri_reverse: Vec<u8>,
    tri_material: Vec<u32>,
    sph_o2w: Vec<f32>,
    sph_w2o: Vec<f32>,
    sph_params: Vec<f32>,
    sph_reverse: Vec<u8>,
    sph_material: Vec<u32>,
    materials: Vec<sys::bg_material>,
    // (is_sphere, index among its kind, emission)
    lights: Vec<(bool, usize, Emission)>,
}

impl GpuSceneBuilder {
    pub fn new() -> Self {
        Self::default()
    }

    /// A material record (MatteMaterial with constant Kd and sigma 0, the conductor, or
    /// PlasticMaterial with constant textures); returns its index.
    pub fn material(&mut self, m: sys::bg_material) -> u32 {
        self.materials.push(m);
        (self.materials.len() - 1) as u32
    }

    /// One triangle of a TriangleMesh (world-space vertices, triangle.rs:93-105), with the
    /// GeometricPrimitive built over it.
    pub fn triangle(&mut self, prim: Arc<dyn Primitive>, p: [Point3; 3], reverse_orientation: bool, material: u32,
                    emission: Option<Emission>) {
        for v in p.iter() {
            self.tri_positions.extend_from_slice(&[v.x as f32, v.y as f32, v.z as f32]);
        }
        self.tri_reverse.push(reverse_orientation as u8);
        self.tri_material.push(material);
        if let Some(e) = emission {
            self.lights.push((false, self.tris.len(), e));
        }
        self.tris.push(prim);
    }

    /// A Sphere (SphereOptions: transform, radius, z_min, z_max, phi_max in degrees, sphere.rs:38-65)
    /// with the GeometricPrimitive built over it. The transform must be affine.
    pub fn sphere(&mut self, prim: Arc<dyn Primitive>, object_to_world: &Transform, radius: Float, z_min: Float,
                  z_max: Float, phi_max: Float, reverse_orientation: bool, material: u32,
                  emission: Option<Emission>) {
        self.sph_o2w.extend_from_slice(&mat16(&object_to_world.m));
        self.sph_w2o.extend_from_slice(&mat16(&object_to_world.m_inverse));
        self.sph_params.extend_from_slice(&[radius as f32, z_min as f32, z_max as f32, phi_max as f32]);
        self.sph_reverse.push(reverse_orientation as u8);
        self.sph_material.push(material);
        if let Some(e) = emission {
            self.lights.push((true, self.spheres.len(), e));
        }
        self.spheres.push(prim);
    }

    /// Builds the SAH BVH on the host (bvh.rs:91-269 with its bucket index fixed), flattens it and
    /// uploads the scene to `device`.
    pub fn build(self, device: i32) -> Result<GpuBvh, GpuError> {
        let n_tris = self.tris.len();
        let lights: Vec<sys::bg_area_light> = self
            .lights
            .iter()
            .map(|&(sph, k, e)| sys::bg_area_light {
                prim_id: (if sph { n_tris + k } else { k }) as u32,
                radiance: e.radiance,
                double_sided: e.double_sided as u32,
            })
            .collect();
        let desc = sys::bg_scene_desc {
            n_tris: n_tris as u64,
            tri_positions: self.tri_positions.as_ptr(),
            tri_reverse_orientation: self.tri_reverse.as_ptr(),
            tri_material: self.tri_material.as_ptr(),
            n_spheres: self.spheres.len() as u64,
            sphere_object_to_world: self.sph_o2w.as_ptr(),
            sphere_world_to_object: self.sph_w2o.as_ptr(),
            sphere_params: self.sph_params.as_ptr(),
            sphere_reverse_orientation: self.sph_reverse.as_ptr(),
            sphere_material: self.sph_material.as_ptr(),
            n_materials: self.materials.len() as u32,
            materials: self.materials.as_ptr(),
            n_lights: lights.len() as u32,
            lights: lights.as_ptr(),
        };
        let mut ctx: *mut sys::bg_ctx = std::ptr::null_mut();
        check(std::ptr::null(), unsafe { sys::bg_create(device, &mut ctx) })?;
        let mut scene: *mut sys::bg_scene = std::ptr::null_mut();
        if let Err(e) = check(ctx, unsafe { sys::bg_scene_create(ctx, &desc, &mut scene) }) {
            unsafe { sys::bg_destroy(ctx) };
            return Err(e);
        }
        let dev = Device { ctx, scene };
        let mut info = sys::bg_scene_info::default();
        check(ctx, unsafe { sys::bg_scene_info_get(scene, &mut info) })?;
        // Primitive::bounds of the BVH (bvh.rs:272-279): the root node's box, the default when empty
        let bounds = if info.n_nodes == 0 {
            Bounds3::default()
        } else {
            let mut nodes = vec![sys::bg_bvh_node::default(); info.n_nodes as usize];
            check(ctx, unsafe { sys::bg_scene_bvh(scene, nodes.as_mut_ptr(), std::ptr::null_mut()) })?;
            let r = nodes[0];
            Bounds3 {
                min: Point3::new(r.bmin[0] as Float, r.bmin[1] as Float, r.bmin[2] as Float),
                max: Point3::new(r.bmax[0] as Float, r.bmax[1] as Float, r.bmax[2] as Float),
            }
        };
        let mut prims = self.tris;
        prims.extend(self.spheres);
        Ok(GpuBvh { dev: Mutex::new(dev), prims, bounds })
    }
}

/// The device context and scene; a context is used by one thread at a time, hence the Mutex.
pub struct Device {
    pub ctx: *mut sys::bg_ctx,
    pub scene: *mut sys::bg_scene,
}
unsafe impl Send for Device {}

pub struct GpuBvh {
    dev: Mutex<Device>,
    prims: Vec<Arc<dyn Primitive>>,
    bounds: Bounds3,
}

impl Drop for GpuBvh {
    fn drop(&mut self) {
        let d = self.dev.get_mut().unwrap();
        unsafe {
            sys::bg_scene_destroy(d.scene);
            sys::bg_destroy(d.ctx);
        }
    }
}

/// The reference Ray (ray.rs:10-17) as the 32-byte GPU ray; differentials decide no hit.
pub fn to_bg_ray(r: &Ray) -> sys::bg_ray {
    sys::bg_ray {
        o: [r.origin.x as f32, r.origin.y as f32, r.origin.z as f32],
        t_max: r.t_max as f32,
        d: [r.direction.x as f32, r.direction.y as f32, r.direction.z as f32],
        time: r.time as f32,
    }
}

impl GpuBvh {
    pub fn device(&self) -> MutexGuard<'_, Device> {
        self.dev.lock().unwrap()
    }

    /// Closest hits (bvh.rs:281-352) or, with `any_hit`, the first hit in traversal order
    /// (bvh.rs:354-423) of every ray; prim_id BG_MISS on a miss.
    pub fn intersect_batch(&self, rays: &[sys::bg_ray], any_hit: bool) -> Result<Vec<sys::bg_hit>, GpuError> {
        let mut hits = vec![sys::bg_hit::default(); rays.len()];
        let d = self.device();
        let flags = if any_hit { sys::BG_ANY } else { sys::BG_CLOSEST };
        check(d.ctx, unsafe {
            sys::bg_intersect(d.ctx, d.scene, rays.as_ptr(), rays.len(), hits.as_mut_ptr(), flags, std::ptr::null_mut())
        })?;
        Ok(hits)
    }

    /// The geometry of each hit's Interaction (point, error, t, normals, dpdu, uv; NaN for misses).
    pub fn surfaces(&self, rays: &[sys::bg_ray], hits: &[sys::bg_hit]) -> Result<Vec<sys::bg_surface>, GpuError> {
        assert_eq!(rays.len(), hits.len());
        let mut out = vec![sys::bg_surface::default(); rays.len()];
        let d = self.device();
        check(d.ctx, unsafe {
            sys::bg_surface_interactions(d.ctx, d.scene, rays.as_ptr(), hits.as_ptr(), rays.len(), out.as_mut_ptr())
        })?;
        Ok(out)
    }
}

impl Primitive for GpuBvh {
    fn bounds(&self) -> Bounds3 {
        self.bounds
    }

    fn intersect(&self, ray: &mut Ray, it: &mut Interaction) -> bool {
        let hit = self.intersect_batch(&[to_bg_ray(ray)], false).expect("GpuBvh::intersect")[0];
        if hit.prim_id == sys::BG_MISS {
            return false;
        }
        // the hit primitive fills the Interaction and shrinks t_max to the same t (geometric.rs:27-38)
        self.prims[hit.prim_id as usize].intersect(ray, it)
    }

    fn intersect_test(&self, ray: &Ray) -> bool {
        self.intersect_batch(&[to_bg_ray(ray)], true).expect("GpuBvh::intersect_test")[0].prim_id != sys::BG_MISS
    }

    fn compute_scattering_functions(
        &self,
        _it: &mut Interaction,
        _transport_mode: TransportMode,
        _allow_multiple_lobes: bool,
    ) {
        unimplemented!(); // as BVH (bvh.rs:425-432): only the hit primitive has a material
    }

    fn material(&self) -> Option<&dyn Material> {
        None
    }

    fn area_light(&self) -> Option<&dyn AreaLight> {
        None
    }
}
