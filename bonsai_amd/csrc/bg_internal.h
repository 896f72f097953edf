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
};

struct LaunchShape {
    int block = 128;
    int stack = 16;
    int blocks_per_cu = 8;
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

}  // namespace bg

struct bg_scene {
    bg_ctx* ctx = nullptr;
    bg_scene_info info{};
    bg::HostBVH bvh;
    uint64_t n_tris = 0, n_spheres = 0;
    // device
    float4* d_nodes = nullptr;
    float* d_tri_verts = nullptr;
    uint32_t* d_slot_prim = nullptr;
    bg::SphereDev* d_spheres = nullptr;
    bg::SphereAux* d_sphere_aux = nullptr;
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

struct bg_ctx {
    int device = 0;
    int num_cus = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    bg::LaunchShape shape;
    // reusable device buffers
    uint32_t* spill = nullptr;
    unsigned long long* counters = nullptr;
    size_t spill_bytes = 0;
    void* io = nullptr;
    size_t io_bytes = 0;
};
