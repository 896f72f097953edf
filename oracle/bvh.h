// Oracle (test infrastructure only): bonsai's BVH (accelerators/bvh.rs)
// restated on the CPU: the recursive SAH build with the bucket-index and
// bucket-initialisation defects fixed (bvh.rs:149-155, 203-204), the DFS
// flatten (bvh.rs:249-269), and the reference-order closest-hit and any-hit
// traversals (bvh.rs:281-423) with per-ray visit counters.
#pragma once

#include <vector>

#include "shapes.h"

namespace orc {

// One primitive of the scene: triangles first (ids [0, n_tris)), then spheres.
struct Scene {
    std::vector<TriVerts> tris;
    std::vector<uint8_t> tri_flip;  // reverse_orientation ^ transform_swaps_handedness
    std::vector<Sphere> spheres;
    size_t n_prims() const { return tris.size() + spheres.size(); }
    bool is_sphere(uint32_t id) const { return id >= tris.size(); }
    Bounds3 prim_bounds(uint32_t id) const {
        if (id < tris.size()) {
            const TriVerts& t = tris[id];
            return bunion(Bounds3(t.p0, t.p1), t.p2);  // triangle.rs:128-133
        }
        return sphere_world_bounds(spheres[id - tris.size()]);
    }
};

// bvh.rs:25-31 (the flattened node as the reference stores it).
struct RefNode {
    Bounds3 bounds;
    uint32_t primitive_offset = 0, second_child_offset = 0, count = 0, axis = 0;
};

struct RefBVH {
    std::vector<RefNode> nodes;
    std::vector<uint32_t> ordered;  // ordered_primitives as original primitive ids
};

RefBVH build_bvh(const Scene& scene);

struct Counters {
    uint32_t nodes = 0, tris = 0, spheres = 0;
};

struct Hit {
    uint32_t prim = 0xFFFFFFFFu;
    Float t = 0, b1 = 0, b2 = 0;
};

// bvh.rs:281-352. On return ray->t_max holds the closest accepted t (as
// GeometricPrimitive::intersect leaves it, geometric.rs:32).
bool bvh_intersect(const Scene& s, const RefBVH& bvh, Ray* ray, Hit* hit, Isect* it, Counters* c);
// bvh.rs:354-423; the hit records the first primitive that reported a hit.
bool bvh_intersect_test(const Scene& s, const RefBVH& bvh, const Ray& ray, Hit* hit, Counters* c);

}  // namespace orc
