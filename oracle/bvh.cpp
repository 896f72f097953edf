// Oracle (test infrastructure only): see bvh.h.
#include "bvh.h"

#include <algorithm>
#include <memory>

namespace orc {

namespace {

const size_t MAX_PRIMITIVES_IN_NODE = 256;  // bvh.rs:17
const int PARTITION_BUCKET_SIZE = 16;       // bvh.rs:18

// bvh.rs:33-37, 435-443
struct PrimInfo {
    uint32_t index;
    Bounds3 bounds;
    V3 centroid;
};

// bvh.rs:39-45
struct BuildNode {
    Bounds3 bounds;
    std::unique_ptr<BuildNode> children[2];
    uint32_t split_axis = 0, offset = 0, count = 0;
};

struct Bucket {
    Float count = 0.0f;
    Bounds3 bounds;
};

// Bucket index with the evident pbrt-v3 intent: floor(16 * offset), clamped.
inline int bucket_of(const Bounds3& cb, const V3& c, int dim) {
    int b = (int)((Float)PARTITION_BUCKET_SIZE * boffset(cb, c)[dim]);
    if (b == PARTITION_BUCKET_SIZE) b = PARTITION_BUCKET_SIZE - 1;
    return b;
}

// itertools 0.10 `partition`: two-ended scan that swaps a failing front
// element with the last passing back element. Returns the split index.
template <class P>
size_t two_ended_partition(PrimInfo* a, size_t n, P pred) {
    size_t split = 0, front = 0, back = n;
    while (front < back) {
        PrimInfo* f = &a[front++];
        if (!pred(*f)) {
            bool swapped = false;
            while (front < back) {
                PrimInfo* b = &a[--back];
                if (pred(*b)) { std::swap(*f, *b); swapped = true; break; }
            }
            if (!swapped) return split;
        }
        split += 1;
    }
    return split;
}

// bvh.rs:91-247
std::unique_ptr<BuildNode> build(PrimInfo* info, size_t size, size_t* count, std::vector<uint32_t>* ordered) {
    std::unique_ptr<BuildNode> node(new BuildNode());
    *count += 1;
    Bounds3 bounds;
    for (size_t i = 0; i < size; ++i) bounds = bunion(bounds, info[i].bounds);

    auto make_leaf = [&]() {
        node->offset = (uint32_t)ordered->size();
        for (size_t i = 0; i < size; ++i) ordered->push_back(info[i].index);
        node->count = (uint32_t)size;
        node->bounds = bounds;
    };

    if (size == 1) { make_leaf(); return node; }
    Bounds3 cb;
    for (size_t i = 0; i < size; ++i) cb = bunion(cb, info[i].centroid);
    int dim = maximum_extent(cb);
    if (cb.mx[dim] == cb.mn[dim]) { make_leaf(); return node; }

    size_t mid = size / 2;
    if (size <= 2) {
        // select_nth_unstable_by(mid) over total_cmp on the split-axis centroid.
        if (info[1].centroid[dim] < info[0].centroid[dim]) std::swap(info[0], info[1]);
    } else {
        Bucket buckets[PARTITION_BUCKET_SIZE];
        for (size_t i = 0; i < size; ++i) {
            int b = bucket_of(cb, info[i].centroid, dim);
            buckets[b].count += 1.0f;
            buckets[b].bounds = bunion(buckets[b].bounds, info[i].bounds);
        }
        Float cost[PARTITION_BUCKET_SIZE - 1];
        for (int i = 0; i < PARTITION_BUCKET_SIZE - 1; ++i) {
            Bounds3 b0, b1;
            Float count0 = 0.0f, count1 = 0.0f;
            for (int j = 0; j <= i; ++j) { b0 = bunion(b0, buckets[j].bounds); count0 += buckets[j].count; }
            for (int j = i + 1; j < PARTITION_BUCKET_SIZE; ++j) { b1 = bunion(b1, buckets[j].bounds); count1 += buckets[j].count; }
            cost[i] = 1.0f + (count0 * surface_area(b0) + count1 * surface_area(b1)) / surface_area(bounds);
        }
        Float min_cost = cost[0];
        int split = 0;
        for (int i = 1; i < PARTITION_BUCKET_SIZE - 1; ++i)
            if (cost[i] < min_cost) { min_cost = cost[i]; split = i; }
        Float leaf_cost = (Float)size;
        if (size > MAX_PRIMITIVES_IN_NODE || min_cost < leaf_cost) {
            mid = two_ended_partition(info, size, [&](const PrimInfo& p) { return bucket_of(cb, p.centroid, dim) <= split; });
        } else {
            make_leaf();
            return node;
        }
    }
    node->split_axis = (uint32_t)dim;
    node->children[0] = build(info, mid, count, ordered);
    node->children[1] = build(info + mid, size - mid, count, ordered);
    node->bounds = bunion(node->children[0]->bounds, node->children[1]->bounds);
    return node;
}

// bvh.rs:249-269
uint32_t flatten(std::vector<RefNode>* nodes, const BuildNode* n, uint32_t* offset) {
    uint32_t cur = (*offset)++;
    (*nodes)[cur].bounds = n->bounds;
    if (n->count > 0) {
        (*nodes)[cur].primitive_offset = n->offset;
        (*nodes)[cur].count = n->count;
    } else {
        (*nodes)[cur].axis = n->split_axis;
        (*nodes)[cur].count = 0;
        flatten(nodes, n->children[0].get(), offset);
        (*nodes)[cur].second_child_offset = flatten(nodes, n->children[1].get(), offset);
    }
    return cur;
}

}  // namespace

// bvh.rs:53-89
RefBVH build_bvh(const Scene& scene) {
    RefBVH out;
    size_t n = scene.n_prims();
    if (n == 0) return out;
    std::vector<PrimInfo> info(n);
    for (size_t i = 0; i < n; ++i) {
        info[i].index = (uint32_t)i;
        info[i].bounds = scene.prim_bounds((uint32_t)i);
        info[i].centroid = 0.5f * info[i].bounds.mn + 0.5f * info[i].bounds.mx;
    }
    size_t total = 0;
    out.ordered.reserve(n);
    std::unique_ptr<BuildNode> root = build(info.data(), n, &total, &out.ordered);
    out.nodes.resize(total);
    uint32_t off = 0;
    flatten(&out.nodes, root.get(), &off);
    return out;
}

namespace {

inline bool prim_intersect(const Scene& s, uint32_t id, Ray* ray, Hit* hit, Isect* it, Counters* c) {
    Float t, b[3];
    if (id < s.tris.size()) {
        if (c) c->tris++;
        if (!tri_intersect(s.tris[id], *ray, s.tri_flip[id] != 0, &t, b, it)) return false;
        hit->b1 = b[1]; hit->b2 = b[2];
    } else {
        if (c) c->spheres++;
        if (!sphere_intersect(s.spheres[id - s.tris.size()], *ray, &t, b, it)) return false;
        hit->b1 = b[0]; hit->b2 = b[1];
    }
    ray->t_max = t;  // geometric.rs:32
    hit->prim = id;
    hit->t = t;
    return true;
}

inline bool prim_test(const Scene& s, uint32_t id, const Ray& ray, Float* t, Counters* c) {
    if (id < s.tris.size()) {
        if (c) c->tris++;
        return tri_intersect_test(s.tris[id], ray, t);
    }
    if (c) c->spheres++;
    return sphere_intersect_test(s.spheres[id - s.tris.size()], ray, t);
}

}  // namespace

bool bvh_intersect(const Scene& s, const RefBVH& bvh, Ray* ray, Hit* hit, Isect* it, Counters* c) {
    if (bvh.nodes.empty()) return false;
    bool any = false;
    V3 inv(1.0f / ray->d.x, 1.0f / ray->d.y, 1.0f / ray->d.z);
    int neg[3] = {inv.x < 0.0f, inv.y < 0.0f, inv.z < 0.0f};
    std::vector<uint32_t> stack(64);
    size_t to_visit = 0;
    uint32_t cur = 0;
    for (;;) {
        if (to_visit >= stack.size()) stack.resize(stack.size() + 64);
        const RefNode& node = bvh.nodes[cur];
        if (c) c->nodes++;
        if (intersect_range_precomputed(node.bounds, *ray, inv, neg)) {
            if (node.count > 0) {
                for (uint32_t i = 0; i < node.count; ++i)
                    if (prim_intersect(s, bvh.ordered[node.primitive_offset + i], ray, hit, it, c)) any = true;
                if (to_visit == 0) break;
                cur = stack[--to_visit];
            } else {
                if (neg[node.axis]) {
                    stack[to_visit] = cur + 1;
                    cur = node.second_child_offset;
                } else {
                    stack[to_visit] = node.second_child_offset;
                    cur = cur + 1;
                }
                to_visit++;
            }
        } else {
            if (to_visit == 0) break;
            cur = stack[--to_visit];
        }
    }
    return any;
}

bool bvh_intersect_test(const Scene& s, const RefBVH& bvh, const Ray& ray, Hit* hit, Counters* c) {
    if (bvh.nodes.empty()) return false;
    V3 inv(1.0f / ray.d.x, 1.0f / ray.d.y, 1.0f / ray.d.z);
    int neg[3] = {inv.x < 0.0f, inv.y < 0.0f, inv.z < 0.0f};
    std::vector<uint32_t> stack(64);
    size_t to_visit = 0;
    uint32_t cur = 0;
    for (;;) {
        if (to_visit >= stack.size()) stack.resize(stack.size() + 64);
        const RefNode& node = bvh.nodes[cur];
        if (c) c->nodes++;
        if (intersect_range_precomputed(node.bounds, ray, inv, neg)) {
            if (node.count > 0) {
                for (uint32_t i = 0; i < node.count; ++i) {
                    uint32_t id = bvh.ordered[node.primitive_offset + i];
                    Float t = 0.0f;
                    if (prim_test(s, id, ray, &t, c)) {
                        if (hit) { hit->prim = id; hit->t = t; hit->b1 = 0.0f; hit->b2 = 0.0f; }
                        return true;
                    }
                }
                if (to_visit == 0) break;
                cur = stack[--to_visit];
            } else {
                if (neg[node.axis]) {
                    stack[to_visit] = cur + 1;
                    cur = node.second_child_offset;
                } else {
                    stack[to_visit] = node.second_child_offset;
                    cur = cur + 1;
                }
                to_visit++;
            }
        } else {
            if (to_visit == 0) break;
            cur = stack[--to_visit];
        }
    }
    return false;
}

}  // namespace orc
