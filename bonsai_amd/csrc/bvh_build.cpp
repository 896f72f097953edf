// Host SAH BVH builder (accelerators/bvh.rs:53-269).
//
// The reference builds a recursive node tree and flattens it depth-first.
// Here each subtree is built straight into its own flat node array; the two
// children of a split are built independently (concurrently near the root)
// and concatenated left-then-right behind their parent, which is exactly the
// depth-first layout of bvh.rs:249-269. Bucket selection follows bvh.rs:141-225
// with the two defects fixed: the bucket index is floor(16 * offset) and the
// buckets start empty. Float arithmetic is kept in the reference's order so the
// layout is reproducible on any host.
#include <algorithm>
#include <limits>
#include <cfloat>
#include <cstring>
#include <future>

#include "bg_internal.h"

namespace bg {
namespace {

constexpr size_t kMaxPrimsInNode = 256;  // bvh.rs:17
constexpr int kBuckets = 16;             // bvh.rs:18

// Rust f32::min / f32::max (NaN operand -> other operand; ties -> second).
inline float rmin(float a, float b) {
    if (a != a) return b;
    if (b != b) return a;
    return a < b ? a : b;
}
inline float rmax(float a, float b) {
    if (a != a) return b;
    if (b != b) return a;
    return a > b ? a : b;
}

struct Box {
    float mn[3], mx[3];
    static Box empty() {
        Box b;
        for (int i = 0; i < 3; ++i) { b.mn[i] = FLT_MAX; b.mx[i] = -FLT_MAX; }
        return b;
    }
    void grow(const Box& o) {
        for (int i = 0; i < 3; ++i) { mn[i] = rmin(mn[i], o.mn[i]); mx[i] = rmax(mx[i], o.mx[i]); }
    }
    void grow(const float p[3]) {
        for (int i = 0; i < 3; ++i) { mn[i] = rmin(mn[i], p[i]); mx[i] = rmax(mx[i], p[i]); }
    }
    // bounds3.rs:160-163
    float area() const {
        float dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
        return 2.0f * (dx * dy + dx * dz + dy * dz);
    }
    // bounds3.rs:175-184
    int widest() const {
        float dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
        if (dx > dy && dx > dz) return 0;
        if (dy > dz) return 1;
        return 2;
    }
};

struct Ref {
    uint32_t id;
    Box box;
    float c[3];
};

struct Sub {
    std::vector<bg_bvh_node> nodes;   // offsets relative to this subtree
    std::vector<uint32_t> order;      // primitive ids, relative slot order
    uint32_t depth = 0;               // pending-far-child depth
    uint32_t max_leaf = 0;
    bool sphere_leafy = false;
};

struct Builder {
    const std::vector<uint8_t>& is_sphere;
    int parallel_levels;

    // bounds3.rs:194-206 offset along one axis, then floor(16 * offset).
    static int bucket(const Box& cb, const float c[3], int dim) {
        float o = c[dim] - cb.mn[dim];
        if (cb.mx[dim] > cb.mn[dim]) o = o * (1.0f / (cb.mx[dim] - cb.mn[dim]));
        int b = (int)((float)kBuckets * o);
        return b == kBuckets ? kBuckets - 1 : b;
    }

    void leaf(Ref* r, size_t n, const Box& bounds, Sub* out) const {
        bg_bvh_node nd;
        std::memcpy(nd.bmin, bounds.mn, sizeof(nd.bmin));
        std::memcpy(nd.bmax, bounds.mx, sizeof(nd.bmax));
        nd.offset = 0;
        bool sph = false;
        for (size_t i = 0; i < n; ++i) {
            out->order.push_back(r[i].id);
            sph = sph || is_sphere[r[i].id];
        }
        nd.meta = ((uint32_t)n << 3) | (sph ? 4u : 0u);
        out->nodes.push_back(nd);
        out->depth = 0;
        out->max_leaf = (uint32_t)n;
    }

    // itertools::partition (two-ended swap), as used at bvh.rs:202-213.
    template <class P>
    static size_t partition(Ref* a, size_t n, P pred) {
        size_t split = 0, front = 0, back = n;
        while (front < back) {
            Ref* f = &a[front++];
            if (!pred(*f)) {
                bool swapped = false;
                while (front < back) {
                    Ref* b = &a[--back];
                    if (pred(*b)) { std::swap(*f, *b); swapped = true; break; }
                }
                if (!swapped) return split;
            }
            ++split;
        }
        return split;
    }

    void build(Ref* r, size_t n, int level, Sub* out) const {
        Box bounds = Box::empty();
        for (size_t i = 0; i < n; ++i) bounds.grow(r[i].box);
        if (n == 1) { leaf(r, n, bounds, out); return; }
        Box cb = Box::empty();
        for (size_t i = 0; i < n; ++i) cb.grow(r[i].c);
        int dim = cb.widest();
        if (cb.mx[dim] == cb.mn[dim]) { leaf(r, n, bounds, out); return; }

        size_t mid = n / 2;
        if (n <= 2) {
            if (r[1].c[dim] < r[0].c[dim]) std::swap(r[0], r[1]);
        } else {
            float cnt[kBuckets];
            Box bb[kBuckets];
            for (int b = 0; b < kBuckets; ++b) { cnt[b] = 0.0f; bb[b] = Box::empty(); }
            for (size_t i = 0; i < n; ++i) {
                int b = bucket(cb, r[i].c, dim);
                cnt[b] += 1.0f;
                bb[b].grow(r[i].box);
            }
            // Prefix/suffix sweeps give the same unions and sums as the
            // reference's per-split loops (same operands, same order).
            float cost[kBuckets - 1];
            Box left = Box::empty();
            float c0 = 0.0f;
            float lsa[kBuckets - 1], lcnt[kBuckets - 1];
            for (int i = 0; i < kBuckets - 1; ++i) {
                left.grow(bb[i]);
                c0 += cnt[i];
                lsa[i] = left.area();
                lcnt[i] = c0;
            }
            float total_area = bounds.area();
            for (int i = 0; i < kBuckets - 1; ++i) {
                Box right = Box::empty();
                float c1 = 0.0f;
                for (int j = i + 1; j < kBuckets; ++j) { right.grow(bb[j]); c1 += cnt[j]; }
                cost[i] = 1.0f + (lcnt[i] * lsa[i] + c1 * right.area()) / total_area;
            }
            float min_cost = cost[0];
            int split = 0;
            for (int i = 1; i < kBuckets - 1; ++i)
                if (cost[i] < min_cost) { min_cost = cost[i]; split = i; }
            if (n > kMaxPrimsInNode || min_cost < (float)n) {
                mid = partition(r, n, [&](const Ref& p) { return bucket(cb, p.c, dim) <= split; });
            } else {
                leaf(r, n, bounds, out);
                return;
            }
        }

        Sub a, b;
        if (level < parallel_levels && n > 4096) {
            auto fut = std::async(std::launch::async, [&] { build(r, mid, level + 1, &a); });
            build(r + mid, n - mid, level + 1, &b);
            fut.get();
        } else {
            build(r, mid, level + 1, &a);
            build(r + mid, n - mid, level + 1, &b);
        }
        // Parent, then the first child's subtree, then the second's.
        bg_bvh_node nd;
        for (int i = 0; i < 3; ++i) {
            nd.bmin[i] = rmin(a.nodes[0].bmin[i], b.nodes[0].bmin[i]);
            nd.bmax[i] = rmax(a.nodes[0].bmax[i], b.nodes[0].bmax[i]);
        }
        const uint32_t second = 1u + (uint32_t)a.nodes.size();
        nd.offset = second;
        nd.meta = (uint32_t)dim;
        out->nodes.reserve(1 + a.nodes.size() + b.nodes.size());
        out->nodes.push_back(nd);
        append(a, 1u, 0u, out);
        append(b, second, (uint32_t)a.order.size(), out);
        out->order.reserve(a.order.size() + b.order.size());
        out->order.insert(out->order.end(), a.order.begin(), a.order.end());
        out->order.insert(out->order.end(), b.order.begin(), b.order.end());
        out->depth = 1u + std::max(a.depth, b.depth);
        out->max_leaf = std::max(a.max_leaf, b.max_leaf);
    }

    static void append(const Sub& s, uint32_t node_base, uint32_t slot_base, Sub* out) {
        for (const bg_bvh_node& nd : s.nodes) {
            bg_bvh_node c = nd;
            if ((c.meta >> 3) != 0) c.offset += slot_base;  // leaf: first slot
            else c.offset += node_base;                     // interior: second child
            out->nodes.push_back(c);
        }
    }
};

}  // namespace

void build_bvh(const std::vector<PrimBounds>& bounds, const std::vector<uint8_t>& is_sphere, int threads,
               HostBVH* out) {
    out->nodes.clear();
    out->slot_prim.clear();
    out->max_depth = 0;
    out->max_leaf = 0;
    const size_t n = bounds.size();
    if (n == 0) return;
    std::vector<Ref> refs(n);
    for (size_t i = 0; i < n; ++i) {
        refs[i].id = (uint32_t)i;
        std::memcpy(refs[i].box.mn, bounds[i].mn, sizeof(refs[i].box.mn));
        std::memcpy(refs[i].box.mx, bounds[i].mx, sizeof(refs[i].box.mx));
        // bvh.rs:438: centroid = 0.5 * min + 0.5 * max
        for (int k = 0; k < 3; ++k) refs[i].c[k] = 0.5f * bounds[i].mn[k] + 0.5f * bounds[i].mx[k];
    }
    int levels = 0;
    while ((1 << levels) < threads && levels < 6) ++levels;
    Builder b{is_sphere, levels};
    Sub root;
    b.build(refs.data(), n, 0, &root);
    out->nodes = std::move(root.nodes);
    out->slot_prim = std::move(root.order);
    out->max_depth = root.depth;
    out->max_leaf = root.max_leaf;
}


void build_pairs(const HostBVH& bvh, std::vector<PairRecord>* pairs, std::vector<uint32_t>* leaves) {
    const size_t n = bvh.nodes.size();
    std::vector<uint32_t> ref(n);
    uint32_t n_int = 0, n_leaf = 0;
    for (size_t i = 0; i < n; ++i) {
        if ((bvh.nodes[i].meta >> 3) == 0) ref[i] = 1 + n_int++;
        else ref[i] = kRefLeaf | n_leaf++;
    }
    pairs->assign(1 + n_int, PairRecord{});
    leaves->assign(2 * (size_t)n_leaf, 0u);
    const float qnan = std::numeric_limits<float>::quiet_NaN();
    auto put_box = [](PairRecord& r, int k, const bg_bvh_node& nd) {
        for (int c = 0; c < 3; ++c) {
            r.b[6 * k + c] = nd.bmin[c];
            r.b[6 * k + 3 + c] = nd.bmax[c];
        }
    };
    // the virtual root: child 0 is the BVH root, child 1 a NaN box that no ray enters
    PairRecord& root = (*pairs)[0];
    for (int c = 0; c < 12; ++c) root.b[c] = qnan;
    root.ref0 = root.ref1 = kRefNone;
    root.axis = 3;
    if (n > 0) {
        put_box(root, 0, bvh.nodes[0]);
        root.ref0 = ref[0];
    }
    for (size_t i = 0; i < n; ++i) {
        const bg_bvh_node& nd = bvh.nodes[i];
        const uint32_t count = nd.meta >> 3;
        if (count > 0) {
            const uint32_t id = ref[i] & ~kRefLeaf;
            (*leaves)[2 * id] = nd.offset;
            (*leaves)[2 * id + 1] = count << 1 | ((nd.meta >> 2) & 1u);
            continue;
        }
        PairRecord& r = (*pairs)[ref[i]];
        const size_t c0 = i + 1, c1 = nd.offset;
        put_box(r, 0, bvh.nodes[c0]);
        put_box(r, 1, bvh.nodes[c1]);
        r.ref0 = ref[c0];
        r.ref1 = ref[c1];
        r.axis = nd.meta & 3u;
        r.pad = 0;
    }
    // Renumber: the first kTopRecords records breadth-first from the virtual root (child 0 before
    // child 1), the others in their depth-first order after them. Only the numbering changes; every
    // record keeps its boxes, children and split axis.
    const size_t n_rec = pairs->size();
    std::vector<uint32_t> id(n_rec, kRefNone), bfs;
    bfs.reserve(std::min<size_t>(n_rec, 2 * (size_t)kTopRecords + 2));
    bfs.push_back(0);
    uint32_t next = 0;
    for (size_t head = 0; head < bfs.size() && next < kTopRecords; ++head) {
        const uint32_t r = bfs[head];
        id[r] = next++;
        for (uint32_t c : {(*pairs)[r].ref0, (*pairs)[r].ref1})
            if (c != kRefNone && !(c & kRefLeaf)) bfs.push_back(c);
    }
    for (size_t r = 0; r < n_rec; ++r)
        if (id[r] == kRefNone) id[r] = next++;
    auto remap = [&](uint32_t c) { return (c == kRefNone || (c & kRefLeaf)) ? c : id[c]; };
    std::vector<PairRecord> out(n_rec);
    for (size_t r = 0; r < n_rec; ++r) {
        PairRecord q = (*pairs)[r];
        q.ref0 = remap(q.ref0);
        q.ref1 = remap(q.ref1);
        out[id[r]] = q;
    }
    pairs->swap(out);
}

}  // namespace bg
