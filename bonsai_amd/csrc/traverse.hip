// Closest-hit and any-hit BVH traversal kernels for gfx950.
//
// One lane traces one ray through the flattened BVH in exactly the order of
// accelerators/bvh.rs:281-423: each visited node's box is tested with the
// ray's current t_max; interior nodes push the far child and descend to the
// near one (chosen by the direction sign on the split axis); leaves test
// their primitives in slot order and, for closest hit, shrink t_max on every
// accepted hit (primitives/geometric.rs:32). Because each lane replays that
// sequence exactly, the accepted primitive, t and barycentrics equal the
// reference's, ties included.
//
// The per-lane traversal stack lives in LDS, laid out [entry][lane] so lanes
// of a wavefront hit distinct banks. It is a ring of `STACK` entries; on
// overflow its oldest entry moves to a per-thread global spill row (sized
// from the tree height at scene build), so any depth is handled.
#include <hip/hip_runtime.h>

#include "bg_internal.h"
#include "dev_geom.h"
#include "dev_trav.h"

namespace bg {
namespace {

using namespace dev;

template <int BLOCK, int STACK, bool ANY, bool COUNT, bool SPH, bool QUEUE = false>
__global__ __launch_bounds__(BLOCK) void k_traverse(TraverseArgs a) {
    __shared__ uint32_t lds[STACK * BLOCK];
    const int tid = threadIdx.x;
    const uint64_t gtid = (uint64_t)blockIdx.x * BLOCK + tid;
    const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
    uint32_t* __restrict__ spill = a.spill + gtid * a.spill_depth;
    uint32_t* __restrict__ ring = lds + tid;
    constexpr uint32_t kMask = STACK - 1;
    uint64_t c_nodes = 0, c_tris = 0, c_sph = 0;

    // Rays come in grid-stride order, or (work queue) in chunks of 64 x kChunk that each
    // wavefront takes from an atomic counter when it has finished its previous chunk.
    constexpr uint32_t kChunk = 16;
    const int lane = tid & 63;
    constexpr bool queued = QUEUE;
    uint64_t qbase = 0, ray = gtid;
    uint32_t qk = kChunk;
    auto next_queued = [&]() {
        if (qk == kChunk) {
            unsigned long long b0 = 0;
            if (lane == 0) b0 = atomicAdd(a.queue, 64ull * kChunk);
            qbase = __shfl(b0, 0);
            qk = 0;
        }
        ray = qbase + (uint64_t)qk * 64 + lane;
        ++qk;
    };
    if (queued) next_queued();
    while (queued ? qbase < a.n_rays : ray < a.n_rays) {
        // (queued: lanes past the end of the last chunk skip the body; every lane of the
        // wavefront still reaches next_queued() together)
        if (ray < a.n_rays) {
        const float4 r0 = a.rays[2 * ray];
        const float4 r1 = a.rays[2 * ray + 1];
        RayState rs;
        rs.ox = r0.x; rs.oy = r0.y; rs.oz = r0.z; rs.t_max = r0.w;
        rs.dx = r1.x; rs.dy = r1.y; rs.dz = r1.z;
        ray_setup(rs);

        int best = -1;
        float bt = 0.0f, bb1 = 0.0f, bb2 = 0.0f;
        uint32_t sp = 0, base = 0;  // logical stack size, first logical entry held in LDS
        uint32_t cur = a.roots ? a.roots[ray] : 0u;
        for (;;) {
            const float4 n0 = a.nodes[2 * cur];
            const float4 n1 = a.nodes[2 * cur + 1];
            if (COUNT) ++c_nodes;
            bool descend = false;
            if (slab(rs, n0.x, n0.y, n0.z, n0.w, n1.x, n1.y)) {
                const uint32_t off = __float_as_uint(n1.z);
                const uint32_t meta = __float_as_uint(n1.w);
                const uint32_t count = meta >> 3;
                if (count > 0) {
                    const bool mixed = (meta & 4u) != 0;
                    bool done = false;
                    for (uint32_t s = off; s < off + count; ++s) {
                        float t, b1, b2;
                        bool hit;
                        uint32_t pid = 0;
                        if (SPH && mixed && (pid = a.slot_prim[s]) >= a.n_tris) {
                            if (COUNT) ++c_sph;
                            const SphereDev& sp_ = a.spheres[pid - a.n_tris];
                            SphereHit h;
                            hit = sphere_hit(rs, sp_.w2o, sp_.radius, sp_.z_min, sp_.z_max, sp_.phi_max, &h);
                            if (hit) {
                                t = h.t;
                                b1 = h.phi / sp_.phi_max;
                                const SphereAux ax = a.sphere_aux[pid - a.n_tris];
                                float c = h.pz / sp_.radius;
                                c = c < -1.0f ? -1.0f : (c > 1.0f ? 1.0f : c);
                                const float theta = acosf(c);
                                b2 = (theta - ax.theta_min) / (ax.theta_max - ax.theta_min);
                            }
                        } else {
                            if (COUNT) ++c_tris;
                            hit = tri_hit(rs, a.tri_verts + 9 * (uint64_t)s, &t, &b1, &b2);
                        }
                        if (hit) {
                            best = (int)s;
                            bt = t;
                            if (ANY) {
                                bb1 = 0.0f; bb2 = 0.0f;
                                done = true;
                                break;
                            }
                            bb1 = b1; bb2 = b2;
                            rs.t_max = t;
                        }
                    }
                    if (ANY && done) break;
                } else {
                    // interior: push far child, go to near child (bvh.rs:331-339)
                    const uint32_t axis = meta & 3u;
                    const int neg = axis == 0 ? rs.neg0() : (axis == 1 ? rs.neg1() : rs.neg2());
                    const uint32_t far = neg ? cur + 1 : off;
                    cur = neg ? off : cur + 1;
                    if (sp - base == STACK) {
                        spill[base] = ring[(base & kMask) * BLOCK];
                        ++base;
                    }
                    ring[(sp & kMask) * BLOCK] = far;
                    ++sp;
                    descend = true;
                }
            }
            if (!descend) {
                if (sp == 0) break;
                --sp;
                if (sp >= base) {
                    cur = ring[(sp & kMask) * BLOCK];
                } else {
                    cur = spill[sp];
                    base = sp;
                }
            }
        }
        uint4 out;
        out.x = best >= 0 ? a.slot_prim[best] : 0xFFFFFFFFu;
        out.y = __float_as_uint(bt);
        out.z = __float_as_uint(bb1);
        out.w = __float_as_uint(bb2);
        a.hits[ray] = out;
        }
        if (queued) next_queued();
        else ray += stride;
    }
    if (COUNT) {
        atomicAdd(&a.counters[0], (unsigned long long)c_nodes);
        atomicAdd(&a.counters[1], (unsigned long long)c_tris);
        atomicAdd(&a.counters[2], (unsigned long long)c_sph);
    }
}

// Same traversal with per-lane refill: the loop advances one node per
// iteration, and a lane whose ray is finished stores its hit and starts the
// next ray of its grid-stride sequence at once instead of idling until every
// lane of its wavefront has finished. The per-ray visit sequence is unchanged.
template <int BLOCK, int STACK, bool ANY, bool COUNT, bool SPH>
__global__ __launch_bounds__(BLOCK) void k_traverse_refill(TraverseArgs a) {
    __shared__ uint32_t lds[STACK * BLOCK];
    const int tid = threadIdx.x;
    const uint64_t gtid = (uint64_t)blockIdx.x * BLOCK + tid;
    const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
    uint32_t* __restrict__ spill = a.spill + gtid * a.spill_depth;
    uint32_t* __restrict__ ring = lds + tid;
    constexpr uint32_t kMask = STACK - 1;
    uint64_t c_nodes = 0, c_tris = 0, c_sph = 0;

    uint64_t ray = gtid;
    if (ray < a.n_rays) {
        RayState rs;
        int best = -1;
        float bt = 0.0f, bb1 = 0.0f, bb2 = 0.0f;
        uint32_t sp = 0, base = 0, cur = 0;
        auto start = [&](uint64_t r) {
            const float4 r0 = a.rays[2 * r];
            const float4 r1 = a.rays[2 * r + 1];
            rs.ox = r0.x; rs.oy = r0.y; rs.oz = r0.z; rs.t_max = r0.w;
            rs.dx = r1.x; rs.dy = r1.y; rs.dz = r1.z;
            ray_setup(rs);
            best = -1;
            bt = 0.0f; bb1 = 0.0f; bb2 = 0.0f;
            sp = 0; base = 0; cur = a.roots ? a.roots[r] : 0u;
        };
        start(ray);
        for (;;) {
            const float4 n0 = a.nodes[2 * cur];
            const float4 n1 = a.nodes[2 * cur + 1];
            if (COUNT) ++c_nodes;
            bool descend = false, finished = false;
            if (slab(rs, n0.x, n0.y, n0.z, n0.w, n1.x, n1.y)) {
                const uint32_t off = __float_as_uint(n1.z);
                const uint32_t meta = __float_as_uint(n1.w);
                const uint32_t count = meta >> 3;
                if (count > 0) {
                    const bool mixed = (meta & 4u) != 0;
                    for (uint32_t s = off; s < off + count; ++s) {
                        float t, b1, b2;
                        bool hit;
                        uint32_t pid = 0;
                        if (SPH && mixed && (pid = a.slot_prim[s]) >= a.n_tris) {
                            if (COUNT) ++c_sph;
                            const SphereDev& sp_ = a.spheres[pid - a.n_tris];
                            SphereHit h;
                            hit = sphere_hit(rs, sp_.w2o, sp_.radius, sp_.z_min, sp_.z_max, sp_.phi_max, &h);
                            if (hit) {
                                t = h.t;
                                b1 = h.phi / sp_.phi_max;
                                const SphereAux ax = a.sphere_aux[pid - a.n_tris];
                                float c = h.pz / sp_.radius;
                                c = c < -1.0f ? -1.0f : (c > 1.0f ? 1.0f : c);
                                const float theta = acosf(c);
                                b2 = (theta - ax.theta_min) / (ax.theta_max - ax.theta_min);
                            }
                        } else {
                            if (COUNT) ++c_tris;
                            hit = tri_hit(rs, a.tri_verts + 9 * (uint64_t)s, &t, &b1, &b2);
                        }
                        if (hit) {
                            best = (int)s;
                            bt = t;
                            if (ANY) {
                                bb1 = 0.0f; bb2 = 0.0f;
                                finished = true;
                                break;
                            }
                            bb1 = b1; bb2 = b2;
                            rs.t_max = t;
                        }
                    }
                } else {
                    const uint32_t axis = meta & 3u;
                    const int neg = axis == 0 ? rs.neg0() : (axis == 1 ? rs.neg1() : rs.neg2());
                    const uint32_t far = neg ? cur + 1 : off;
                    cur = neg ? off : cur + 1;
                    if (sp - base == STACK) {
                        spill[base] = ring[(base & kMask) * BLOCK];
                        ++base;
                    }
                    ring[(sp & kMask) * BLOCK] = far;
                    ++sp;
                    descend = true;
                }
            }
            if (!descend && !finished) {
                if (sp == 0) {
                    finished = true;
                } else {
                    --sp;
                    if (sp >= base) {
                        cur = ring[(sp & kMask) * BLOCK];
                    } else {
                        cur = spill[sp];
                        base = sp;
                    }
                }
            }
            if (finished) {
                uint4 out;
                out.x = best >= 0 ? a.slot_prim[best] : 0xFFFFFFFFu;
                out.y = __float_as_uint(bt);
                out.z = __float_as_uint(bb1);
                out.w = __float_as_uint(bb2);
                a.hits[ray] = out;
                ray += stride;
                if (ray >= a.n_rays) break;
                start(ray);
            }
        }
    }
    if (COUNT) {
        atomicAdd(&a.counters[0], (unsigned long long)c_nodes);
        atomicAdd(&a.counters[1], (unsigned long long)c_tris);
        atomicAdd(&a.counters[2], (unsigned long long)c_sph);
    }
}

// Chunk of the refill kernels' ray sequences: rays / wavefronts rounded down to a
// power of two, within [64, 512], so that every wavefront has rays and a
// wavefront's consecutive rays come from one stretch of the ray array.
__host__ __device__ inline uint32_t refill_shift(uint64_t n_rays, uint64_t n_waves) {
    const uint64_t per = n_rays / (n_waves ? n_waves : 1);
    uint32_t sh = 6;
    while (sh < 9 && (2ull << sh) <= per) ++sh;
    return sh;
}

// Pair-record traversal (PairRecord): at an interior record both child boxes
// are tested with one 64-byte fetch; the near child (by the sign of the ray
// direction on the split axis, bvh.rs:331-339) is entered at once if its box
// is hit, the far child is pushed with its entry distance and entered when
// popped if that distance is still below t_max. This is the reference's visit
// order and its box results, without fetching a far child only to find its box
// missed. Leaves are processed in wavefront-wide leaf steps (pair_step). With
// COUNT, box tests are counted where the reference makes them (the near child
// on arrival, the far child when popped), so the totals equal the node-record
// kernels'. Degenerate triangles are collapsed in the slot vertices at scene
// build, so the leaf test skips the reference's per-test degeneracy check. The
// triangle-only instantiations (not counting, not queued) are held to 64 VGPRs
// (8 waves per SIMD, no spills); the others would spill there and run faster at
// the register count the compiler picks (5-7 waves). Rays are taken in grid-stride order, or
// (QUEUE) in chunks of 64 x 16 that each wavefront takes from an atomic counter;
// with an index list (a.idx) logical ray j is ray a.idx[j], for j < *a.n_idx.
template <int BLOCK, int STACK, bool ANY, bool COUNT, bool SPH, bool QUEUE>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(!SPH && !COUNT && !QUEUE ? 8 : 1, 8)))
void k_trav_pairs(TraverseArgs a) {
    __shared__ uint2 lds[STACK * BLOCK];
    __shared__ float4 top_[SPH ? 1 : 4 * kBlockTop<BLOCK, SPH>];
    LdsF* top = (LdsF*)top_;
    const uint32_t n_top = load_top<BLOCK, SPH>(a, top);
    const int tid = threadIdx.x;
    const uint64_t gtid = (uint64_t)blockIdx.x * BLOCK + tid;
    const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
    uint2* __restrict__ spill = reinterpret_cast<uint2*>(a.spill) + gtid * a.spill_depth;
    uint2* __restrict__ ring = lds + tid;
    PairCounts cnt;

    constexpr uint32_t kChunk = 16;
    const int lane = tid & 63;
    uint64_t qbase = 0, ray = gtid;
    uint32_t qk = kChunk;
    auto next_queued = [&]() {
        if (qk == kChunk) {
            unsigned long long b0 = 0;
            if (lane == 0) b0 = atomicAdd(a.queue, 64ull * kChunk);
            qbase = __shfl(b0, 0);
            qk = 0;
        }
        ray = qbase + (uint64_t)qk * 64 + lane;
        ++qk;
    };
    const uint64_t n_rays = a.n_idx ? (uint64_t)*a.n_idx : a.n_rays;
    if (QUEUE) next_queued();
    while (QUEUE ? qbase < n_rays : ray < n_rays) {
        if (ray < n_rays) {
            const uint64_t rid = a.idx ? (uint64_t)a.idx[ray] : ray;
            PairLane L;
            pair_start(a, rid, L);
            for (;;) {
                // (single compares, so the ballots stay scalar)
                const uint64_t m_leaf = __builtin_amdgcn_ballot_w64((int32_t)L.ref < (int32_t)kRefNone);
                const uint64_t m_int = __builtin_amdgcn_ballot_w64((int32_t)L.ref >= 0);
                if ((m_leaf | m_int) == 0) break;
                pair_step<BLOCK, STACK, ANY, COUNT, SPH>(a, L, ring, spill, top, n_top, m_leaf, m_int, cnt);
            }
            pair_store(a, rid, L);
        }
        if (QUEUE) next_queued();
        else ray += stride;
    }
    flush_counts<COUNT>(a, cnt);
}

// The same traversal with batched refill: each wavefront owns a sequence of
// rays (chunks of 2^chunk_shift consecutive rays, chunk c belonging to
// wavefront c mod wavefronts), and once at least refill_min of its lanes have
// finished their rays (or all have), those lanes store their hits and take the
// next rays of the sequence in lane order. Lanes therefore do not idle from
// their ray's end until the slowest lane of the wavefront ends; neighbouring
// rays stay together in a wavefront (consecutive rays of a chunk). Each lane
// still traces each ray exactly as above, so hits and visit counts are the
// same; only which lane traces which ray, and when, differ. The triangle-only
// instantiations are held to 64 VGPRs (8 waves per SIMD, no spills).
template <int BLOCK, int STACK, bool ANY, bool COUNT, bool SPH>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(!SPH ? 8 : 1, 8)))
void k_trav_pairs_refill(TraverseArgs a) {
    __shared__ uint2 lds[STACK * BLOCK];
    __shared__ float4 top_[SPH ? 1 : 4 * kBlockTop<BLOCK, SPH>];
    LdsF* top = (LdsF*)top_;
    const uint32_t n_top = load_top<BLOCK, SPH>(a, top);
    const int tid = threadIdx.x;
    const uint64_t gtid = (uint64_t)blockIdx.x * BLOCK + tid;
    uint2* __restrict__ spill = reinterpret_cast<uint2*>(a.spill) + gtid * a.spill_depth;
    uint2* __restrict__ ring = lds + tid;
    PairCounts cnt;

    const uint64_t n_waves = (uint64_t)gridDim.x * (BLOCK / 64);
    // With an index list the live-ray count is on the device: the chunk and the refill threshold
    // follow from it here (below four rays per lane a wavefront refills only when all its lanes
    // are free, taking its rays in rounds as the grid-stride kernel does).
    const uint64_t n_rays = a.n_idx ? (uint64_t)*a.n_idx : a.n_rays;
    const uint32_t shift = a.n_idx ? refill_shift(n_rays, n_waves) : a.chunk_shift;
    const uint32_t refill_min = a.n_idx && n_rays < 256 * n_waves ? 64u : a.refill_min;
    const uint64_t chunk = 1ull << shift;
    const uint64_t wave = (uint64_t)blockIdx.x * (BLOCK / 64) + (uint32_t)__builtin_amdgcn_readfirstlane(tid >> 6);
    // position s of this wavefront's sequence -> logical ray (increasing in s)
    auto seq_ray = [&](uint64_t s) { return (((s >> shift) * n_waves + wave) << shift) + (s & (chunk - 1)); };
    uint64_t pos = 0;               // next sequence position (the same in every lane)
    uint64_t ray = ~0ull;           // this lane's ray, ~0 for none
    PairLane L;
    L.ref = kRefNone;
    // (one path back to the loop head: a refill that re-entered the loop head through its own
    // edge made the compiler copy the whole lane state between two register sets on every step.
    // One refill serves every free lane the sequence still has rays for, and pair_start puts a
    // lane at a record, never at kRefNone, so after a refill a lane is free only once the
    // sequence has no rays left: the refill condition cannot hold again right after it, and a
    // wavefront with no lane at a record has finished.)
    for (;;) {
        uint64_t m_leaf = __builtin_amdgcn_ballot_w64((int32_t)L.ref < (int32_t)kRefNone);
        uint64_t m_int = __builtin_amdgcn_ballot_w64((int32_t)L.ref >= 0);
        const uint64_t m_free = ~(m_leaf | m_int);
        const uint32_t n_free = popc64(m_free);
        if (n_free != 0 && (n_free >= refill_min || n_free == 64) && seq_ray(pos) < n_rays) {
            if (L.ref == kRefNone) {
                if (ray != ~0ull) pair_store(a, ray, L);
                const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(m_free >> 32),
                                                                __builtin_amdgcn_mbcnt_lo((uint32_t)m_free, 0u));
                const uint64_t r = seq_ray(pos + rank);
                ray = ~0ull;
                if (r < n_rays) {
                    ray = a.idx ? (uint64_t)a.idx[r] : r;
                    pair_start(a, ray, L);
                }
            }
            pos += n_free;
            m_leaf = __builtin_amdgcn_ballot_w64((int32_t)L.ref < (int32_t)kRefNone);
            m_int = __builtin_amdgcn_ballot_w64((int32_t)L.ref >= 0);
        }
        if ((m_leaf | m_int) == 0) break;
        pair_step<BLOCK, STACK, ANY, COUNT, SPH>(a, L, ring, spill, top, n_top, m_leaf, m_int, cnt);
    }
    if (ray != ~0ull) pair_store(a, ray, L);
    flush_counts<COUNT>(a, cnt);
}

// Persistent traversal with per-lane refill from an atomic work queue (variant 5): the grid stays
// resident; each wavefront takes runs of kQueueRun consecutive logical rays from a.queue (one
// atomic per run, zero at launch) and, as in k_trav_pairs_refill, once refill_min of its lanes
// (or all) have finished, those lanes store their hits and take the next rays of the wavefront's
// current run (fetching the next run when it cannot serve them all). Unlike the static sequences
// of k_trav_pairs_refill, a wavefront whose rays end early takes more, so scene batches with
// uneven rays and compacted launches (an index list, the ray count on the device) balance to the
// end. Each lane traces each ray exactly as above: the hits are the same.
constexpr uint64_t kQueueRun = 256;  // rays per atomic fetch: 4 per lane
template <int BLOCK, int STACK, bool ANY, bool SPH>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(!SPH ? 8 : 1, 8)))
void k_trav_pairs_queue(TraverseArgs a) {
    __shared__ uint2 lds[STACK * BLOCK];
    __shared__ float4 top_[SPH ? 1 : 4 * kBlockTop<BLOCK, SPH>];
    LdsF* top = (LdsF*)top_;
    const uint32_t n_top = load_top<BLOCK, SPH>(a, top);
    const int tid = threadIdx.x;
    const uint64_t gtid = (uint64_t)blockIdx.x * BLOCK + tid;
    uint2* __restrict__ spill = reinterpret_cast<uint2*>(a.spill) + gtid * a.spill_depth;
    uint2* __restrict__ ring = lds + tid;
    PairCounts cnt;

    const int lane = tid & 63;
    const uint64_t n = a.n_idx ? (uint64_t)*a.n_idx : a.n_rays;
    // the wavefront's unassigned logical rays [cur, end) of its current run, and whether the queue
    // has run dry (all wavefront-uniform)
    uint64_t cur = 0, end = 0;
    bool dry = false;
    uint64_t ray = ~0ull;  // this lane's ray (physical id), ~0 for none
    PairLane L;
    L.ref = kRefNone;
    // (one path back to the loop head, as in k_trav_pairs_refill; lanes a refill leaves free are
    // served at a later step, and a wavefront with no lane at a record after a refill has no rays
    // left: a lane that got none drew a position at or past the end of the queue)
    for (;;) {
        uint64_t m_leaf = __builtin_amdgcn_ballot_w64((int32_t)L.ref < (int32_t)kRefNone);
        uint64_t m_int = __builtin_amdgcn_ballot_w64((int32_t)L.ref >= 0);
        const uint64_t m_free = ~(m_leaf | m_int);
        const uint32_t n_free = popc64(m_free);
        if (n_free != 0 && (n_free >= a.refill_min || n_free == 64) && !(dry && cur >= end)) {
            const uint64_t rem = end - cur;
            uint64_t run = 0;
            bool fetched = false;
            if (rem < n_free && !dry) {
                unsigned long long b0 = 0;
                if (lane == 0) b0 = atomicAdd(a.queue, (unsigned long long)kQueueRun);
                run = __builtin_amdgcn_readfirstlane((uint32_t)b0) | ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(b0 >> 32)) << 32);
                if (run >= n) dry = true;
                else fetched = true;
            }
            if (L.ref == kRefNone) {
                if (ray != ~0ull) pair_store(a, ray, L);
                const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(m_free >> 32),
                                                                __builtin_amdgcn_mbcnt_lo((uint32_t)m_free, 0u));
                const uint64_t r = rank < rem ? cur + rank : (fetched ? run + (rank - rem) : ~0ull);
                ray = ~0ull;
                if (r < n) {
                    ray = a.idx ? (uint64_t)a.idx[r] : r;
                    pair_start(a, ray, L);
                }
            }
            if (rem >= n_free) {
                cur += n_free;
            } else if (fetched) {
                cur = run + (n_free - rem);
                end = run + kQueueRun;
            } else {
                cur = end;
            }
            m_leaf = __builtin_amdgcn_ballot_w64((int32_t)L.ref < (int32_t)kRefNone);
            m_int = __builtin_amdgcn_ballot_w64((int32_t)L.ref >= 0);
        }
        if ((m_leaf | m_int) == 0) break;
        pair_step<BLOCK, STACK, ANY, false, SPH>(a, L, ring, spill, top, n_top, m_leaf, m_int, cnt);
    }
    if (ray != ~0ull) pair_store(a, ray, L);
}

uint32_t refill_chunk_shift(uint64_t n_rays, uint64_t n_waves) { return refill_shift(n_rays, n_waves); }

// Variant 4 refills on triangle scenes: over an index list (the kernel takes the live count from
// the device), or over every ray where there are at least four rays per lane.
bool refill_applies(const TraverseArgs& a, uint64_t n_waves) {
    return a.spheres == nullptr && (a.idx != nullptr || a.n_rays >= 256 * n_waves);
}

template <int BLOCK, int STACK, bool ANY, bool COUNT>
void launch4(const TraverseArgs& a_in, int variant, int grid, hipStream_t s) {
    // scenes without spheres use instantiations with the sphere code compiled out
    // variants 0 and 2: pair records (2: work queue, a.queue set and zeroed by the caller; counting
    // runs take rays in grid-stride order); 4: pair records with batched refill; 1: node records
    // with per-lane refill; 3: node records
    TraverseArgs a = a_in;
    // variant 5 needs a queue counter and is not a counting kernel: otherwise it runs as 4
    if (variant == 5 && (COUNT || a.queue == nullptr || a.pairs == nullptr)) variant = 4;
    if (variant == 5) {
        launch_zero_u64(a.queue, s);
        if (a.spheres != nullptr) hipLaunchKernelGGL((k_trav_pairs_queue<BLOCK, STACK, ANY, true>), dim3(grid), dim3(BLOCK), 0, s, a);
        else hipLaunchKernelGGL((k_trav_pairs_queue<BLOCK, STACK, ANY, false>), dim3(grid), dim3(BLOCK), 0, s, a);
        return;
    }
    if (variant == 2 && a.queue != nullptr && !COUNT) launch_zero_u64(a.queue, s);
    // compacted launches (an index list) run the grid-stride, queue or (triangle scenes) refill
    // pair kernel
    if (a.idx != nullptr && a.pairs != nullptr && variant != 2 && !(variant == 4 && a.spheres == nullptr)) variant = 0;
    const bool q = variant == 2 && a.queue != nullptr && !COUNT;
    const bool pairs = (variant == 0 || variant == 2 || variant == 4) && a.pairs != nullptr;
    // Variant 4 refills only where lanes have rays to refill from: at least four rays per lane
    // and no spheres (measured: 1.13x on the 2^24-ray triangle intersection sets; on launches of
    // about one ray per lane, and on the sphere-and-triangle workloads even at eight rays per
    // lane, the grid-stride kernel was 3-18% faster, so those run as 0).
    const uint64_t n_waves = (uint64_t)grid * (BLOCK / 64);
    const bool refill = variant == 4 && pairs && refill_applies(a, n_waves);
    if (refill) a.chunk_shift = refill_chunk_shift(a.n_rays, n_waves);
    if (a.spheres != nullptr) {
        if (refill) hipLaunchKernelGGL((k_trav_pairs_refill<BLOCK, STACK, ANY, COUNT, true>), dim3(grid), dim3(BLOCK), 0, s, a);
        else if (variant == 1) hipLaunchKernelGGL((k_traverse_refill<BLOCK, STACK, ANY, COUNT, true>), dim3(grid), dim3(BLOCK), 0, s, a);
        else if (pairs && q) hipLaunchKernelGGL((k_trav_pairs<BLOCK, STACK, ANY, false, true, true>), dim3(grid), dim3(BLOCK), 0, s, a);
        else if (pairs) hipLaunchKernelGGL((k_trav_pairs<BLOCK, STACK, ANY, COUNT, true, false>), dim3(grid), dim3(BLOCK), 0, s, a);
        else hipLaunchKernelGGL((k_traverse<BLOCK, STACK, ANY, COUNT, true>), dim3(grid), dim3(BLOCK), 0, s, a);
    } else {
        if (refill) hipLaunchKernelGGL((k_trav_pairs_refill<BLOCK, STACK, ANY, COUNT, false>), dim3(grid), dim3(BLOCK), 0, s, a);
        else if (variant == 1) hipLaunchKernelGGL((k_traverse_refill<BLOCK, STACK, ANY, COUNT, false>), dim3(grid), dim3(BLOCK), 0, s, a);
        else if (pairs && q) hipLaunchKernelGGL((k_trav_pairs<BLOCK, STACK, ANY, false, false, true>), dim3(grid), dim3(BLOCK), 0, s, a);
        else if (pairs) hipLaunchKernelGGL((k_trav_pairs<BLOCK, STACK, ANY, COUNT, false, false>), dim3(grid), dim3(BLOCK), 0, s, a);
        else hipLaunchKernelGGL((k_traverse<BLOCK, STACK, ANY, COUNT, false>), dim3(grid), dim3(BLOCK), 0, s, a);
    }
}

template <int BLOCK, int STACK>
hipError_t launch2(const TraverseArgs& a, bool any, bool count, int variant, int grid, hipStream_t s) {
    if (count) {
        if (any) launch4<BLOCK, STACK, true, true>(a, variant, grid, s);
        else launch4<BLOCK, STACK, false, true>(a, variant, grid, s);
    } else {
        if (any) launch4<BLOCK, STACK, true, false>(a, variant, grid, s);
        else launch4<BLOCK, STACK, false, false>(a, variant, grid, s);
    }
    return hipGetLastError();
}

template <int BLOCK>
hipError_t launch1(const TraverseArgs& a, bool any, bool count, int stack, int variant, int grid, hipStream_t s) {
    switch (stack) {
        case 8: return launch2<BLOCK, 8>(a, any, count, variant, grid, s);
        case 16: return launch2<BLOCK, 16>(a, any, count, variant, grid, s);
        case 32: return launch2<BLOCK, 32>(a, any, count, variant, grid, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

bool traverse_compacts(const TraverseArgs& a, const LaunchShape& shape, int grid) {
    // (below about 1M slots the three compaction launches cost more than the empty slots)
    return shape.compact && a.pairs != nullptr && a.n_rays >= (1ull << 20);
}

bool traverse_shape_supported(const LaunchShape& s) {
    const bool b = s.block == 64 || s.block == 128 || s.block == 256;
    const bool st = s.stack == 8 || s.stack == 16 || s.stack == 32;
    return b && st && s.blocks_per_cu >= 0 && s.variant >= 0 && s.variant <= 5;
}

hipError_t launch_traverse(const TraverseArgs& a, bool any_hit, bool count, const LaunchShape& shape, int grid,
                           hipStream_t stream) {
    if (a.n_rays == 0) return hipSuccess;
    switch (shape.block) {
        case 64: return launch1<64>(a, any_hit, count, shape.stack, shape.variant, grid, stream);
        case 128: return launch1<128>(a, any_hit, count, shape.stack, shape.variant, grid, stream);
        case 256: return launch1<256>(a, any_hit, count, shape.stack, shape.variant, grid, stream);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace bg
