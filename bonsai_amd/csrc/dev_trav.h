// Device functions of the pair-record traversal (pair records: both child boxes in the parent),
// shared by the traversal kernels (traverse.hip) and the persistent path kernel (path.hip). One
// lane traces one ray in exactly the order of accelerators/bvh.rs:281-423 (near child first by
// the sign of the ray direction on the split axis, the far child's box re-tested against the
// current t_max when it is popped, leaves in slot order with t_max shrinking on every accepted
// hit), so the accepted primitive, t and barycentrics are the reference's.
#pragma once

#include <hip/hip_runtime.h>

#include "bg_internal.h"
#include "dev_geom.h"

namespace bg {
namespace dev {

// Per-lane state of the pair-record traversal.
struct PairLane {
    RayState rs;
    // the record or leaf the lane is at; kRefNone once its traversal has ended (a leaf ref is
    // below it: primitive counts are held under 2^31 - 1 at scene build)
    uint32_t ref;
    uint32_t sp, base;  // stack top, and the first entry still in the LDS ring (older ones spilled)
    int best;           // slot of the accepted primitive, -1 for none
    float bt, bb1, bb2;
};

// A float in LDS, by address space.
typedef __attribute__((address_space(3))) float LdsF;

struct PairCounts {
    uint64_t nodes = 0, tris = 0, sph = 0;
};

// Starts ray `ray` on this lane at the scene's virtual root record.
__device__ __forceinline__ void pair_start(const TraverseArgs& a, uint64_t ray, PairLane& L) {
    const float4 r0 = a.rays[2 * ray];
    const float4 r1 = a.rays[2 * ray + 1];
    L.rs.ox = r0.x; L.rs.oy = r0.y; L.rs.oz = r0.z; L.rs.t_max = r0.w;
    L.rs.dx = r1.x; L.rs.dy = r1.y; L.rs.dz = r1.z;
    ray_setup(L.rs);
    L.best = -1;
    L.bt = 0.0f; L.bb1 = 0.0f; L.bb2 = 0.0f;
    L.sp = 0; L.base = 0;
    L.ref = a.roots ? a.roots[ray] : 0u;
}

__device__ __forceinline__ void pair_store(const TraverseArgs& a, uint64_t ray, const PairLane& L) {
    uint4 out;
    out.x = L.best >= 0 ? a.slot_prim[L.best] : 0xFFFFFFFFu;
    out.y = __float_as_uint(L.bt);
    out.z = __float_as_uint(L.bb1);
    out.w = __float_as_uint(L.bb2);
    a.hits[ray] = out;
}

__device__ __forceinline__ uint32_t popc64(uint64_t m) {
    // (two 32-bit counts keep the comparisons of counts in scalar registers)
    return __builtin_popcount((uint32_t)m) + __builtin_popcount((uint32_t)(m >> 32));
}

// One wavefront-wide step. Every step advances the lanes at an interior record; the lanes at a
// leaf (m_leaf, from the caller's ballot, as m_int) wait until they are at least as many as the
// interior lanes, so the triangle code runs for many lanes at once instead of in nearly every
// step. Each lane still takes its own nodes and leaves in the reference's order, so hits and
// visit counts are unchanged.
template <int BLOCK, int STACK, bool ANY, bool COUNT, bool SPH>
__device__ __forceinline__ void pair_step(const TraverseArgs& a, PairLane& L, uint2* __restrict__ ring,
                                          uint2* __restrict__ spill, const LdsF* __restrict__ top,
                                          uint32_t n_top, uint64_t m_leaf, uint64_t m_int, PairCounts& c) {
    constexpr uint32_t kMask = STACK - 1;
    const bool at_leaf = (int32_t)L.ref < (int32_t)kRefNone;
    const bool at_int = (int32_t)L.ref >= 0;
    const bool leaf_step = popc64(m_leaf) >= popc64(m_int);
    bool next = false;  // this lane needs the next pushed entry
    if (leaf_step) {
        if (at_leaf) {
            const uint2 lf = a.leaves[L.ref & ~kRefLeaf];
            const uint32_t off = lf.x, count = lf.y >> 1;
            const bool mixed = (lf.y & 1u) != 0;
            for (uint32_t s = off; s < off + count; ++s) {
                float t, b1, b2;
                bool hit;
                uint32_t pid = 0;
                if (SPH && mixed && (pid = a.slot_prim[s]) >= a.n_tris) {
                    if (COUNT) ++c.sph;
                    const SphereDev& sp_ = a.spheres[pid - a.n_tris];
                    SphereHit h;
                    hit = sphere_hit(L.rs, sp_.w2o, sp_.radius, sp_.z_min, sp_.z_max, sp_.phi_max, &h);
                    if (hit) {
                        t = h.t;
                        b1 = h.phi / sp_.phi_max;
                        const SphereAux ax = a.sphere_aux[pid - a.n_tris];
                        float cz = h.pz / sp_.radius;
                        cz = cz < -1.0f ? -1.0f : (cz > 1.0f ? 1.0f : cz);
                        const float theta = acosf(cz);
                        b2 = (theta - ax.theta_min) / (ax.theta_max - ax.theta_min);
                    }
                } else {
                    if (COUNT) ++c.tris;
                    hit = tri_hit(L.rs, a.tri_verts + 9 * (uint64_t)s, &t, &b1, &b2);
                }
                if (hit) {
                    L.best = (int)s;
                    L.bt = t;
                    if (ANY) {
                        L.bb1 = 0.0f; L.bb2 = 0.0f;
                        L.ref = kRefNone;
                        break;
                    }
                    L.bb1 = b1; L.bb2 = b2;
                    L.rs.t_max = t;
                }
            }
            next = L.ref != kRefNone;
        }
    }
    if (at_int) {
        // the record, from the LDS copy of the top levels where it is one of them
        float4 q0, q1, q2;
        uint4 q3;
        if (L.ref < n_top) {
            // (an LDS-typed pointer: the two branches then stay apart, LDS reads here and global
            // loads below, instead of one generic load that would take both paths for every lane)
            const LdsF* T = top + 16 * L.ref;
            q0 = make_float4(T[0], T[1], T[2], T[3]);
            q1 = make_float4(T[4], T[5], T[6], T[7]);
            q2 = make_float4(T[8], T[9], T[10], T[11]);
            q3 = make_uint4(__float_as_uint(T[12]), __float_as_uint(T[13]), __float_as_uint(T[14]), __float_as_uint(T[15]));
        } else {
            const float4* P = a.pairs + 4 * (uint64_t)L.ref;
            q0 = P[0]; q1 = P[1]; q2 = P[2];
            q3 = *reinterpret_cast<const uint4*>(P + 3);
        }
        const bool swap = (L.rs.negbits >> q3.z) & 1u;  // near child by the split axis (3: none)
        const uint32_t rn = swap ? q3.y : q3.x, rf = swap ? q3.x : q3.y;
        float t0, t1;
        const bool e0 = slab_enter(L.rs, q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, &t0);
        const bool e1 = slab_enter(L.rs, q1.z, q1.w, q2.x, q2.y, q2.z, q2.w, &t1);
        // each child's box admitted at the current t_max (entered, entry below t_max)
        const bool h0 = e0 && t0 < L.rs.t_max, h1 = e1 && t1 < L.rs.t_max;
        const bool hn = swap ? h1 : h0, hf = swap ? h0 : h1;
        if (COUNT) ++c.nodes;  // the near child's box test
        // counting pushes every far child, with a NaN entry for a box not entered (NaN < t_max
        // is false when it is popped); otherwise only admitted far children are pushed
        if (rf != kRefNone && (COUNT || hf)) {
            if (L.sp - L.base == STACK) {
                spill[L.base] = ring[(L.base & kMask) * BLOCK];
                ++L.base;
            }
            const float tf = swap ? t0 : t1;
            const bool ef = swap ? e0 : e1;
            ring[(L.sp & kMask) * BLOCK] = make_uint2(rf, __float_as_uint(COUNT && !ef ? __builtin_nanf("") : tf));
            ++L.sp;
        }
        if (hn && rn != kRefNone) L.ref = rn;
        else next = true;
    }
    if (next) {
        // the next pushed entry that the current t_max still admits
        bool found = false;
        while (L.sp > 0) {
            --L.sp;
            if (L.sp < L.base) {  // the ring is empty: the entry comes back from the spill row
                ring[(L.sp & kMask) * BLOCK] = spill[L.sp];
                L.base = L.sp;
            }
            const uint2 e = ring[(L.sp & kMask) * BLOCK];
            if (COUNT) ++c.nodes;  // the far child's box test, where the reference makes it
            if (__uint_as_float(e.y) < L.rs.t_max) {
                L.ref = e.x;
                found = true;
                break;
            }
        }
        if (!found) L.ref = kRefNone;
    }
}

// Records of the top levels a block holds in LDS: BLOCK / 16 (16 for 256 threads: the virtual root
// and the BVH's top four levels, 1 KB beside the 16 KB stack ring of an 8-entry stack). Measured on
// the 2^24-ray sets (256 threads, 8-entry stack, 8 blocks per CU): 16 records 5,117 Mrays/s
// against 5,005 with none; 64 records (4 KB) served more loads from LDS (incoherent closest-hit
// 4.03 -> 3.99 ms) but, at 20 KB per block, cost the coherent sets more (2.89 -> 3.00 ms):
// 5,085 Mrays/s.
// The sphere-capable kernels hold none: the LDS branch of the record fetch took the closest-hit
// one from 96 to 98 VGPRs, past the 5-waves-per-SIMD limit (1080p path tracer 0.144 -> 0.148 s).
template <int BLOCK, bool SPH>
constexpr uint32_t kBlockTop = SPH ? 0 : BLOCK / 16;

// Copies the first n_top records (the top levels of the tree, build_pairs numbering) into LDS, one
// float4 per thread. Every thread of the block calls it once, before any traversal.
template <int BLOCK, bool SPH>
__device__ __forceinline__ uint32_t load_top(const TraverseArgs& a, LdsF* __restrict__ top) {
    if (kBlockTop<BLOCK, SPH> == 0) return 0;
    const uint32_t n_top = a.n_top < kBlockTop<BLOCK, SPH> ? a.n_top : kBlockTop<BLOCK, SPH>;
    if (threadIdx.x < 4 * n_top) {
        const float4 v = a.pairs[threadIdx.x];
        LdsF* d = top + 4 * threadIdx.x;
        d[0] = v.x; d[1] = v.y; d[2] = v.z; d[3] = v.w;
    }
    __syncthreads();
    return n_top;
}

template <bool COUNT>
__device__ __forceinline__ void flush_counts(const TraverseArgs& a, const PairCounts& c) {
    if (COUNT) {
        atomicAdd(&a.counters[0], (unsigned long long)c.nodes);
        atomicAdd(&a.counters[1], (unsigned long long)c.tris);
        atomicAdd(&a.counters[2], (unsigned long long)c.sph);
    }
}

// Traces ray i of `a` on this lane when `live` (its hit stored; nothing written otherwise) with
// wavefront-wide steps until no lane of the wavefront has a ray left. Every lane of the wavefront
// calls it together (the steps take the wavefront's ballots).
template <int BLOCK, int STACK, bool ANY, bool SPH>
__device__ __forceinline__ void trace_lane(const TraverseArgs& a, uint64_t i, bool live, uint2* __restrict__ ring,
                                           uint2* __restrict__ spill, PairCounts& cnt) {
    PairLane L;
    L.ref = kRefNone;
    if (live) pair_start(a, i, L);
    for (;;) {
        const uint64_t m_leaf = __builtin_amdgcn_ballot_w64((int32_t)L.ref < (int32_t)kRefNone);
        const uint64_t m_int = __builtin_amdgcn_ballot_w64((int32_t)L.ref >= 0);
        if ((m_leaf | m_int) == 0) break;
        pair_step<BLOCK, STACK, ANY, false, SPH>(a, L, ring, spill, nullptr, 0u, m_leaf, m_int, cnt);
    }
    if (live) pair_store(a, i, L);
}

}  // namespace dev
}  // namespace bg
