// Wavefront path integrator for gfx950 (integrators/path.rs:51-146 with
// base/integrator.rs:35-133, 309-437).
//
// A pass holds n_paths camera samples (window pixels x samples of the pass),
// path-major state in HBM. Each bounce runs: extend (closest-hit traversal of
// every live path's ray), shade (surface interaction, emission, light and
// BSDF sampling for next-event estimation, BSDF sampling of the continuation,
// Russian roulette), the two next-event queries (any-hit shadow rays and
// closest-hit MIS-branch rays, both plain traversal launches), and resolve
// (adds the estimate to the path's radiance). The film kernel then gathers
// every sample whose box-filter footprint covers a pixel, in a fixed order,
// so the film is deterministic without atomics.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "bg_internal.h"
#include "dev_camera.h"
#include "dev_guide.h"
#include "dev_light.h"
#include "dev_scan.h"
#include "dev_shade.h"
#include "dev_trav.h"

namespace bg {
namespace {

using namespace dev;
using dev::block_exclusive;

// state.y flags; kQuiet: dead since an earlier bounce, with its rays killed and no NEE pending, so
// shading has nothing to write for it. Every ray a live path issues has t_max > 0 (a camera ray
// +inf less its error offset, an extension ray +inf or the re-intersection distance, a shadow
// ray 1 - 1e-4, an MIS ray 1 - 1e-4 or +inf), so the kernels that read a hit read only hits of
// rays the ray masks mark, and a compacted launch need not write misses for the others.
constexpr uint32_t kAlive = 1u, kSpecFlag = 2u, kQuiet = 4u, kItemShift = 16u;
// nee_a.w flags
constexpr uint32_t kHasNee = 1u, kHasLight = 2u, kHasBsdf = 4u;

// A path the network is evaluated for at this bounce: alive, with a hit.
__device__ __forceinline__ bool guide_live(const PathBuffers& b, uint64_t i) {
    return (b.state[i].y & kAlive) && b.hit[i].x != 0xFFFFFFFFu;
}

// Training-record slot of path i at this bounce: every rec_stride-th path up to
// rec_paths of them records one slot per bounce < max_depth, path-major.
__device__ __forceinline__ bool rec_slot(const RenderParamsDev& p, uint64_t i, uint32_t bounce, uint64_t* r) {
    if (!p.guided || p.rec_paths == 0 || bounce >= p.max_depth || i % p.rec_stride != 0) return false;
    const uint64_t k = i / p.rec_stride;
    if (k >= p.rec_paths) return false;
    *r = k * p.max_depth + bounce;
    return true;
}

// Batch item owning path i (items sorted by path_base) or pixel t (by pix_base).
__device__ __forceinline__ uint32_t item_of_path(const RenderParamsDev& p, uint64_t i) {
    uint32_t lo = 0, hi = p.n_batch;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (p.batch[mid].path_base <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}
__device__ __forceinline__ uint32_t item_of_pixel(const RenderParamsDev& p, uint64_t t) {
    uint32_t lo = 0, hi = p.n_batch;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (p.batch[mid].pix_base <= t) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Stores, from lane 0 of each wavefront, one word per ray set k < NSETS: bit l for slot
// 64 w + l set where bit k of that lane's `live` is (the ray written there can hit). Every lane
// of the wavefront must call it; lanes past the slot count pass 0.
template <int NSETS>
__device__ __forceinline__ void store_ray_masks(const PathBuffers& b, uint64_t i, uint32_t live) {
#pragma unroll
    for (int k = 0; k < NSETS; ++k) {
        const uint64_t m = __ballot((live >> k) & 1u);
        if ((threadIdx.x & 63) == 0) b.wmask[k][i >> 6] = m;
    }
}

// One path's camera ray and state; returns the ray's t_max.
__device__ __forceinline__ float init_path(const RenderParamsDev& p, const PathBuffers& b, uint64_t i) {
    const CameraDev* cam = &p.cam;
    uint64_t seed = p.seed, li = i;
    uint32_t x0 = p.x0, y0 = p.y0, x1 = p.x1, y1 = p.y1, item = 0;
    if (p.n_batch) {  // scene batch: the item's camera, seed and whole film
        item = item_of_path(p, i);
        const BatchItemDev& it = p.batch[item];
        cam = &it.cam;
        seed = it.seed;
        x0 = 0; y0 = 0; x1 = it.cam.width; y1 = it.cam.height;
        li = i - it.path_base;
        b.root[i] = it.root;
    }
    const uint32_t first = p.pass_first ? *p.pass_first : p.first_sample;
    const uint32_t ww = x1 - x0, wh = y1 - y0;
    const uint64_t npix = (uint64_t)ww * wh;
    const uint32_t local = (uint32_t)(li % npix);
    const uint32_t s = first + (uint32_t)(li / npix);
    const uint32_t px = x0 + local % ww, py = y0 + local / ww;
    const uint64_t key = stream_key(seed, (uint64_t)py * cam->width + px, s);
    uint32_t dim = 0;
    // sampler.rs:14-20: film = pixel + get_2d, lens = get_2d, time = get_1d
    const float fx = (float)px + draw(key, dim);
    const float fy = (float)py + draw(key, dim);
    draw(key, dim);
    draw(key, dim);
    const float tu = draw(key, dim);
    const CamRay r = generate_ray(*cam, fx, fy, tu);
    b.ray[2 * i] = make_float4(r.ox, r.oy, r.oz, r.t_max);
    b.ray[2 * i + 1] = make_float4(r.dx, r.dy, r.dz, r.time);
    b.beta[i] = make_float4(1.0f, 1.0f, 1.0f, __uint_as_float(1u));  // .w: rays issued (camera ray)
    b.L[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    b.key[i] = make_uint2((uint32_t)key, (uint32_t)(key >> 32));
    // the batch item rides in the flags; a camera ray that can hit nothing (t_max NaN or <= 0, from a
    // degenerate camera transform) ends the path here, as its miss would at the first shading
    b.state[i] = make_uint2(dim, (r.t_max > 0.0f ? kAlive : 0u) | (item << kItemShift));
    b.film[i] = make_float2(fx, fy);
    return r.t_max;
}

__global__ __launch_bounds__(256) void k_init_paths(RenderParamsDev p, PathBuffers b) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const uint32_t live = i < p.n_paths && init_path(p, b, i) > 0.0f ? 1u : 0u;
    store_ray_masks<1>(b, i, live);
}

// One path's vertex; returns which of the rays it wrote can hit: bit 0 the extension ray, bit 1 the
// shadow ray, bit 2 the MIS-branch ray (a killed ray has t_max = -inf).
template <bool GUIDED, bool MULTI, bool SPL>
__device__ __forceinline__ uint32_t shade_path(const RenderParamsDev& p, const ShadeSceneDev& sc, const PathBuffers& b,
                                               uint32_t bounce, uint64_t i) {
    uint2 st = b.state[i];
    uint32_t flags = st.y;
    uint64_t rec = 0;
    const bool recording = rec_slot(p, i, bounce, &rec);
    if (recording) b.rec_tgt[rec] = make_float4(0.0f, 0.0f, 1.0f, 0.0f);  // .w = 0: no sample at this slot
    if (flags & kQuiet) return 0u;
    b.nee_a[i].w = 0.0f;
    kill_ray(b.sray, i);
    kill_ray(b.mray, i);
    if (!(flags & kAlive)) {  // its extension ray was killed when it ended
        b.state[i] = make_uint2(st.x, kQuiet);
        return 0u;
    }
    const bool ref = p.mode == 0;
    float ext_t = -__builtin_inff(), shadow_t = -__builtin_inff(), mis_t = -__builtin_inff();
    const uint4 h = b.hit[i];
    const float4 r0 = b.ray[2 * i], r1 = b.ray[2 * i + 1];
    if (h.x == 0xFFFFFFFFu) {
        b.state[i] = make_uint2(st.x, 0u);
        kill_ray(b.ray, i);
        return 0u;
    }
    const uint32_t prim = h.x;
    Surf s;
    float t;
    if (!hit_surface(sc, r0, r1, prim, &s, &t)) {  // unreachable for a primitive the traversal accepted on this ray
        b.state[i] = make_uint2(st.x, 0u);
        kill_ray(b.ray, i);
        return 0u;
    }
    const uint64_t key = key_of(b.key[i]);
    uint32_t dim = st.x;
    const float4 b4 = b.beta[i];
    // rays this path has issued: low 16 bits all, high 16 bits any-hit (shadow) queries
    uint32_t nrays = __float_as_uint(b4.w);
    C3 beta = C3{b4.x, b4.y, b4.z};
    // radiance so far (.w: the continuation pdf of the previous vertex, guided only). The plain
    // estimator changes it here only where the ray hit an emitter, so only there does it read and
    // write it.
    float4 l4 = GUIDED ? b.L[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    C3 L = C3{l4.x, l4.y, l4.z};
    bool l_changed = GUIDED;
    const V rd = mkv(r1.x, r1.y, r1.z);

    // path.rs:78-86
    const uint2 info = sc.prim_info[prim];
    const int32_t hit_light = (int32_t)info.x;
    if (hit_light >= 0) {
        if (bounce == 0 || (flags & kSpecFlag)) {
            if (!GUIDED) {
                l4 = b.L[i];
                L = C3{l4.x, l4.y, l4.z};
                l_changed = true;
            }
            L = L + beta * emission(sc.lights[hit_light], s.n, -rd);
        } else if (GUIDED) {
            // guided estimator: emission reached by the continuation, MIS-weighted against light
            // sampling (power heuristic; both pdfs in solid angle, light pdf with its 1/n_lights pick)
            const LightDev& lt = sc.lights[hit_light];
            const float cosl = __builtin_fabsf(dot(s.n, rd));
            const float pc = l4.w;  // continuation pdf of the previous vertex
            if (cosl > 0.0f && pc > 0.0f) {
                const float pl = light_pdf_of_hit<SPL>(sc, lt, r0, t, cosl) / (float)sc.n_lights;
                L = L + beta * emission(lt, s.n, -rd) * power_heuristic(pc, pl);
            }
        }
    }
    bool alive = bounce < p.max_depth;
    float cont_pdf = 0.0f;
    if (alive) {
        // matte.rs:20-41 / conductor / plastic.rs:24-52; bsdf.rs:33-52
        const MaterialDev& m = sc.mats[info.y & ~kPrimFlip];
        const BsdfN<MULTI ? 2 : 1> bs = make_bsdf<MULTI ? 2 : 1>(m, s, ref);
        const uint32_t nonspec = kAll & ~kSpecular;
        // network mixture of this vertex (guided estimator with guide_alpha > 0)
        const float4* gm = nullptr;
        float alpha = 0.0f;
        // the network is sampled where the material's lobes are of the kinds selected (diffuse / glossy)
        bool lobe_ok = bs.n > 0;
        for (int k = 0; k < (MULTI ? 2 : 1); ++k)
            if (k < bs.n) lobe_ok = lobe_ok && ((bs.lobe(k).type & kDiffuse) ? (p.guide_lobes & 1u) : (p.guide_lobes & 2u));
        // past the guided bounces there is no network output (and no training record) for this vertex
        const uint32_t gslot = GUIDED && bounce < p.guide_bounces ? b.gslot[i] : 0xFFFFFFFFu;
        if (GUIDED && p.guide_alpha > 0.0f && lobe_ok && !bs.any_specular() && gslot != 0xFFFFFFFFu) {
            gm = reinterpret_cast<const float4*>(b.gout) + kMixRecord * (uint64_t)gslot;
            if (guide_ok(gm)) alpha = p.guide_alpha;
        }
        uint32_t light_base = 0, n_lights = sc.n_lights;
        if (p.n_batch) {  // the item's own lights
            const BatchItemDev& it = p.batch[flags >> kItemShift];
            light_base = it.light_base;
            n_lights = it.n_lights;
        }
        if (bs.ncomp(nonspec) > 0 && n_lights > 0) {
            // integrator.rs:309-328
            const float nl = (float)n_lights;
            const float u = draw(key, dim);
            const uint32_t idx = light_base + (uint32_t)rmin(u * nl, nl - 1.0f);
            const float light_pdf = 1.0f / nl;
            const float ul0 = draw(key, dim), ul1 = draw(key, dim);
            // the guided estimator has no separate BSDF-sampling branch (its continuation is that branch)
            const float us0 = GUIDED ? 0.0f : draw(key, dim), us1 = GUIDED ? 0.0f : draw(key, dim);
            const LightDev& light = sc.lights[idx];
            uint32_t nee = kHasNee;
            // diffuse.rs:134-167: a point on the light's triangle (triangle.rs:522-559)
            const LightPoint lpt = sample_light_point<SPL>(sc, light, s, ul0, ul1, ref);
            const float lpdf = lpt.pdf;
            const V lwi = lpt.wi;
            const C3 lrad = lpt.rad;
            float scattering_pdf = 0.0f;
            C3 term = mkc(0.0f);
            if (lpdf > 0.0f && !black(lrad)) {
                const C3 f = bs.f(s.wo, lwi, nonspec) * __builtin_fabsf(dot(lwi, bs.ns));
                scattering_pdf = bs.pdf(s.wo, lwi, nonspec);
                if (!black(f)) {
                    shadow_t = write_shadow_ray(b.sray, i, s, lpt);
                    float w;
                    if (GUIDED) {
                        const float pc = (1.0f - alpha) * scattering_pdf + (alpha > 0.0f ? alpha * guide_pdf(gm, lwi) : 0.0f);
                        w = power_heuristic(lpdf * light_pdf, pc);
                    } else {
                        w = power_heuristic(lpdf, scattering_pdf);
                    }
                    term = f * lrad * w / lpdf;
                    nee |= kHasLight;
                    nrays += 0x10001u;
                }
            }
            // BSDF-sampling branch (integrator.rs:379-434)
            const BxS bsm = GUIDED ? BxS{mkv(0, 0, 0), mkc(0.0f), 0.0f, 0} : bs.sample(s.wo, us0, us1, nonspec);
            const C3 fb = bsm.f * __builtin_fabsf(dot(bsm.wi, bs.ns));
            const float spdf = ref ? scattering_pdf : bsm.pdf;
            const V mdir = ref ? lwi : bsm.wi;  // the reference traces the light-sample direction
            if (!black(fb) && spdf > 0.0f) {
                const float lpp = light_point_pdf<SPL>(sc, light, s, mdir, ref);
                if (lpp != 0.0f) {
                    const float w = power_heuristic(spdf, lpp);
                    V o;
                    float tm;
                    spawn(s, mdir, ref, &o, &tm);
                    write_ray(b.mray, i, o, tm, mdir, s.time);
                    mis_t = tm;
                    const C3 kb = beta * (fb * (w / spdf) / light_pdf);
                    b.nee_b[i] = make_float4(kb.r, kb.g, kb.b, __uint_as_float(idx));
                    nee |= kHasBsdf;
                    nrays += 1u;
                }
            }
            const C3 ka = beta * (term / light_pdf);
            b.nee_a[i] = make_float4(ka.r, ka.g, ka.b, __uint_as_float(nee));
        }
        // path.rs:111-142; guided: one-sample MIS of the network's mixture and the BSDF
        V wi;
        C3 fs;
        float pdf;
        uint32_t stype = 0;
        if (GUIDED) {
            // the BSDF draws come first, so with guide_alpha == 0 the streams match the plain integrator's
            const float u0 = draw(key, dim), u1 = draw(key, dim);
            const float uc = p.guide_alpha > 0.0f ? draw(key, dim) : 1.0f;
            const float ul = p.guide_alpha > 0.0f ? draw(key, dim) : 0.0f;
            const float a = alpha;
            float pb;
            if (uc < a) {
                wi = guide_sample(gm, ul, u0, u1);
                fs = bs.f(-rd, wi, kAll);
                pb = bs.pdf(-rd, wi, kAll);
            } else {
                const BxS bsc = bs.sample(-rd, u0, u1, kAll);
                wi = bsc.wi;
                fs = bsc.f;
                pb = bsc.pdf;
                stype = bsc.type;
            }
            pdf = (1.0f - a) * pb + (a > 0.0f ? a * guide_pdf(gm, wi) : 0.0f);
            cont_pdf = pdf;
            if (recording && !black(fs) && pdf > 0.0f && gslot != 0xFFFFFFFFu) {
                const float4* gi = b.gin + 3 * (uint64_t)gslot;
                b.rec_in[3 * rec] = gi[0];
                b.rec_in[3 * rec + 1] = gi[1];
                b.rec_in[3 * rec + 2] = gi[2];
                b.rec_tgt[rec] = make_float4(wi.x, wi.y, wi.z, 1.0f);
                b.rec_aux[2 * rec] = make_float4(beta.r, beta.g, beta.b, 0.0f);
                b.rec_aux[2 * rec + 1] = make_float4(L.r, L.g, L.b, 0.0f);
            }
        } else {
            const float u0 = draw(key, dim), u1 = draw(key, dim);
            const BxS bsc = bs.sample(-rd, u0, u1, kAll);
            wi = bsc.wi;
            fs = bsc.f;
            pdf = bsc.pdf;
            stype = bsc.type;
        }
        if (black(fs) || !(pdf > 0.0f)) {
            alive = false;
        } else {
            beta = beta * (fs * __builtin_fabsf(dot(wi, bs.ns)) / pdf);
            flags = (stype & kSpecular) ? (flags | kSpecFlag) : (flags & ~kSpecFlag);
            if (cmax(beta) < p.rr && bounce >= p.rr_start) {
                const float q = rmax(1.0f - cmax(beta), 0.5f);
                if (draw(key, dim) < q) alive = false;
                else beta = beta / (1.0f - q);
            }
            // At bounce + 1 == max_depth the next intersection can add nothing (no specular lobes), so
            // stop; the guided estimator still adds that intersection's MIS-weighted emission.
            if (alive && bounce + 1 >= p.max_depth && !GUIDED) alive = false;
            if (alive) {
                if (ref) {
                    b.ray[2 * i].w = t;  // Q1: the camera ray is re-intersected with t_max = t_hit
                    ext_t = t;
                } else {
                    V o;
                    float tm;
                    spawn(s, wi, false, &o, &tm);
                    write_ray(b.ray, i, o, tm, wi, s.time);
                    ext_t = tm;
                }
            }
        }
    }
    if (!alive) kill_ray(b.ray, i);
    else nrays += 1u;
    b.beta[i] = make_float4(beta.r, beta.g, beta.b, __uint_as_float(nrays));
    if (l_changed) b.L[i] = make_float4(L.r, L.g, L.b, cont_pdf);  // .w: continuation pdf, for the next vertex's MIS
    b.state[i] = make_uint2(dim, alive ? (flags | kAlive) : 0u);
    return (alive && ext_t > 0.0f ? 1u : 0u) | (shadow_t > 0.0f ? 2u : 0u) | (mis_t > 0.0f ? 4u : 0u);
}

// MULTI: the scene has a two-lobe material (plastic)
template <bool GUIDED, bool MULTI, bool SPL>
__global__ __launch_bounds__(128) void k_shade(RenderParamsDev p, ShadeSceneDev sc, PathBuffers b, uint32_t bounce) {
    const uint64_t i = (uint64_t)blockIdx.x * 128 + threadIdx.x;
    const uint32_t live = i < p.n_paths ? shade_path<GUIDED, MULTI, SPL>(p, sc, b, bounce, i) : 0u;
    store_ray_masks<3>(b, i, live);
}

// One path's next-event estimate of this vertex, from its shadow and MIS-branch hits, added to
// its radiance.
template <bool SPL>
__device__ __forceinline__ void nee_resolve_path(const RenderParamsDev& p, const ShadeSceneDev& sc,
                                                 const PathBuffers& b, uint32_t bounce, uint64_t i) {
    const float4 a = b.nee_a[i];
    const uint32_t nee = __float_as_uint(a.w);
    if (!(nee & kHasNee)) return;
    // integrator.rs:309-434: beta (term_light + f_bsdf Le w / pdf_bsdf) / light_pdf, the scaling
    // by beta / light_pdf applied when the terms were stored
    C3 add = mkc(0.0f);
    if ((nee & kHasLight) && b.shit[i].x == 0xFFFFFFFFu) add = add + C3{a.x, a.y, a.z};
    if (nee & kHasBsdf) {
        const uint32_t hp = b.mhit[i].x;
        if (hp != 0xFFFFFFFFu) {
            const int32_t li = (int32_t)sc.prim_info[hp].x;
            const float4 kb = b.nee_b[i];
            const bool same = p.mode == 0 ? (li >= 0) : (li == (int32_t)__float_as_uint(kb.w));
            if (same) {
                const C3 Le = emission_at<SPL>(sc, li, hp, b.mray[2 * i], b.mray[2 * i + 1]);
                if (!black(Le)) add = add + C3{kb.x, kb.y, kb.z} * Le;
            }
        }
    }
    const float4 l4 = b.L[i];
    const C3 L = C3{l4.x, l4.y, l4.z} + add;
    b.L[i] = make_float4(L.r, L.g, L.b, l4.w);
    uint64_t rec;
    if (rec_slot(p, i, bounce, &rec) && b.rec_tgt[rec].w != 0.0f) b.rec_aux[2 * rec + 1] = make_float4(L.r, L.g, L.b, 0.0f);
}

template <bool SPL>
__global__ __launch_bounds__(256) void k_nee_resolve(RenderParamsDev p, ShadeSceneDev sc, PathBuffers b,
                                                     uint32_t bounce) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < p.n_paths) nee_resolve_path<SPL>(p, sc, b, bounce, i);
}

// The persistent path kernel of a plain pass. Each lane of a wavefront holds one path at a time,
// at its own bounce; a lane whose path has ended takes the next slot of the run of 64 consecutive
// slots its wavefront last claimed from an atomic counter (zeroed before the launch), claiming the
// next run once that one is handed out, so lanes keep working while the wavefront's other paths
// are short or long. One round of the loop traces the lanes' extension (or camera) rays, shades
// their vertices, traces the shadow and MIS-branch rays the vertices wrote, and resolves the
// next-event estimates. These are the per-path functions the wavefront pass launches one kernel at
// a time (init_path, shade_path, nee_resolve_path), called for each path in the same order on the
// same path-state arrays, and each ray is traced by the same pair traversal, so every path's
// radiance, film point and ray count equal the wavefront render's (a path that has ended is not
// shaded again: those shadings change none of them). The three ray sets share the block's LDS stack
// ring and one spill row per thread, being traced one after another. Path slots run item-major,
// then sample, then pixel, so a run is one scene item's pixels at one sample.
// The kernel holds the shading's registers and the traversal's together: 201-247 VGPRs (2 waves
// per SIMD) as the compiler picks them; it is held to 4 waves (128 VGPRs, the rest spilled), which
// measured fastest on the 64-scene batch when each wavefront ran 64 whole paths in lockstep
// (2 / 3 / 4 / 5 / 6 waves 0.72 / 0.57 / 0.52 / 0.53 / 0.56 s).
template <bool MULTI, bool SPL, bool SPH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_path_persistent(RenderParamsDev p, ShadeSceneDev sc, PathBuffers b,
                                                         TraverseArgs ext, TraverseArgs sha, TraverseArgs mis,
                                                         unsigned long long* queue) {
    constexpr int kBlock = 256, kStack = 8;
    __shared__ uint2 lds[kStack * kBlock];
    const uint64_t gtid = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    uint2* __restrict__ ring = lds + threadIdx.x;
    uint2* __restrict__ spill = reinterpret_cast<uint2*>(ext.spill) + gtid * ext.spill_depth;
    const uint32_t lane = threadIdx.x & 63;
    PairCounts cnt;
    uint64_t i = 0;          // this lane's path slot, while `have`
    bool have = false;
    bool ext_live = false;   // its extension (or camera) ray is to be traced
    uint32_t bounce = 0;
    uint64_t run = 0, run_end = 0;  // slots claimed and not yet handed out (uniform)
    bool drained = false;           // the counter has passed the last slot (uniform)
    for (;;) {
        // hand out slots to the lanes without a path, in lane order
        for (;;) {
            const uint64_t idle = __builtin_amdgcn_ballot_w64(!have);
            if (idle == 0 || drained) break;
            if (run == run_end) {
                unsigned long long b0 = 0;
                if (lane == 0) b0 = atomicAdd(queue, 64ull);
                run = (uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)b0) |
                      ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(b0 >> 32)) << 32);
                if (run >= p.n_paths) {
                    drained = true;
                    break;
                }
                run_end = run + 64 < p.n_paths ? run + 64 : p.n_paths;
            }
            const uint64_t avail = run_end - run;
            const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
            if (!have && rank < avail) {
                i = run + rank;
                have = true;
                bounce = 0;
                ext_live = init_path(p, b, i) > 0.0f;
            }
            const uint64_t given = (uint64_t)__builtin_popcountll(idle);
            run += given < avail ? given : avail;
        }
        if (__builtin_amdgcn_ballot_w64(have) == 0) break;  // no slots left and every path done
        trace_lane<kBlock, kStack, false, SPH>(ext, i, have && ext_live, ring, spill, cnt);
        const uint32_t rays = have ? shade_path<false, MULTI, SPL>(p, sc, b, bounce, i) : 0u;
        trace_lane<kBlock, kStack, true, SPH>(sha, i, (rays & 2u) != 0, ring, spill, cnt);
        trace_lane<kBlock, kStack, false, SPH>(mis, i, (rays & 4u) != 0, ring, spill, cnt);
        if (have && sc.n_lights > 0 && bounce < p.max_depth) nee_resolve_path<SPL>(p, sc, b, bounce, i);
        // the path goes on with its extension ray (none is written at max_depth), else the lane is free
        ext_live = have && (rays & 1u) != 0 && bounce < p.max_depth;
        if (ext_live) ++bounce;
        else have = false;
    }
}

// Network input at each live path's hit: position, outgoing direction and the
// geometric normal turned toward it.
__global__ __launch_bounds__(256) void k_guide_inputs(RenderParamsDev p, ShadeSceneDev sc, PathBuffers b) {
    __shared__ uint32_t wt[4];
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const bool live = i < p.n_paths && guide_live(b, i);
    uint32_t total;
    const uint32_t rank = block_exclusive(live ? 1u : 0u, wt, &total);
    if (i >= p.n_paths) return;
    if (!live) {
        b.gslot[i] = 0xFFFFFFFFu;
        return;
    }
    const uint4 h = b.hit[i];
    const uint32_t slot = b.gblk[blockIdx.x] + rank;  // compacted, in path order
    b.gslot[i] = slot;
    float4 gp, gw, gn;
    {
        const float4 r0 = b.ray[2 * i], r1 = b.ray[2 * i + 1];
        const float t = __uint_as_float(h.y);
        const V d = mkv(r1.x, r1.y, r1.z);
        const V hp = mkv(r0.x, r0.y, r0.z) + d * t;
        V n;
        if (h.x < sc.n_tris) {
            V p0, p1, p2;
            tri_verts(sc.tri_orig + 9 * (uint64_t)h.x, &p0, &p1, &p2);
            n = cross(p1 - p0, p2 - p0);
        } else {
            const float* m = sc.spheres[h.x - sc.n_tris].o2w;
            n = hp - mkv(m[3], m[7], m[11]);
        }
        const float l2 = len2(n);
        n = l2 > 0.0f ? n * (1.0f / __builtin_sqrtf(l2)) : mkv(0.0f, 0.0f, 1.0f);
        if (dot(n, d) > 0.0f) n = -n;
        gp = make_float4(hp.x, hp.y, hp.z, 0.0f);
        gw = make_float4(-d.x, -d.y, -d.z, 0.0f);
        gn = make_float4(n.x, n.y, n.z, 0.0f);
    }
    b.gin[3 * (uint64_t)slot] = gp;
    b.gin[3 * (uint64_t)slot + 1] = gw;
    b.gin[3 * (uint64_t)slot + 2] = gn;
}

// After a pass: the training weight of each recorded direction,
// luminance((L_final - L_after_vertex) / beta_before_scattering)
// = luminance(f cos L_i / pdf) of the sampled direction (0 for empty slots).
__global__ __launch_bounds__(256) void k_guide_targets(RenderParamsDev p, PathBuffers b) {
    const uint64_t r = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= (uint64_t)p.rec_paths * p.max_depth) return;
    float4 t = b.rec_tgt[r];
    float w = 0.0f;
    if (t.w != 0.0f) {
        const uint64_t i = (r / p.max_depth) * p.rec_stride;
        const float4 lf = b.L[i], be = b.rec_aux[2 * r], ls = b.rec_aux[2 * r + 1];
        const float cr = be.x > 0.0f ? fmaxf(lf.x - ls.x, 0.0f) / be.x : 0.0f;
        const float cg = be.y > 0.0f ? fmaxf(lf.y - ls.y, 0.0f) / be.y : 0.0f;
        const float cb = be.z > 0.0f ? fmaxf(lf.z - ls.z, 0.0f) / be.z : 0.0f;
        w = luminance(C3{cr, cg, cb});
        if (!__builtin_isfinite(w)) w = 0.0f;
    } else {
        b.rec_in[3 * r] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        b.rec_in[3 * r + 1] = make_float4(0.0f, 0.0f, 1.0f, 0.0f);
        b.rec_in[3 * r + 2] = make_float4(0.0f, 0.0f, 1.0f, 0.0f);
    }
    t.w = w;
    b.rec_tgt[r] = t;
}

// Sums the per-path ray counts of a pass into counters[0] (all queries) and
// counters[1] (any-hit queries): grid-stride partial sums, two atomics per block.
__global__ __launch_bounds__(256) void k_count_rays(uint64_t n, const float4* __restrict__ beta,
                                                    unsigned long long* __restrict__ counters) {
    __shared__ unsigned long long part[2][4];
    unsigned long long all = 0, shadow = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
        const uint32_t c = __float_as_uint(beta[i].w);
        all += c & 0xFFFFu;
        shadow += c >> 16;
    }
    for (int off = 32; off > 0; off >>= 1) {
        all += __shfl_down(all, off);
        shadow += __shfl_down(shadow, off);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) { part[0][wave] = all; part[1][wave] = shadow; }
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd(&counters[0], part[0][0] + part[0][1] + part[0][2] + part[0][3]);
        atomicAdd(&counters[1], part[1][0] + part[1][1] + part[1][2] + part[1][3]);
    }
}

// Live paths of this bounce (alive, with a hit), compacted in path order for the
// network: per-block counts, one-block exclusive scan, then k_guide_inputs
// writes each live path at its block's offset plus its rank in the block.

// Rays that can report a hit: t_max > 0 (every primitive test needs 0 < t < t_max,
// so a ray with t_max <= 0 or NaN, such as a killed path's, can only miss).
__device__ __forceinline__ bool ray_can_hit(const float4* __restrict__ rays, uint64_t i) { return rays[2 * i].w > 0.0f; }

__global__ __launch_bounds__(256) void k_ray_count(const float4* __restrict__ rays, uint64_t n,
                                                   uint32_t* __restrict__ blk, uint64_t* __restrict__ mask) {
    __shared__ uint32_t wt[4];
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const bool live = i < n && ray_can_hit(rays, i);
    const uint64_t m = __ballot(live);
    if ((threadIdx.x & 63) == 0) mask[i >> 6] = m;  // one word per wavefront: bit l for ray i + l
    uint32_t total;
    block_exclusive(live ? 1u : 0u, wt, &total);
    if (threadIdx.x == 0) blk[blockIdx.x] = total;
}

// Per group of 4096 rays (64 mask words), how many can hit, from the wavefront masks the kernel that
// wrote the rays stored (store_ray_masks): one wavefront per group, lane l reading word 64 g + l
// (words from ceil(n / 64) on count as empty).
__global__ __launch_bounds__(256) void k_mask_count(const uint64_t* __restrict__ mask, uint64_t n,
                                                    uint32_t* __restrict__ gcnt) {
    const uint64_t g = ((uint64_t)blockIdx.x * 256 + threadIdx.x) >> 6;
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t nw = (n + 63) / 64, ng = (nw + 63) / 64;
    if (g >= ng) return;  // (whole wavefronts)
    const uint64_t w = 64 * g + lane;
    uint32_t c = w < nw ? (uint32_t)__popcll(mask[w]) : 0u;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
    if (lane == 0) gcnt[g] = c;
}

// After k_live_scan over the group counts: each ray that can hit at its group's offset, plus the
// live rays of the group's earlier words (a wavefront-wide scan over the group's 64 words), plus
// its rank in its word; ascending ray order. No misses are written (see kQuiet).
__global__ __launch_bounds__(256) void k_mask_compact(uint64_t n, const uint32_t* __restrict__ gcnt,
                                                      const uint64_t* __restrict__ mask, uint32_t* __restrict__ idx) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t word = ((uint64_t)blockIdx.x * 256 + threadIdx.x) >> 6;  // this wavefront's word
    const uint64_t nw = (n + 63) / 64;
    if (word >= nw) return;  // (whole wavefronts)
    const uint64_t g = word >> 6;
    const uint64_t w = 64 * g + lane;
    const uint32_t c = w < nw ? (uint32_t)__popcll(mask[w]) : 0u;
    uint32_t x = c;  // inclusive scan over the group's words
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if (lane >= (uint32_t)d) x += y;
    }
    const uint32_t k = (uint32_t)(word & 63);
    const uint32_t before = __shfl(x - c, k, 64);  // live rays in the group's words before this one
    const uint64_t m = mask[word];
    if ((m >> lane) & 1ull)
        idx[gcnt[g] + before + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = (uint32_t)(64 * word + lane);
}

// After k_live_scan over blk: each ray that can hit at its block's offset plus its rank in the
// block (ascending ray order); every other ray gets the miss the traversal would report. The
// ranks come from k_ray_count's wavefront masks (bits past n are zero), so the rays are not read
// again.
__global__ __launch_bounds__(256) void k_ray_compact(uint64_t n, const uint32_t* __restrict__ blk,
                                                     const uint64_t* __restrict__ mask, uint32_t* __restrict__ idx,
                                                     uint4* __restrict__ hits) {
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if ((i & ~63ull) >= n) return;  // a wavefront wholly past the rays (uniform)
    const uint64_t* bm = mask + (uint64_t)blockIdx.x * 4;
    uint32_t off = blk[blockIdx.x];
    for (uint32_t w = 0; w < wave; ++w) off += (uint32_t)__popcll(bm[w]);
    const uint64_t m = bm[wave];
    if ((m >> lane) & 1ull) idx[off + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = (uint32_t)i;
    else if (i < n) hits[i] = make_uint4(0xFFFFFFFFu, 0u, 0u, 0u);
}

__global__ __launch_bounds__(256) void k_live_count(uint64_t n, PathBuffers b) {
    __shared__ uint32_t wt[4];
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    uint32_t total;
    block_exclusive(i < n && guide_live(b, i) ? 1u : 0u, wt, &total);
    if (threadIdx.x == 0) b.gblk[blockIdx.x] = total;
}

// Exclusive prefix of nblk counts in place, and their sum, in one block (dev::scan_counts).
__global__ __launch_bounds__(1024) void k_live_scan(uint32_t nblk, uint32_t* __restrict__ blk, uint32_t* __restrict__ total) {
    __shared__ uint32_t wsum[16];
    const uint32_t all = dev::scan_counts(nblk, blk, wsum);
    if (threadIdx.x == 0) *total = all;
}

// film.rs:232-284 with a radius-0.5 box filter, gathered per pixel: a sample at
// film point f covers pixels [ceil(f - 1), floor(f) + 1) clipped to its tile's
// pixel bounds [tile_min - 1, tile_max) (film.rs:121-136) and the film.
// Adds every sample of the pass whose filter footprint covers pixel (px, py) of a
// W x H film, in a fixed order: samples s, then rows dy, then columns dx. The
// pass's paths of this film start at path_base, npix = window pixels per sample.
__device__ __forceinline__ void gather_pixel(const PathBuffers& b, int px, int py, int x0, int y0, int x1, int y1,
                                             int W, int H, uint64_t path_base, uint32_t n_samples, float4* acc,
                                             double4* acc64) {
    const uint32_t ww = (uint32_t)(x1 - x0);
    const uint64_t npix = (uint64_t)ww * (uint32_t)(y1 - y0);
    for (uint32_t s = 0; s < n_samples; ++s) {
        for (int dy = -1; dy <= 1; ++dy) {
            const int qy = py + dy;
            if (qy < y0 || qy >= y1) continue;
            for (int dx = -1; dx <= 1; ++dx) {
                const int qx = px + dx;
                if (qx < x0 || qx >= x1) continue;
                const uint64_t path = path_base + (uint64_t)s * npix + (uint64_t)(qy - y0) * ww + (qx - x0);
                const float2 f = b.film[path];
                const float fpx = f.x - 0.5f, fpy = f.y - 0.5f;
                const int tx0 = (qx / 16) * 16, ty0 = (qy / 16) * 16;
                const int lx = max((int)ceilf(fpx - 0.5f), max(tx0 - 1, 0));
                const int ly = max((int)ceilf(fpy - 0.5f), max(ty0 - 1, 0));
                const int hx = min((int)floorf(fpx + 0.5f) + 1, min(tx0 + 16, W));
                const int hy = min((int)floorf(fpy + 0.5f) + 1, min(ty0 + 16, H));
                if (px < lx || px >= hx || py < ly || py >= hy) continue;
                const float4 l4 = b.L[path];
                C3 L = C3{l4.x, l4.y, l4.z};
                // integrator.rs:92-117
                const float y = luminance(L);
                if (L.r != L.r || L.g != L.g || L.b != L.b) L = mkc(0.0f);
                else if (y < -1e-5f) L = mkc(0.0f);
                else if (__builtin_isinf(y)) L = mkc(0.0f);
                acc->x += L.r; acc->y += L.g; acc->z += L.b; acc->w += 1.0f;
                if (acc64) { acc64->x += L.r; acc64->y += L.g; acc64->z += L.b; acc64->w += 1.0; }
            }
        }
    }
}

// film.rs:232-284 with a radius-0.5 box filter, gathered per pixel: a sample at
// film point f covers pixels [ceil(f - 1), floor(f) + 1) clipped to its tile's
// pixel bounds [tile_min - 1, tile_max) (film.rs:121-136) and the film.
__global__ __launch_bounds__(256) void k_film(RenderParamsDev p, PathBuffers b, uint32_t W, uint32_t H, float4* film,
                                              double4* film64) {
    const uint32_t ww = p.x1 - p.x0, wh = p.y1 - p.y0;
    const uint64_t npix = (uint64_t)ww * wh;
    const uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= npix) return;
    const int px = (int)(p.x0 + t % ww), py = (int)(p.y0 + t / ww);
    float4 acc = film[(uint64_t)py * W + px];
    double4 acc64 = film64 ? film64[(uint64_t)py * W + px] : double4{0, 0, 0, 0};
    gather_pixel(b, px, py, (int)p.x0, (int)p.y0, (int)p.x1, (int)p.y1, (int)W, (int)H, 0, p.n_samples, &acc,
                 film64 ? &acc64 : nullptr);
    film[(uint64_t)py * W + px] = acc;
    if (film64) film64[(uint64_t)py * W + px] = acc64;
}

// The films of a scene batch, concatenated (item k's pixels from pix_base).
__global__ __launch_bounds__(256) void k_batch_film(RenderParamsDev p, PathBuffers b, uint64_t total_pix, float4* film) {
    const uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= total_pix) return;
    const BatchItemDev& it = p.batch[item_of_pixel(p, t)];
    const int W = (int)it.cam.width, H = (int)it.cam.height;
    const uint64_t local = t - it.pix_base;
    const int px = (int)(local % (uint64_t)W), py = (int)(local / (uint64_t)W);
    float4 acc = film[t];
    gather_pixel(b, px, py, 0, 0, W, H, W, H, it.path_base, p.n_samples, &acc, nullptr);
    film[t] = acc;
}

__global__ void k_pass_advance(uint32_t* pass_first, uint32_t spp) {
    if (threadIdx.x == 0) *pass_first += spp;
}
__global__ void k_zero_u64(unsigned long long* q) {
    if (threadIdx.x == 0) *q = 0ull;
}

}  // namespace

hipError_t launch_init_paths(const RenderParamsDev& p, const PathBuffers& b, hipStream_t s) {
    if (p.n_paths == 0) return hipSuccess;
    hipLaunchKernelGGL(k_init_paths, dim3((uint32_t)((p.n_paths + 255) / 256)), dim3(256), 0, s, p, b);
    return hipGetLastError();
}
hipError_t launch_shade(const RenderParamsDev& p, const ShadeSceneDev& sc, const PathBuffers& b, uint32_t bounce,
                        hipStream_t s) {
    if (p.n_paths == 0) return hipSuccess;
    const dim3 grid((uint32_t)((p.n_paths + 127) / 128));
#define BG_SHADE(G, M)                                                                                  \
    do {                                                                                                \
        if (sc.sphere_lights) hipLaunchKernelGGL((k_shade<G, M, true>), grid, dim3(128), 0, s, p, sc, b, bounce); \
        else hipLaunchKernelGGL((k_shade<G, M, false>), grid, dim3(128), 0, s, p, sc, b, bounce);                 \
    } while (0)
    if (p.guided) {
        if (sc.multi_lobe) BG_SHADE(true, true);
        else BG_SHADE(true, false);
    } else {
        if (sc.multi_lobe) BG_SHADE(false, true);
        else BG_SHADE(false, false);
    }
#undef BG_SHADE
    return hipGetLastError();
}
hipError_t launch_nee_resolve(const RenderParamsDev& p, const ShadeSceneDev& sc, const PathBuffers& b,
                              uint32_t bounce, hipStream_t s) {
    if (p.n_paths == 0) return hipSuccess;
    const dim3 grid((uint32_t)((p.n_paths + 255) / 256));
    if (sc.sphere_lights) hipLaunchKernelGGL(k_nee_resolve<true>, grid, dim3(256), 0, s, p, sc, b, bounce);
    else hipLaunchKernelGGL(k_nee_resolve<false>, grid, dim3(256), 0, s, p, sc, b, bounce);
    return hipGetLastError();
}
hipError_t launch_path_persistent(const RenderParamsDev& p, const ShadeSceneDev& sc, const PathBuffers& b,
                                  const TraverseArgs& ext, const TraverseArgs& sha, const TraverseArgs& mis,
                                  unsigned long long* queue, int grid, hipStream_t s) {
    if (p.n_paths == 0) return hipSuccess;
    hipError_t e = launch_zero_u64(queue, s);
    if (e != hipSuccess) return e;
    const bool sph = ext.spheres != nullptr;
#define BG_PERSIST(M, L, S) \
    hipLaunchKernelGGL((k_path_persistent<M, L, S>), dim3(grid), dim3(256), 0, s, p, sc, b, ext, sha, mis, queue)
    if (sc.multi_lobe) {
        if (sc.sphere_lights) { if (sph) BG_PERSIST(true, true, true); else BG_PERSIST(true, true, false); }
        else { if (sph) BG_PERSIST(true, false, true); else BG_PERSIST(true, false, false); }
    } else {
        if (sc.sphere_lights) { if (sph) BG_PERSIST(false, true, true); else BG_PERSIST(false, true, false); }
        else { if (sph) BG_PERSIST(false, false, true); else BG_PERSIST(false, false, false); }
    }
#undef BG_PERSIST
    return hipGetLastError();
}
int path_persistent_blocks_per_cu(const ShadeSceneDev& sc, bool spheres) {
    int n = 0;
    hipError_t e;
#define BG_OCC(M, L, S) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_path_persistent<M, L, S>, 256, 0)
    if (sc.multi_lobe) {
        if (sc.sphere_lights) { if (spheres) BG_OCC(true, true, true); else BG_OCC(true, true, false); }
        else { if (spheres) BG_OCC(true, false, true); else BG_OCC(true, false, false); }
    } else {
        if (sc.sphere_lights) { if (spheres) BG_OCC(false, true, true); else BG_OCC(false, true, false); }
        else { if (spheres) BG_OCC(false, false, true); else BG_OCC(false, false, false); }
    }
#undef BG_OCC
    return e == hipSuccess && n > 0 ? n : 1;
}
hipError_t launch_count_rays(const RenderParamsDev& p, const PathBuffers& b, unsigned long long* counters,
                             hipStream_t s) {
    if (p.n_paths == 0) return hipSuccess;
    const uint64_t blocks = std::min<uint64_t>((p.n_paths + 255) / 256, 1024);
    hipLaunchKernelGGL(k_count_rays, dim3((uint32_t)blocks), dim3(256), 0, s, p.n_paths, b.beta, counters);
    return hipGetLastError();
}
hipError_t launch_ray_compact(const float4* rays, uint64_t n, uint32_t* blk, uint64_t* mask, uint32_t* idx,
                              uint32_t* count, uint4* hits, hipStream_t s, const uint64_t* produced) {
    if (n == 0) return hipSuccess;
    const uint32_t nblk = (uint32_t)((n + 255) / 256);
    if (produced) {
        // counts per 4096 rays, so the one-block scan covers 16 times fewer counts
        const uint32_t ngrp = (uint32_t)((n + 4095) / 4096);
        hipLaunchKernelGGL(k_mask_count, dim3((ngrp + 3) / 4), dim3(256), 0, s, produced, n, blk);
        hipLaunchKernelGGL(k_live_scan, dim3(1), dim3(1024), 0, s, ngrp, blk, count);
        hipLaunchKernelGGL(k_mask_compact, dim3(nblk), dim3(256), 0, s, n, blk, produced, idx);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_ray_count, dim3(nblk), dim3(256), 0, s, rays, n, blk, mask);
    hipLaunchKernelGGL(k_live_scan, dim3(1), dim3(1024), 0, s, nblk, blk, count);
    hipLaunchKernelGGL(k_ray_compact, dim3(nblk), dim3(256), 0, s, n, blk, mask, idx, hits);
    return hipGetLastError();
}
hipError_t launch_guide_inputs(const RenderParamsDev& p, const ShadeSceneDev& sc, const PathBuffers& b, hipStream_t s) {
    if (p.n_paths == 0) return hipSuccess;
    const uint32_t nblk = (uint32_t)((p.n_paths + 255) / 256);
    hipLaunchKernelGGL(k_live_count, dim3(nblk), dim3(256), 0, s, p.n_paths, b);
    hipLaunchKernelGGL(k_live_scan, dim3(1), dim3(1024), 0, s, nblk, b.gblk, b.gcount);
    hipLaunchKernelGGL(k_guide_inputs, dim3((uint32_t)((p.n_paths + 255) / 256)), dim3(256), 0, s, p, sc, b);
    return hipGetLastError();
}
hipError_t launch_guide_targets(const RenderParamsDev& p, const PathBuffers& b, hipStream_t s) {
    const uint64_t n = (uint64_t)p.rec_paths * p.max_depth;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_guide_targets, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, p, b);
    return hipGetLastError();
}
hipError_t launch_batch_film(const RenderParamsDev& p, const PathBuffers& b, uint64_t total_pix, float4* film,
                             hipStream_t s) {
    if (total_pix == 0) return hipSuccess;
    hipLaunchKernelGGL(k_batch_film, dim3((uint32_t)((total_pix + 255) / 256)), dim3(256), 0, s, p, b, total_pix, film);
    return hipGetLastError();
}
hipError_t launch_pass_advance(uint32_t* pass_first, uint32_t spp, hipStream_t s) {
    hipLaunchKernelGGL(k_pass_advance, dim3(1), dim3(64), 0, s, pass_first, spp);
    return hipGetLastError();
}
hipError_t launch_zero_u64(unsigned long long* q, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, q);
    return hipGetLastError();
}
hipError_t launch_film(const RenderParamsDev& p, const PathBuffers& b, uint32_t W, uint32_t H, float4* film,
                       double4* film64, hipStream_t s) {
    const uint64_t npix = (uint64_t)(p.x1 - p.x0) * (p.y1 - p.y0);
    if (npix == 0) return hipSuccess;
    hipLaunchKernelGGL(k_film, dim3((uint32_t)((npix + 255) / 256)), dim3(256), 0, s, p, b, W, H, film, film64);
    return hipGetLastError();
}

}  // namespace bg
