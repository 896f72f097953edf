// Direct-lighting and Whitted integrators for gfx950 (integrators/direct.rs:72-113,
// integrators/whitted.rs:41-93, base/integrator.rs:274-437), on the same
// wavefront pieces as the path integrator.
//
// A pass holds n_paths camera samples. After the camera rays' closest-hit
// launch, k_direct_shade computes each path's surface, emission and BSDF and
// fills its light slots: one slot per light for UniformSampleAll and Whitted,
// one for UniformSampleOne. A slot holds a shadow ray with the light-sampling
// term it carries if unoccluded and, for direct lighting, the BSDF-sampling
// (MIS) ray with its weight. Slots are slot-major (slot j of path i at
// j * n_paths + i), so each traversal launch covers one contiguous range of
// n_paths * n_slots rays. k_direct_resolve then adds each path's slots in the
// reference's light order.
//
// Every primitive has a material, so the reference's bsdf.is_none()
// continuation never runs; area lights give escaping rays no radiance
// (light.rs:32-34); and the supported materials have no specular lobes, so
// specular_reflect / specular_transmit (integrator.rs:143-271) add nothing
// and are not traced.
#include <hip/hip_runtime.h>

#include "bg_internal.h"
#include "dev_light.h"
#include "dev_shade.h"

namespace bg {
namespace {

using namespace dev;

// DirectSlots::a.w flags
constexpr uint32_t kSlotLight = 1u, kSlotBsdf = 2u;

// One estimate_direct (integrator.rs:330-437) for light idx: writes the slot's
// shadow ray and light-branch term, and its MIS ray, weight and pdf.
template <int NL, bool SPL>
__device__ __forceinline__ void direct_slot(const ShadeSceneDev& sc, const DirectSlots& d, uint64_t slot,
                                            const Surf& s, const BsdfN<NL>& bs, uint32_t idx, float us0, float us1,
                                            float ul0, float ul1, bool ref) {
    const uint32_t nonspec = kAll & ~kSpecular;
    const LightDev& light = sc.lights[idx];
    uint32_t flags = 0;
    const LightPoint lpt = sample_light_point<SPL>(sc, light, s, ul0, ul1, ref);
    float scattering_pdf = 0.0f;
    C3 term = mkc(0.0f);
    if (lpt.pdf > 0.0f && !black(lpt.rad)) {
        const C3 f = bs.f(s.wo, lpt.wi, nonspec) * __builtin_fabsf(dot(lpt.wi, bs.ns));
        scattering_pdf = bs.pdf(s.wo, lpt.wi, nonspec);
        if (!black(f)) {
            write_shadow_ray(d.sray, slot, s, lpt);
            const float w = power_heuristic(lpt.pdf, scattering_pdf);
            term = f * lpt.rad * w / lpt.pdf;
            flags |= kSlotLight;
        }
    }
    // BSDF-sampling branch; the reference traces the light-sample direction with its BSDF pdf
    const BxS bsm = bs.sample(s.wo, us0, us1, nonspec);
    const C3 fb = bsm.f * __builtin_fabsf(dot(bsm.wi, bs.ns));
    const float spdf = ref ? scattering_pdf : bsm.pdf;
    const V mdir = ref ? lpt.wi : bsm.wi;
    if (!black(fb) && spdf > 0.0f) {
        const float lpp = light_point_pdf<SPL>(sc, light, s, mdir, ref);
        if (lpp != 0.0f) {
            V o;
            float tm;
            spawn(s, mdir, ref, &o, &tm);
            write_ray(d.mray, slot, o, tm, mdir, s.time);
            d.b[slot] = make_float4(fb.r, fb.g, fb.b, power_heuristic(spdf, lpp));
            d.c[slot] = make_float2(spdf, __uint_as_float(idx));
            flags |= kSlotBsdf;
        }
    }
    d.a[slot] = make_float4(term.r, term.g, term.b, __uint_as_float(flags));
}

// KIND: 0 direct lighting, all lights; 1 direct lighting, one light; 2 Whitted.
template <int KIND, bool MULTI, bool SPL>
__global__ __launch_bounds__(128) void k_direct_shade(RenderParamsDev p, ShadeSceneDev sc, PathBuffers b,
                                                      DirectSlots d) {
    const uint64_t i = (uint64_t)blockIdx.x * 128 + threadIdx.x;
    if (i >= p.n_paths) return;
    const uint32_t n_slots = KIND == 1 ? 1u : sc.n_lights;
    for (uint32_t j = 0; j < n_slots; ++j) {
        const uint64_t slot = (uint64_t)j * p.n_paths + i;
        kill_ray(d.sray, slot);
        if (KIND != 2) kill_ray(d.mray, slot);
        d.a[slot].w = 0.0f;
    }
    const bool ref = p.mode == 0;
    const uint4 h = b.hit[i];
    const float4 r0 = b.ray[2 * i], r1 = b.ray[2 * i + 1];
    Surf s;
    float t;
    if (h.x == 0xFFFFFFFFu || !hit_surface(sc, r0, r1, h.x, &s, &t)) {
        b.L[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);  // no radiance from the area lights
        return;
    }
    const uint32_t prim = h.x;
    const uint64_t key = key_of(b.key[i]);
    uint32_t dim = b.state[i].x;
    C3 L = mkc(0.0f);
    const uint2 info = sc.prim_info[prim];
    const int32_t hit_light = (int32_t)info.x;
    if (hit_light >= 0) L = L + emission(sc.lights[hit_light], s.n, s.wo);  // it.emitted_radiance(&it.direction)
    const BsdfN<MULTI ? 2 : 1> bs = make_bsdf<MULTI ? 2 : 1>(sc.mats[info.y & ~kPrimFlip], s, ref);
    const uint32_t nl = sc.n_lights;
    if (KIND == 0) {  // integrator.rs:274-306, one sample per light (light.rs:52-54)
        for (uint32_t j = 0; j < nl; ++j) {
            const float ul0 = draw(key, dim), ul1 = draw(key, dim);
            const float us0 = draw(key, dim), us1 = draw(key, dim);
            direct_slot<MULTI ? 2 : 1, SPL>(sc, d, (uint64_t)j * p.n_paths + i, s, bs, j, us0, us1, ul0, ul1, ref);
        }
    } else if (KIND == 1) {  // integrator.rs:309-328
        if (nl > 0) {
            const float fn = (float)nl;
            const float u = draw(key, dim);
            const uint32_t idx = (uint32_t)rmin(u * fn, fn - 1.0f);
            const float ul0 = draw(key, dim), ul1 = draw(key, dim);
            const float us0 = draw(key, dim), us1 = draw(key, dim);
            direct_slot<MULTI ? 2 : 1, SPL>(sc, d, i, s, bs, idx, us0, us1, ul0, ul1, ref);
        }
    } else {  // whitted.rs:67-84
        for (uint32_t j = 0; j < nl; ++j) {
            const uint64_t slot = (uint64_t)j * p.n_paths + i;
            const float u0 = draw(key, dim), u1 = draw(key, dim);
            const LightPoint lpt = sample_light_point<SPL>(sc, sc.lights[j], s, u0, u1, ref);
            if (black(lpt.rad) || lpt.pdf == 0.0f) continue;
            const C3 f = bs.f(s.wo, lpt.wi, kAll);
            if (!black(f)) {
                write_shadow_ray(d.sray, slot, s, lpt);
                const C3 term = f * lpt.rad * __builtin_fabsf(dot(lpt.wi, bs.ns)) / lpt.pdf;
                d.a[slot] = make_float4(term.r, term.g, term.b, __uint_as_float(kSlotLight));
            }
        }
    }
    b.L[i] = make_float4(L.r, L.g, L.b, 0.0f);
}

// Adds each path's light slots to its radiance in light order and records the
// rays it issued (beta.w: low 16 bits all queries, high 16 bits any-hit ones).
template <int KIND, bool SPL>
__global__ __launch_bounds__(256) void k_direct_resolve(RenderParamsDev p, ShadeSceneDev sc, PathBuffers b,
                                                        DirectSlots d) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= p.n_paths) return;
    const uint32_t n_slots = KIND == 1 ? 1u : sc.n_lights;
    const float4 l4 = b.L[i];
    C3 L = C3{l4.x, l4.y, l4.z};
    C3 all = mkc(0.0f);
    uint32_t queries = 1, shadow = 0;
    for (uint32_t j = 0; j < n_slots; ++j) {
        const uint64_t slot = (uint64_t)j * p.n_paths + i;
        const float4 a = d.a[slot];
        const uint32_t f = __float_as_uint(a.w);
        if (f == 0) continue;
        C3 ed = mkc(0.0f);
        if (f & kSlotLight) {
            ++queries;
            ++shadow;
            if (d.shit[slot].x == 0xFFFFFFFFu) ed = ed + C3{a.x, a.y, a.z};
        }
        if (f & kSlotBsdf) {
            ++queries;
            const uint32_t hp = d.mhit[slot].x;
            if (hp != 0xFFFFFFFFu) {
                const int32_t li = (int32_t)sc.prim_info[hp].x;
                const float2 c = d.c[slot];
                // integrator.rs:412-419: the reference accepts any area light, pbrt only the sampled one
                const bool same = p.mode == 0 ? (li >= 0) : (li == (int32_t)__float_as_uint(c.y));
                if (same) {
                    const C3 Le = emission_at<SPL>(sc, li, hp, d.mray[2 * slot], d.mray[2 * slot + 1]);
                    if (!black(Le)) {
                        const float4 fw = d.b[slot];
                        ed = ed + C3{fw.x, fw.y, fw.z} * Le * fw.w / c.x;
                    }
                }
            }
        }
        if (KIND == 0) all = all + (mkc(0.0f) + ed) / 1.0f;  // direct_output / sample_count
        else if (KIND == 1) L = L + ed / (1.0f / (float)sc.n_lights);
        else L = L + ed;
    }
    if (KIND == 0) L = L + all;
    b.L[i] = make_float4(L.r, L.g, L.b, 0.0f);
    b.beta[i].w = __uint_as_float(queries | (shadow << 16));
}

}  // namespace

hipError_t launch_direct_shade(const RenderParamsDev& p, const ShadeSceneDev& sc, const PathBuffers& b,
                               const DirectSlots& d, int kind, hipStream_t s) {
    if (p.n_paths == 0) return hipSuccess;
    const dim3 grid((uint32_t)((p.n_paths + 127) / 128)), block(128);
    const bool m = sc.multi_lobe != 0, spl = sc.sphere_lights != 0;
#define BG_DIRECT(K)                                                                                   \
    do {                                                                                               \
        if (m && spl) hipLaunchKernelGGL((k_direct_shade<K, true, true>), grid, block, 0, s, p, sc, b, d);     \
        else if (m) hipLaunchKernelGGL((k_direct_shade<K, true, false>), grid, block, 0, s, p, sc, b, d);      \
        else if (spl) hipLaunchKernelGGL((k_direct_shade<K, false, true>), grid, block, 0, s, p, sc, b, d);    \
        else hipLaunchKernelGGL((k_direct_shade<K, false, false>), grid, block, 0, s, p, sc, b, d);            \
    } while (0)
    switch (kind) {
        case 0: BG_DIRECT(0); break;
        case 1: BG_DIRECT(1); break;
        default: BG_DIRECT(2);
    }
#undef BG_DIRECT
    return hipGetLastError();
}

hipError_t launch_direct_resolve(const RenderParamsDev& p, const ShadeSceneDev& sc, const PathBuffers& b,
                                 const DirectSlots& d, int kind, hipStream_t s) {
    if (p.n_paths == 0) return hipSuccess;
    const dim3 grid((uint32_t)((p.n_paths + 255) / 256)), block(256);
    const bool spl = sc.sphere_lights != 0;
    if (kind == 0 && spl) hipLaunchKernelGGL((k_direct_resolve<0, true>), grid, block, 0, s, p, sc, b, d);
    else if (kind == 0) hipLaunchKernelGGL((k_direct_resolve<0, false>), grid, block, 0, s, p, sc, b, d);
    else if (kind == 1 && spl) hipLaunchKernelGGL((k_direct_resolve<1, true>), grid, block, 0, s, p, sc, b, d);
    else if (kind == 1) hipLaunchKernelGGL((k_direct_resolve<1, false>), grid, block, 0, s, p, sc, b, d);
    else if (spl) hipLaunchKernelGGL((k_direct_resolve<2, true>), grid, block, 0, s, p, sc, b, d);
    else hipLaunchKernelGGL((k_direct_resolve<2, false>), grid, block, 0, s, p, sc, b, d);
    return hipGetLastError();
}

}  // namespace bg
