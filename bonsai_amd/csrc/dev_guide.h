// The guiding network's output as a directional density: an 8-lobe von
// Mises-Fisher mixture (no counterpart in the reference).
//
// Raw output per sample (40 floats): lobe j at 5j = (mixture logit,
// direction 3-vector m, concentration parameter s). pi = softmax(logits),
// mu = m / sqrt(|m|^2 + 1e-12), kappa = exp(1.5 + 5.5 tanh((s - 1.5) / 5.5)).
// vMF density: kappa / (2 pi (1 - e^{-2 kappa})) * exp(kappa (mu.w - 1)).
//
// The inference kernel turns the raw output into a mixture record of 10
// float4 that the shading kernel reads directly: lobe j as
// (kappa mu, log(pi_j norm_j) - kappa) so that p(w) = sum_j exp(a_j.w + c_j),
// then pi_0..pi_7 in two float4 (pi_0 < 0 marks a non-finite output).
// Sampling: pick a lobe by pi, cos(theta) = 1 + log(u + (1 - u) e^{-2 kappa}) /
// kappa (inverse CDF), phi = 2 pi u', around mu.
#pragma once

#include "dev_shade.h"

namespace bg {
namespace dev {

constexpr int kMixRecord = 10;  // float4 per sample

// Half of a sample's mixture record, computed by the two lanes that hold the
// sample's outputs (lanes l and l + 32 of a wavefront): this lane has the
// outputs of lobes 4h .. 4h + 3 (o[5j + c] for lobe 4h + j, c = logit,
// direction 3-vector, concentration parameter). The lanes exchange the logit
// maximum, the exp sum and finiteness; lane h writes lobes 4h .. 4h + 3 and
// the mixture weights float4 8 + h. All 64 lanes must call this (it shuffles);
// only lanes with `write` store. A sample with a non-finite output gets a
// zero record with pi_0 = -1 (guide_ok then fails).
__device__ __forceinline__ void guide_mixture_half(const float* __restrict__ o, int h, bool write,
                                                   float4* __restrict__ rec) {
    bool finite = true;
    float amax = -3.0e38f;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        amax = fmaxf(amax, o[5 * j]);
#pragma unroll
        for (int k = 0; k < 5; ++k) finite = finite && __builtin_isfinite(o[5 * j + k]);
    }
    amax = fmaxf(amax, __shfl_xor(amax, 32));
    finite = (__shfl_xor((int)finite, 32) != 0) && finite;
    float pi[4], sum = 0.0f;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        pi[j] = __expf(o[5 * j] - amax);
        sum += pi[j];
    }
    sum += __shfl_xor(sum, 32);
    if (!write) return;
    if (!finite) {
#pragma unroll
        for (int j = 0; j < 4; ++j) rec[4 * h + j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        rec[8 + h] = make_float4(h == 0 ? -1.0f : 0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    const float inv = 1.0f / sum;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        pi[j] *= inv;
        const V m = mkv(o[5 * j + 1], o[5 * j + 2], o[5 * j + 3]);
        const V mu = m * (1.0f / __builtin_sqrtf(len2(m) + 1e-12f));
        const float k = expf(1.5f + 5.5f * tanhf((o[5 * j + 4] - 1.5f) / 5.5f));
        const float norm = k / (2.0f * kPi * -expm1f(-2.0f * k));
        rec[4 * h + j] = make_float4(k * mu.x, k * mu.y, k * mu.z, logf(pi[j] * norm) - k);
    }
    rec[8 + h] = make_float4(pi[0], pi[1], pi[2], pi[3]);
}

__device__ __forceinline__ bool guide_ok(const float4* __restrict__ rec) { return rec[8].x >= 0.0f; }

__device__ __forceinline__ float guide_pdf(const float4* __restrict__ rec, V w) {
    float p = 0.0f;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const float4 a = rec[j];
        p += __expf(a.x * w.x + a.y * w.y + a.z * w.z + a.w);
    }
    return p;
}

// Picks a lobe with ul, then samples it with (u0, u1).
__device__ __forceinline__ V guide_sample(const float4* __restrict__ rec, float ul, float u0, float u1) {
    const float4 p0 = rec[8], p1 = rec[9];
    const float pi[8] = {p0.x, p0.y, p0.z, p0.w, p1.x, p1.y, p1.z, p1.w};
    int j = 0;
    float c = pi[0];
    while (j < 7 && ul >= c) c += pi[++j];
    const float4 a = rec[j];
    const float k = __builtin_sqrtf(a.x * a.x + a.y * a.y + a.z * a.z);
    const V m = mkv(a.x, a.y, a.z) * (1.0f / k);
    float ct = 1.0f + logf(u0 + (1.0f - u0) * __expf(-2.0f * k)) / k;
    ct = fminf(fmaxf(ct, -1.0f), 1.0f);
    const float st = __builtin_sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
    const float phi = 2.0f * kPi * u1;
    // orthonormal frame around m (vec3.rs:25-33 coordinate_system)
    V s;
    if (__builtin_fabsf(m.x) > __builtin_fabsf(m.y)) s = mkv(-m.z, 0.0f, m.x) * (1.0f / __builtin_sqrtf(m.x * m.x + m.z * m.z));
    else s = mkv(0.0f, m.z, -m.y) * (1.0f / __builtin_sqrtf(m.y * m.y + m.z * m.z));
    const V t = cross(m, s);
    return unit(s * (st * __cosf(phi)) + t * (st * __sinf(phi)) + m * ct);
}

}  // namespace dev
}  // namespace bg
