// Device-side evaluation of the guiding network's output: an 8-lobe von
// Mises-Fisher mixture over directions (no counterpart in the reference).
//
// Output layout per sample (40 floats): lobe j at 5j = (mixture logit,
// direction 3-vector m, concentration parameter s). pi = softmax(logits),
// mu = m / sqrt(|m|^2 + 1e-12), kappa = exp(1.5 + 5.5 tanh((s - 1.5) / 5.5)).
// vMF density: kappa / (2 pi (1 - e^{-2 kappa})) * exp(kappa (mu.w - 1)).
// Sampling: cos(theta) = 1 + log(u + (1 - u) e^{-2 kappa}) / kappa (inverse CDF),
// phi = 2 pi u', around mu.
#pragma once

#include "dev_shade.h"

namespace bg {
namespace dev {

struct GuideMix {
    float pi[8];      // mixture weights (cumulative sums used for the lobe pick)
    V mu[8];
    float kappa[8];
    float norm[8];    // kappa / (2 pi (1 - e^{-2 kappa}))
};

// Returns false when the outputs are not finite (the caller then samples the BSDF only).
__device__ __forceinline__ bool guide_mix(const float* __restrict__ o, GuideMix* g) {
    float amax = -3.0e38f;
    bool finite = true;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        amax = fmaxf(amax, o[5 * j]);
#pragma unroll
        for (int k = 0; k < 5; ++k) finite = finite && __builtin_isfinite(o[5 * j + k]);
    }
    if (!finite) return false;
    float sum = 0.0f;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        g->pi[j] = __expf(o[5 * j] - amax);
        sum += g->pi[j];
    }
    const float inv = 1.0f / sum;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        g->pi[j] *= inv;
        const V m = mkv(o[5 * j + 1], o[5 * j + 2], o[5 * j + 3]);
        g->mu[j] = m * (1.0f / __builtin_sqrtf(len2(m) + 1e-12f));
        const float k = expf(1.5f + 5.5f * tanhf((o[5 * j + 4] - 1.5f) / 5.5f));
        g->kappa[j] = k;
        g->norm[j] = k / (2.0f * kPi * -expm1f(-2.0f * k));
    }
    return true;
}

__device__ __forceinline__ float guide_pdf(const GuideMix& g, V w) {
    float p = 0.0f;
#pragma unroll
    for (int j = 0; j < 8; ++j) p += g.pi[j] * g.norm[j] * __expf(g.kappa[j] * (dot(g.mu[j], w) - 1.0f));
    return p;
}

// Picks a lobe with ul, then samples it with (u0, u1).
__device__ __forceinline__ V guide_sample(const GuideMix& g, float ul, float u0, float u1) {
    int j = 0;
    float c = g.pi[0];
    while (j < 7 && ul >= c) c += g.pi[++j];
    const float k = g.kappa[j];
    float ct = 1.0f + logf(u0 + (1.0f - u0) * __expf(-2.0f * k)) / k;
    ct = fminf(fmaxf(ct, -1.0f), 1.0f);
    const float st = __builtin_sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
    const float phi = 2.0f * kPi * u1;
    const V m = g.mu[j];
    // orthonormal frame around m (vec3.rs:25-33 coordinate_system)
    V s;
    if (__builtin_fabsf(m.x) > __builtin_fabsf(m.y)) s = mkv(-m.z, 0.0f, m.x) * (1.0f / __builtin_sqrtf(m.x * m.x + m.z * m.z));
    else s = mkv(0.0f, m.z, -m.y) * (1.0f / __builtin_sqrtf(m.y * m.y + m.z * m.z));
    const V t = cross(m, s);
    return unit(s * (st * __cosf(phi)) + t * (st * __sinf(phi)) + m * ct);
}

}  // namespace dev
}  // namespace bg
