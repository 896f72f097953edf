/*
 * bonsai_scenes.h — seeded workload generators for the benchmark and test
 * configurations (displaced icospheres, ray sets). No counterpart in the
 * reference: they produce the synthetic scenes the configurations describe.
 * Only +, -, *, / and sqrt are used, so the output is bit-identical on any
 * x86-64 host.
 */
#ifndef BONSAI_SCENES_H
#define BONSAI_SCENES_H

#include <stdint.h>

#include "bonsai_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Number of triangles of an icosphere subdivided `level` times: 20 * 4^level. */
uint64_t bgx_icosphere_tri_count(int level);

/* Non-indexed triangle soup (9 floats per triangle) of an icosphere of the
 * given centre and radius, each vertex displaced radially by
 * radius * amplitude * noise(frequency * p), with seeded 3-D value noise in
 * [-1, 1]. Winding is outward for amplitude < 1. */
int bgx_icosphere(int level, const float center[3], float radius, float amplitude, float frequency, uint64_t seed,
                  float* out_tris);

/* Incoherent ray set: origins uniform in [bmin, bmax], directions uniform on
 * the unit sphere (Marsaglia), t_max = +inf, time = 0. */
int bgx_rays_incoherent(uint64_t n, const float bmin[3], const float bmax[3], uint64_t seed, bg_ray* out);

/* Counter-based uniform in [0, 1) with 24 random bits for (seed, index). */
float bgx_uniform(uint64_t seed, uint64_t index);

#ifdef __cplusplus
}
#endif

#endif /* BONSAI_SCENES_H */
