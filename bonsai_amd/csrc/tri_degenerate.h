// The triangle degeneracy test of the reference's intersection routines
// (triangle.rs:241-268 / 493-517), shared by device code and the scene
// upload. It depends on the vertices only, so the upload evaluates it once
// per triangle; the sources are built with -ffp-contract=off on both sides.
#pragma once

#include <hip/hip_runtime.h>

namespace bg {

// f64 cross product rounded to f32, then its f32 squared length (vec3.rs:108-122, 103-105).
__host__ __device__ inline float cross64_len2(float ax, float ay, float az, float bx, float by, float bz) {
    const double x = ax, y = ay, z = az, vx = bx, vy = by, vz = bz;
    const float cx = (float)((y * vz) - (z * vy));
    const float cy = (float)((z * vx) - (x * vz));
    const float cz = (float)((x * vy) - (y * vx));
    return cx * cx + cy * cy + cz * cz;
}

// With the default uvs (0,0), (1,0), (1,1): dpdu = (-1*dp02 - (-1)*dp12) * 1,
// dpdv = (-0*dp02 + (-1)*dp12) * 1. Returns true when the triangle is
// degenerate (every intersection with it is rejected).
__host__ __device__ inline bool tri_degenerate(float p0x, float p0y, float p0z, float p1x, float p1y, float p1z,
                                               float p2x, float p2y, float p2z) {
    const float d02x = p0x - p2x, d02y = p0y - p2y, d02z = p0z - p2z;
    const float d12x = p1x - p2x, d12y = p1y - p2y, d12z = p1z - p2z;
    const float m1 = -1.0f, nz = -0.0f;
    const float ux = (m1 * d02x - m1 * d12x) * 1.0f, uy = (m1 * d02y - m1 * d12y) * 1.0f,
                uz = (m1 * d02z - m1 * d12z) * 1.0f;
    const float vx = (nz * d02x + m1 * d12x) * 1.0f, vy = (nz * d02y + m1 * d12y) * 1.0f,
                vz = (nz * d02z + m1 * d12z) * 1.0f;
    if (cross64_len2(ux, uy, uz, vx, vy, vz) != 0.0f) return false;
    return cross64_len2(p2x - p0x, p2y - p0y, p2z - p0z, p1x - p0x, p1y - p0y, p1z - p0z) == 0.0f;
}

}  // namespace bg
