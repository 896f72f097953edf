// Film::write_image (src/base/film.rs:156-202) as one kernel over the film's
// per-pixel sums. One thread per pixel reads a 16-byte (r, g, b, weight) sum
// and writes 12 bytes of RGB; about thirty multiplies and adds per pixel, so
// the kernel is bound by HBM (28 bytes per pixel).
//
// Per pixel, in the reference's order of operations:
//   merge_film_tile (film.rs:138-154): the tile's RGB sum goes to XYZ
//     (rgb_to_xyz, spectrum.rs:40-44) and is added to the pixel's XYZ (zero);
//   write_image: XYZ back to RGB (xyz_to_rgb, spectrum.rs:34-38); if the
//     weight sum is non-zero each channel is multiplied by 1 / weight and
//     clamped with f32::max(x, 0) (a NaN becomes 0); the splat (always zero
//     for these integrators) goes through xyz_to_rgb and is added scaled by
//     splat_scale; every channel is multiplied by the film scale.
// The sources are compiled with -ffp-contract=off, so each product and sum
// rounds once as the reference's f32 arithmetic does.
#include "bg_internal.h"

namespace bg {
namespace {

__device__ __forceinline__ void xyz_to_rgb(const float* xyz, float* rgb) {
    rgb[0] = 3.240479f * xyz[0] - 1.537150f * xyz[1] - 0.498535f * xyz[2];
    rgb[1] = -0.969256f * xyz[0] + 1.875991f * xyz[1] + 0.041556f * xyz[2];
    rgb[2] = 0.055648f * xyz[0] - 0.204043f * xyz[1] + 1.057311f * xyz[2];
}

__global__ __launch_bounds__(256) void k_film_resolve(const float4* __restrict__ sums, uint64_t n, float splat_scale,
                                                      float scale, float* __restrict__ rgb_out) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float4 s = sums[i];
    float xyz[3];
    xyz[0] = 0.412453f * s.x + 0.357580f * s.y + 0.180423f * s.z;
    xyz[1] = 0.212671f * s.x + 0.715160f * s.y + 0.072169f * s.z;
    xyz[2] = 0.019334f * s.x + 0.119193f * s.y + 0.950227f * s.z;
    // Pixel::default() holds zero XYZ; the merge adds the tile's XYZ to it
    const float merged[3] = {0.0f + xyz[0], 0.0f + xyz[1], 0.0f + xyz[2]};
    float rgb[3];
    xyz_to_rgb(merged, rgb);
    if (s.w != 0.0f) {
        const float inv = 1.0f / s.w;
#pragma unroll
        for (int k = 0; k < 3; ++k) rgb[k] = fmaxf(rgb[k] * inv, 0.0f);
    }
    const float zero[3] = {0.0f, 0.0f, 0.0f};
    float splat[3];
    xyz_to_rgb(zero, splat);
#pragma unroll
    for (int k = 0; k < 3; ++k) rgb[k] = (rgb[k] + splat_scale * splat[k]) * scale;
    rgb_out[3 * i + 0] = rgb[0];
    rgb_out[3 * i + 1] = rgb[1];
    rgb_out[3 * i + 2] = rgb[2];
}

}  // namespace

hipError_t launch_film_resolve(const float4* sums, uint64_t n, float splat_scale, float scale, float* rgb_out,
                               hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_film_resolve, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, sums, n, splat_scale, scale,
                       rgb_out);
    return hipGetLastError();
}

}  // namespace bg
