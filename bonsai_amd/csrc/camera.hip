// Camera ray generation: PerspectiveCamera::generate_ray
// (cameras/perspective.rs:117-175) for a pinhole camera, one lane per sample.
// The raster_to_camera matrix comes from camera_host.cpp; the static
// camera_to_world transform is applied as Ray::transform (geometries/ray.rs:50-76),
// including the origin error offset and the t_max reduction.
#include <hip/hip_runtime.h>

#include "bg_internal.h"
#include "dev_camera.h"

namespace bg {
namespace {

__global__ __launch_bounds__(256) void k_camera_rays(CameraDev cam, float fx, float fy, float4* __restrict__ rays) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const uint64_t n = (uint64_t)cam.width * cam.height;
    if (i >= n) return;
    const uint32_t x = (uint32_t)(i % cam.width), y = (uint32_t)(i / cam.width);
    dev::CamRay r = dev::generate_ray(cam, (float)x + fx, (float)y + fy, 0.0f);
    rays[2 * i] = make_float4(r.ox, r.oy, r.oz, r.t_max);
    rays[2 * i + 1] = make_float4(r.dx, r.dy, r.dz, r.time);
}

}  // namespace

hipError_t launch_camera_rays(const CameraDev& cam, float fx, float fy, float4* rays, hipStream_t stream) {
    const uint64_t n = (uint64_t)cam.width * cam.height;
    if (n == 0) return hipSuccess;
    const uint64_t grid = (n + 255) / 256;
    hipLaunchKernelGGL(k_camera_rays, dim3((uint32_t)grid), dim3(256), 0, stream, cam, fx, fy, rays);
    return hipGetLastError();
}

}  // namespace bg
