// Surface interactions of given hits on gfx950: for each (ray, primitive) pair
// the point, its error bounds, the geometric and shading normals, the shading
// dpdu and the uv that the primitive's intersect fills into an Interaction
// (primitives/geometric.rs:27-38 over shapes/triangle.rs:135-388 and
// shapes/sphere.rs), computed as the shading kernels compute them.
#include <hip/hip_runtime.h>

#include "bg_internal.h"
#include "dev_light.h"

namespace bg {
namespace {

using namespace dev;

// out: 5 float4 per ray: (p, t), (p_error, n.x), (n.y, n.z, ns.x, ns.y), (ns.z, dpdu), (u, v, 0, 0);
// all NaN where the hit is a miss.
__global__ __launch_bounds__(256) void k_surface(ShadeSceneDev sc, const float4* __restrict__ rays,
                                                 const uint4* __restrict__ hits, uint64_t n, float4* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint4 h = hits[i];
    float4* o = out + 5 * i;
    Surf s;
    float t;
    if (h.x == 0xFFFFFFFFu || !hit_surface(sc, rays[2 * i], rays[2 * i + 1], h.x, &s, &t)) {
        const float q = __builtin_nanf("");
        for (int k = 0; k < 5; ++k) o[k] = make_float4(q, q, q, q);
        return;
    }
    const float b1 = __uint_as_float(h.z), b2 = __uint_as_float(h.w);
    float u, v;
    if (h.x < sc.n_tris) {
        // triangle.rs:278 with the default uvs (0,0), (1,0), (1,1): b0 * 0 + b1 * 1 + b2 * 1 = b1 + b2, and b2
        u = b1 + b2;
        v = b2;
    } else {  // the sphere's (phi / phi_max, (theta - theta_min) / (theta_max - theta_min)) of the hit
        u = b1;
        v = b2;
    }
    o[0] = make_float4(s.p.x, s.p.y, s.p.z, t);
    o[1] = make_float4(s.perr.x, s.perr.y, s.perr.z, s.n.x);
    o[2] = make_float4(s.n.y, s.n.z, s.ns.x, s.ns.y);
    o[3] = make_float4(s.ns.z, s.dpdu.x, s.dpdu.y, s.dpdu.z);
    o[4] = make_float4(u, v, 0.0f, 0.0f);
}

}  // namespace

hipError_t launch_surface(const ShadeSceneDev& sc, const float4* rays, const uint4* hits, uint64_t n, float4* out,
                          hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_surface, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, sc, rays, hits, n, out);
    return hipGetLastError();
}

}  // namespace bg
