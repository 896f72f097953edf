// Device restatement of PerspectiveCamera::generate_ray for the pinhole case
// (cameras/perspective.rs:117-175) followed by the static Ray::transform
// (geometries/ray.rs:50-76). f32 operation order follows point3.rs:51-69,
// vec3.rs:35-44 and point3.rs:71-104.
#pragma once

#include <hip/hip_runtime.h>

#include "bg_internal.h"
#include "dev_geom.h"

namespace bg {
namespace dev {

struct CamRay {
    float ox, oy, oz, t_max, dx, dy, dz, time;
};

__device__ __forceinline__ CamRay generate_ray(const CameraDev& c, float film_x, float film_y, float time_u) {
    const float* m = c.raster_to_camera;
    // Point3::transform of (film_x, film_y, 0)
    const float z0 = 0.0f;
    const float xp = m[0] * film_x + m[1] * film_y + m[2] * z0 + m[3];
    const float yp = m[4] * film_x + m[5] * film_y + m[6] * z0 + m[7];
    const float zp = m[8] * film_x + m[9] * film_y + m[10] * z0 + m[11];
    const float wp = m[12] * film_x + m[13] * film_y + m[14] * z0 + m[15];
    float cx = xp, cy = yp, cz = zp;
    if (wp != 1.0f) {
        const float iw = 1.0f / wp;
        cx = xp * iw; cy = yp * iw; cz = zp * iw;
    }
    // normalize (vec3.rs:135-137): multiply by the reciprocal length
    const float inv_len = 1.0f / __builtin_sqrtf(cx * cx + cy * cy + cz * cz);
    const float dcx = cx * inv_len, dcy = cy * inv_len, dcz = cz * inv_len;
    // math.rs:23-25 lerp
    const float time = (1.0f - time_u) * c.shutter_open + time_u * c.shutter_close;
    // Ray::transform by camera_to_world of origin (0,0,0)
    const float* w = c.camera_to_world;
    const float px = 0.0f, py = 0.0f, pz = 0.0f;
    const float oxp = (w[0] * px + w[1] * py) + (w[2] * pz + w[3]);
    const float oyp = (w[4] * px + w[5] * py) + (w[6] * pz + w[7]);
    const float ozp = (w[8] * px + w[9] * py) + (w[10] * pz + w[11]);
    const float owp = (w[12] * px + w[13] * py) + (w[14] * pz + w[15]);
    const float ex = kGamma3 * (__builtin_fabsf(w[0] * px) + __builtin_fabsf(w[1] * py) + __builtin_fabsf(w[2] * pz) + __builtin_fabsf(w[3]));
    const float ey = kGamma3 * (__builtin_fabsf(w[4] * px) + __builtin_fabsf(w[5] * py) + __builtin_fabsf(w[6] * pz) + __builtin_fabsf(w[7]));
    const float ez = kGamma3 * (__builtin_fabsf(w[8] * px) + __builtin_fabsf(w[9] * py) + __builtin_fabsf(w[10] * pz) + __builtin_fabsf(w[11]));
    float ox = oxp, oy = oyp, oz = ozp;
    if (owp != 1.0f) {
        const float iw = 1.0f / owp;
        ox = oxp * iw; oy = oyp * iw; oz = ozp * iw;
    }
    const float dx = w[0] * dcx + w[1] * dcy + w[2] * dcz;
    const float dy = w[4] * dcx + w[5] * dcy + w[6] * dcz;
    const float dz = w[8] * dcx + w[9] * dcy + w[10] * dcz;
    float t_max = __builtin_inff();
    const float l2 = dx * dx + dy * dy + dz * dz;
    if (l2 > 0.0f) {
        const float dt = (__builtin_fabsf(dx) * ex + __builtin_fabsf(dy) * ey + __builtin_fabsf(dz) * ez) / l2;
        ox = ox + dx * dt; oy = oy + dy * dt; oz = oz + dz * dt;
        t_max -= dt;
    }
    return CamRay{ox, oy, oz, t_max, dx, dy, dz, time};
}

}  // namespace dev
}  // namespace bg
