// Oracle (test infrastructure only): PerspectiveCamera restated from
// cameras/perspective.rs:49-175 with base/transform.rs and base/sampling.rs.
#pragma once

#include "rt.h"

namespace orc {

struct PerspCamera {
    Transform camera_to_world;
    Transform raster_to_camera;
    Float shutter_open, shutter_close, lens_radius, focal_distance;
};

// perspective.rs:49-114
inline PerspCamera make_camera(const Mat4& c2w, const Mat4& c2w_inv, Float fov, Float near_clip, Float far_clip,
                               Float shutter_open, Float shutter_close, Float lens_radius, Float focal_distance,
                               Float res_x, Float res_y) {
    PerspCamera c;
    c.camera_to_world = Transform{c2w, c2w_inv};
    Transform camera_to_screen = perspective(fov, near_clip, far_clip);
    Float smin_x, smax_x, smin_y, smax_y;
    Float frame = res_x / res_y;
    if (frame > 1.0f) { smin_x = -frame; smax_x = frame; smin_y = -1.0f; smax_y = 1.0f; }
    else { smin_x = -1.0f; smax_x = 1.0f; smin_y = -1.0f / frame; smax_y = 1.0f / frame; }
    Transform screen_to_raster = tmul(tmul(scale(res_x, res_y, 1.0f),
                                           scale(1.0f / (smax_x - smin_x), 1.0f / (smin_y - smax_y), 1.0f)),
                                      translate(V3(-smin_x, -smax_y, 0.0f)));
    Transform raster_to_screen = tinverse(screen_to_raster);
    c.raster_to_camera = tmul(tinverse(camera_to_screen), raster_to_screen);
    c.shutter_open = shutter_open;
    c.shutter_close = shutter_close;
    c.lens_radius = lens_radius;
    c.focal_distance = focal_distance;
    return c;
}

// sampling.rs:181-198
inline void concentric_sample_disk(Float u0, Float u1, Float* dx, Float* dy) {
    Float ox = 2.0f * u0 - 1.0f, oy = 2.0f * u1 - 1.0f;
    if (ox == 0.0f && oy == 0.0f) { *dx = 0.0f; *dy = 0.0f; return; }
    Float theta, r;
    if (std::fabs(ox) > std::fabs(oy)) { theta = (PI / 4.0f) * (oy / ox); r = ox; }
    else { theta = PI_OVER_TWO - (PI / 4.0f) * (ox / oy); r = oy; }
    *dx = r * std::cos(theta);
    *dy = r * std::sin(theta);
}

// perspective.rs:117-175 (differentials omitted: no hot-path result reads them).
inline Ray camera_generate_ray(const PerspCamera& c, Float film_x, Float film_y, Float lens_u, Float lens_v, Float time_u) {
    V3 camera_point = xform_point(c.raster_to_camera, V3(film_x, film_y, 0.0f));
    Ray ray{V3(0.0f, 0.0f, 0.0f), normalize(camera_point), INFINITY, 0.0f};
    if (c.lens_radius > 0.0f) {
        Float lx, ly;
        concentric_sample_disk(lens_u, lens_v, &lx, &ly);
        lx = c.lens_radius * lx; ly = c.lens_radius * ly;
        Float ft = c.focal_distance / ray.d.z;
        V3 focus = ray.o + ray.d * ft;
        ray.o = V3(lx, ly, 0.0f);
        ray.d = normalize(focus - ray.o);
    }
    ray.time = lerp(time_u, c.shutter_open, c.shutter_close);
    return xform_ray(c.camera_to_world, ray);
}

}  // namespace orc
