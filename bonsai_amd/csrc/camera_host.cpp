// Host-side camera setup: PerspectiveCamera::new (cameras/perspective.rs:49-114)
// reduced to the two matrices the ray-generation kernel needs. Matrix algebra
// follows base/transform.rs and geometries/mat4.rs with f32 arithmetic in the
// reference's order (products and the Gauss-Jordan inverse), so the uploaded
// raster_to_camera matrix is the reference's bit for bit.
#include <cmath>
#include <cstring>

#include "bg_internal.h"

namespace bg {

struct M4 {
    float a[4][4];
    static M4 identity() {
        M4 m;
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) m.a[i][j] = (i == j) ? 1.0f : 0.0f;
        return m;
    }
};

// mat4.rs:37-48
M4 mul(const M4& x, const M4& y) {
    M4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            r.a[i][j] = x.a[i][0] * y.a[0][j] + x.a[i][1] * y.a[1][j] + x.a[i][2] * y.a[2][j] + x.a[i][3] * y.a[3][j];
    return r;
}

// mat4.rs:71-143: Gauss-Jordan elimination with full pivoting.
M4 invert(const M4& src) {
    M4 m = src;
    int col_of[4] = {0, 0, 0, 0}, row_of[4] = {0, 0, 0, 0}, used[4] = {0, 0, 0, 0};
    for (int i = 0; i < 4; ++i) {
        int prow = 0, pcol = 0;
        float best = 0.0f;
        for (int j = 0; j < 4; ++j) {
            if (used[j] == 1) continue;
            for (int k = 0; k < 4; ++k) {
                if (used[k] != 0) continue;
                float v = std::fabs(m.a[j][k]);
                if (v >= best) { best = v; prow = j; pcol = k; }
            }
        }
        used[pcol] += 1;
        if (prow != pcol)
            for (int k = 0; k < 4; ++k) std::swap(m.a[prow][k], m.a[pcol][k]);
        row_of[i] = prow;
        col_of[i] = pcol;
        const float inv_pivot = 1.0f / m.a[pcol][pcol];
        m.a[pcol][pcol] = 1.0f;
        for (int j = 0; j < 4; ++j) m.a[pcol][j] *= inv_pivot;
        for (int j = 0; j < 4; ++j) {
            if (j == pcol) continue;
            const float f = m.a[j][pcol];
            m.a[j][pcol] = 0.0f;
            for (int k = 0; k < 4; ++k) m.a[j][k] -= m.a[pcol][k] * f;
        }
    }
    for (int j = 3; j >= 0; --j) {
        if (row_of[j] == col_of[j]) continue;
        for (int k = 0; k < 4; ++k) std::swap(m.a[k][row_of[j]], m.a[k][col_of[j]]);
    }
    return m;
}

namespace {

struct Xf {
    M4 m, inv;
};
// transform.rs:1629-1637
Xf compose(const Xf& x, const Xf& y) { return Xf{mul(x.m, y.m), mul(y.inv, x.inv)}; }
// transform.rs:126-150
Xf scaling(float x, float y, float z) {
    Xf t{M4::identity(), M4::identity()};
    t.m.a[0][0] = x; t.m.a[1][1] = y; t.m.a[2][2] = z;
    t.inv.a[0][0] = 1.0f / x; t.inv.a[1][1] = 1.0f / y; t.inv.a[2][2] = 1.0f / z;
    return t;
}
// transform.rs:87-97
Xf translation(float x, float y, float z) {
    Xf t{M4::identity(), M4::identity()};
    t.m.a[0][3] = x; t.m.a[1][3] = y; t.m.a[2][3] = z;
    t.inv.a[0][3] = -x; t.inv.a[1][3] = -y; t.inv.a[2][3] = -z;
    return t;
}
// transform.rs:48-73; f32::to_radians multiplies by (PI / 180).
Xf perspective(float fov, float n, float f) {
    M4 p = M4::identity();
    p.a[2][2] = f / (f - n);
    p.a[2][3] = -f * n / (f - n);
    p.a[3][2] = 1.0f;
    p.a[3][3] = 0.0f;
    const float pi = 3.14159265358979323846f;
    const float inv_tan = 1.0f / std::tan(fov * (pi / 180.0f) / 2.0f);
    return compose(scaling(inv_tan, inv_tan, 1.0f), Xf{p, invert(p)});
}

}  // namespace

void camera_setup(const bg_camera_desc& d, CameraDev* out) {
    const float rx = (float)d.width, ry = (float)d.height;
    Xf camera_to_screen = perspective(d.fov, d.near_clip, d.far_clip);
    // perspective.rs:62-74
    float wmin_x, wmax_x, wmin_y, wmax_y;
    const float frame = rx / ry;
    if (frame > 1.0f) { wmin_x = -frame; wmax_x = frame; wmin_y = -1.0f; wmax_y = 1.0f; }
    else { wmin_x = -1.0f; wmax_x = 1.0f; wmin_y = -1.0f / frame; wmax_y = 1.0f / frame; }
    // perspective.rs:77-85
    Xf screen_to_raster = compose(compose(scaling(rx, ry, 1.0f),
                                          scaling(1.0f / (wmax_x - wmin_x), 1.0f / (wmin_y - wmax_y), 1.0f)),
                                  translation(-wmin_x, -wmax_y, 0.0f));
    Xf raster_to_screen{screen_to_raster.inv, screen_to_raster.m};
    Xf screen_to_camera{camera_to_screen.inv, camera_to_screen.m};
    Xf raster_to_camera = compose(screen_to_camera, raster_to_screen);
    std::memcpy(out->raster_to_camera, raster_to_camera.m.a, sizeof(out->raster_to_camera));
    std::memcpy(out->camera_to_world, d.camera_to_world, sizeof(out->camera_to_world));
    out->lens_radius = d.lens_radius;
    out->focal_distance = d.focal_distance;
    out->shutter_open = d.shutter_open;
    out->shutter_close = d.shutter_close;
    out->width = d.width;
    out->height = d.height;
}

}  // namespace bg

namespace {

// vec3.rs:108-122 (f64 cross) and vec3.rs:103-137 (length, normalize)
void cross64(const float a[3], const float b[3], float out[3]) {
    const double x = a[0], y = a[1], z = a[2], vx = b[0], vy = b[1], vz = b[2];
    out[0] = (float)((y * vz) - (z * vy));
    out[1] = (float)((z * vx) - (x * vz));
    out[2] = (float)((x * vy) - (y * vx));
}
float len3(const float a[3]) { return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }
void unit3(const float a[3], float out[3]) {
    const float inv = 1.0f / len3(a);
    out[0] = a[0] * inv; out[1] = a[1] * inv; out[2] = a[2] * inv;
}

}  // namespace

extern "C" int bg_look_at(const float pos[3], const float look[3], const float up[3], float m[16], float m_inv[16]) {
    if (!pos || !look || !up || !m || !m_inv) return BG_ERR_INVALID;
    // transform.rs:152-189
    bg::M4 c2w = bg::M4::identity();
    c2w.a[0][3] = pos[0]; c2w.a[1][3] = pos[1]; c2w.a[2][3] = pos[2]; c2w.a[3][3] = 1.0f;
    const float d0[3] = {look[0] - pos[0], look[1] - pos[1], look[2] - pos[2]};
    float dir[3], upn[3], r0[3], right[3], new_up[3];
    unit3(d0, dir);
    unit3(up, upn);
    cross64(upn, dir, r0);
    if (len3(r0) == 0.0f) {
        bg::M4 id = bg::M4::identity();
        std::memcpy(m, id.a, 64);
        std::memcpy(m_inv, id.a, 64);
        return BG_OK;
    }
    unit3(r0, right);
    cross64(dir, right, new_up);
    for (int i = 0; i < 3; ++i) {
        c2w.a[i][0] = right[i];
        c2w.a[i][1] = new_up[i];
        c2w.a[i][2] = dir[i];
    }
    c2w.a[3][0] = 0.0f; c2w.a[3][1] = 0.0f; c2w.a[3][2] = 0.0f;
    const bg::M4 inv = bg::invert(c2w);
    std::memcpy(m, c2w.a, 64);
    std::memcpy(m_inv, inv.a, 64);
    return BG_OK;
}
