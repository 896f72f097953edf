"""Area lights on spheres in the oracle (sphere.rs:328-377 through shape.rs:17-57
and diffuse.rs:134-167), pinned by closed forms: the uniform-cone pdf
1 / (2 pi (1 - sqrt(1 - r^2 / d^2))) from outside; 1 / (4 pi) from the centre;
the reference semantics' unnormalised sample normal (length r), its pdf
d^2 / |n . -wi| without 1/area (Q3) and emission toward +wi (Q4); points on
the sphere uniform in (z, phi) (Archimedes) for area sampling and directions
uniform in the cone for the corrected cone sampling (chi-square); spawn_ray's
t_max 0.9999 hiding the sphere from its own centre once r > 0.9999 (Q2); and
the furnace closed form inside an emitting sphere."""
import numpy as np
import pytest
from scipy.stats import chi2

import cpu_oracle as O
from bonsai_amd.scenes import Camera, SceneData, add_spheres, area_light, matte

REF, CORRECTED = 0, 1
C = np.array([0.3, -0.2, 0.5])


def sphere_scene(r, double_sided=True, le=2.0, reverse=False):
    s = SceneData()
    add_spheres(s, [C], [r], [0])
    if reverse:
        s.sphere_reverse = np.ones(1, np.uint8)
    s.materials = [matte(0.5, 0.5, 0.5)]
    s.lights = [area_light(0, (le, le, le), double_sided)]
    o = O.OracleScene(s)
    o.build()
    o.set_shading(s)
    return o


def outside_points(rng, n, r, dmin=1.2, dmax=10.0):
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    dist = r * rng.uniform(dmin, dmax, (n, 1))
    x = (C + d * dist).astype(np.float32)
    return x, np.zeros_like(x), np.tile(np.float32([0, 0, 1]), (n, 1))


def cone_pdf(r, d2):
    return 1.0 / (2.0 * np.pi * (1.0 - np.sqrt(np.maximum(0.0, 1.0 - r * r / d2))))


def chi2_uniform(a, b, na=16, nb=16):
    """p-value of (a, b) in [0, 1)^2 being uniform on an na x nb grid."""
    h, _, _ = np.histogram2d(a, b, bins=[na, nb], range=[[0, 1], [0, 1]])
    e = a.size / (na * nb)
    return chi2.sf(float(((h - e) ** 2 / e).sum()), na * nb - 1)


def frame(w):
    """An orthonormal frame around each row of w (any frame: phi uniformity does not depend on it)."""
    a = np.where(np.abs(w[:, :1]) > 0.5, [[0.0, 1.0, 0.0]], [[1.0, 0.0, 0.0]])
    x = np.cross(w, a)
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    return x, np.cross(w, x)


@pytest.mark.parametrize("mode", [REF, CORRECTED])
def test_cone_pdf_from_outside(mode):
    r = 0.7
    o = sphere_scene(r)
    rng = np.random.default_rng(1)
    x, e, n = outside_points(rng, 4000, r)
    wi = rng.standard_normal((4000, 3))  # the cone pdf does not depend on wi (sphere.rs:366-372)
    wi /= np.linalg.norm(wi, axis=1, keepdims=True)
    got = o.light_pdf(0, mode, x, e, n, wi).astype(np.float64)
    d2 = ((x.astype(np.float64) - C) ** 2).sum(1)
    np.testing.assert_allclose(got, cone_pdf(np.float32(r), d2), rtol=5e-4)


@pytest.mark.parametrize("reverse", [False, True])
def test_corrected_cone_samples(reverse):
    r = 0.7
    o = sphere_scene(r, double_sided=False, reverse=reverse)
    rng = np.random.default_rng(2)
    m = 60000
    x0, e0, n0 = outside_points(rng, 1, r, 2.0, 2.0)
    x, e, n = np.repeat(x0, m, 0), np.repeat(e0, m, 0), np.repeat(n0, m, 0)
    u0, u1 = rng.uniform(0, 1, m), rng.uniform(0, 1, m)
    out = o.light_sample(0, CORRECTED, x, e, n, u0, u1).astype(np.float64)
    p, nn, pdf, rad = out[:, 0:3], out[:, 3:6], out[:, 6], out[:, 7:10]
    xc = x0[0].astype(np.float64)
    d2 = ((xc - C) ** 2).sum()
    # points on the sphere, on the side facing x (the first point each sampled direction meets)
    np.testing.assert_allclose(np.linalg.norm(p - C, axis=1), r, rtol=2e-6)
    facing = np.einsum("ij,ij->i", (p - C) / r, (xc - p) / np.linalg.norm(xc - p, axis=1, keepdims=True))
    assert facing.min() > -1e-4
    # unit normals, outward (inward when reverse-oriented)
    np.testing.assert_allclose(nn, ((p - C) / r) * (-1.0 if reverse else 1.0), atol=2e-6)
    np.testing.assert_allclose(pdf, cone_pdf(np.float32(r), d2), rtol=5e-5)
    # one-sided emission toward -wi: the facing side emits unless the orientation is reversed
    assert np.all(rad == (0.0 if reverse else 2.0))
    # directions uniform in the cone: (1 - cos theta) / (1 - cos theta_max) and phi uniform
    wc = (C - xc) / np.sqrt(d2)
    wi = (p - xc) / np.linalg.norm(p - xc, axis=1, keepdims=True)
    cos_max = np.sqrt(1.0 - r * r / d2)
    cos_t = wi @ wc
    assert cos_t.min() >= cos_max - 1e-6
    fx, fy = frame(wc[None, :])
    phi = np.mod(np.arctan2(wi @ fy[0], wi @ fx[0]), 2 * np.pi) / (2 * np.pi)
    t = np.clip((1.0 - cos_t) / (1.0 - cos_max), 0.0, 1.0 - 1e-12)
    assert chi2_uniform(t, phi) > 1e-3


def test_area_samples_in_both_semantics():
    r = 0.7
    o = sphere_scene(r, double_sided=False)
    rng = np.random.default_rng(3)
    m = 60000
    x0, e0, n0 = outside_points(rng, 1, r, 3.0, 3.0)
    x, e, n = np.repeat(x0, m, 0), np.repeat(e0, m, 0), np.repeat(n0, m, 0)
    u0, u1 = rng.uniform(0, 1, m), rng.uniform(0, 1, m)
    out = o.light_sample(0, REF, x, e, n, u0, u1).astype(np.float64)
    p, nn, pdf, rad = out[:, 0:3], out[:, 3:6], out[:, 6], out[:, 7:10]
    xc = x0[0].astype(np.float64)
    ok = pdf > 0
    assert ok.mean() > 0.99
    p, nn, pdf, rad = p[ok], nn[ok], pdf[ok], rad[ok]
    np.testing.assert_allclose(np.linalg.norm(p - C, axis=1), r, rtol=2e-6)
    # the reference's sample normal is the transformed, unnormalised object point: length r
    np.testing.assert_allclose(np.linalg.norm(nn, axis=1), r, rtol=2e-6)
    np.testing.assert_allclose(nn / r, (p - C) / r, atol=2e-6)
    # Q3: pdf = d^2 / |n . -wi| with the unnormalised n, no 1/area
    v = p - xc
    d = np.linalg.norm(v, axis=1)
    cosn = np.abs(np.einsum("ij,ij->i", nn / r, -v / d[:, None]))
    away = cosn > 1e-2  # f32 cancellation near grazing
    np.testing.assert_allclose(pdf[away], (d * d / (r * cosn))[away], rtol=3e-5)
    # Q4: a one-sided light emits toward +wi, i.e. only the samples on the side facing away from x
    back = np.einsum("ij,ij->i", p - C, v) > 0
    assert np.array_equal(rad[:, 0] > 0, back)
    # uniform on the sphere: z and phi of the unit offset uniform (Archimedes)
    q = (p - C) / r
    z = np.clip((q[:, 2] + 1.0) / 2.0, 0.0, 1.0 - 1e-12)
    phi = np.mod(np.arctan2(q[:, 1], q[:, 0]), 2 * np.pi) / (2 * np.pi)
    assert chi2_uniform(z, phi) > 1e-3


@pytest.mark.parametrize("r", [0.5, 2.0])
def test_from_the_centre(r):
    o = sphere_scene(r)
    rng = np.random.default_rng(4)
    m = 2000
    x = np.tile(C.astype(np.float32), (m, 1))
    e = np.zeros_like(x)
    n = np.tile(np.float32([0, 0, 1]), (m, 1))
    wi = rng.standard_normal((m, 3))
    wi /= np.linalg.norm(wi, axis=1, keepdims=True)
    inv4pi = 1.0 / (4.0 * np.pi)
    # shape.rs:40-57 from inside: d^2 / (|n . -wi| area) = r^2 / (4 pi r^2); the reference's spawn_ray
    # stops at t 0.9999, short of a sphere of radius 2 (Q2)
    cor = o.light_pdf(0, CORRECTED, x, e, n, wi)
    np.testing.assert_allclose(cor, inv4pi, rtol=2e-5)
    ref = o.light_pdf(0, REF, x, e, n, wi)
    if r < 0.9999:
        np.testing.assert_allclose(ref, inv4pi, rtol=2e-5)
    else:
        assert np.all(ref == 0.0)
    u0, u1 = rng.uniform(0, 1, m), rng.uniform(0, 1, m)
    sc = o.light_sample(0, CORRECTED, x, e, n, u0, u1)
    np.testing.assert_allclose(sc[:, 6], inv4pi, rtol=2e-5)
    # reference: d^2 / |n . -wi| = r^2 / r with the normal of length r
    sr = o.light_sample(0, REF, x, e, n, u0, u1)
    np.testing.assert_allclose(sr[:, 6], r, rtol=2e-5)


def test_corrected_furnace_inside_an_emitting_sphere():
    """Inside a double-sided emitting sphere (radiance 1) that is also a matte reflector of albedo 0.5:
    L = sum_{k<=d} 0.5^k at max depth d, with light sampling from inside the sphere."""
    s = SceneData()
    add_spheres(s, [(0.0, 0.0, 0.0)], [0.8], [0])
    s.materials = [matte(0.5, 0.5, 0.5)]
    s.lights = [area_light(0, (1.0, 1.0, 1.0), True)]
    s.camera = Camera(pos=(0.05, 0.02, 0.01), look=(1.0, 0.3, 0.2), up=(0.0, 1.0, 0.0), fov=60.0, width=16,
                      height=16)
    o = O.OracleScene(s)
    o.build()
    o.set_shading(s)
    for depth in (0, 1, 5):
        rgb, w = o.render(s, 64, depth, mode=CORRECTED)
        img = rgb / w[..., None]
        exact = sum(0.5 ** k for k in range(depth + 1))
        assert abs(img.mean() - exact) < 0.01 * exact, (depth, img.mean(), exact)
