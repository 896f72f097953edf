"""Pins the oracle's triangle and sphere intersection (triangle.rs, sphere.rs,
efloat.rs) against closed forms, invariants and brute force."""
import numpy as np
import pytest

import cpu_oracle as O
from bonsai_amd.scenes import SceneData, add_spheres, icosphere


def tri_scene(*tris):
    return O.OracleScene(SceneData(tri_positions=np.asarray(tris, np.float32).reshape(-1, 9)))


def test_axis_aligned_ray_right_triangle_exact():
    # p0=(0,0,5), p1=(1,0,5), p2=(0,1,5); a +z ray from (0.25, 0.25, 0): every step of the watertight test is
    # exact, so t = 5, b = (0.5, 0.25, 0.25) exactly.
    s = tri_scene([0, 0, 5, 1, 0, 5, 0, 1, 5])
    ok, out = s.prim(0, O.ray((0.25, 0.25, 0), (0, 0, 1)))
    assert ok
    assert out[0] == 5.0 and out[1] == 0.5 and out[2] == 0.25 and out[3] == 0.25


def test_t_max_boundary_is_inclusive():
    # triangle.rs:208: rejects t_scaled > t_max * det, so t == t_max is accepted.
    s = tri_scene([0, 0, 5, 1, 0, 5, 0, 1, 5])
    assert s.prim(0, O.ray((0.25, 0.25, 0), (0, 0, 1), t_max=5.0))[0]
    assert not s.prim(0, O.ray((0.25, 0.25, 0), (0, 0, 1), t_max=np.nextafter(np.float32(5), np.float32(0))))[0]


def test_misses():
    s = tri_scene([0, 0, 5, 1, 0, 5, 0, 1, 5])
    assert not s.prim(0, O.ray((0.75, 0.75, 0), (0, 0, 1)))[0]  # beyond the hypotenuse
    assert not s.prim(0, O.ray((0.25, 0.25, 0), (0, 0, -1)))[0]  # pointing away (t <= 0)
    assert not s.prim(0, O.ray((0.25, 0.25, 6), (0, 0, 1)))[0]  # behind the origin
    assert not s.prim(0, O.ray((0.25, 0.25, 5), (0, 0, 1)))[0]  # origin on the plane: t = 0 <= delta_t


def test_degenerate_triangle_misses():
    # collinear vertices: det == 0 or the zero-area exit (triangle.rs:196-199, 258-264)
    s = tri_scene([0, 0, 5, 1, 0, 5, 2, 0, 5])
    for o in [(0.5, 0.0, 0.0), (0.5, 1e-3, 0.0), (1.0, 0.0, 0.0)]:
        assert not s.prim(0, O.ray(o, (0, 0, 1)))[0]
        assert not s.prim(0, O.ray(o, (0, 0, 1)), kind=1)[0]


def test_barycentrics_reconstruct_hit_point():
    rng = np.random.default_rng(11)
    tris = rng.uniform(-1, 1, (200, 9)).astype(np.float32)
    tris[:, 2::3] += 4.0
    s = tri_scene(*tris)
    hits = 0
    for i in range(200):
        p = tris[i].reshape(3, 3).astype(np.float64)
        w = rng.dirichlet([1, 1, 1])
        target = w @ p
        d = target / np.linalg.norm(target)
        ok, out = s.prim(i, O.ray((0, 0, 0), d))
        if not ok:
            continue
        hits += 1
        t, b0, b1, b2 = map(float, out)
        assert abs(b0 + b1 + b2 - 1) < 1e-5
        np.testing.assert_allclose(b0 * p[0] + b1 * p[1] + b2 * p[2], t * d, atol=2e-5)
    assert hits > 190


def test_intersect_and_test_agree():
    rng = np.random.default_rng(12)
    tris = rng.uniform(-1, 1, (300, 9)).astype(np.float32)
    s = tri_scene(*tris)
    for i in range(300):
        r = O.ray(rng.uniform(-2, 2, 3), rng.standard_normal(3))
        a, oa = s.prim(i, r, 0)
        b, ob = s.prim(i, r, 1)
        assert a == b
        if a:
            assert oa[0] == ob[0]


def test_watertight_closed_mesh():
    # Rays from the centre of a closed icosphere through every vertex and every edge midpoint pass exactly
    # through shared edges/vertices; a watertight test never lets them escape (triangle.rs:135-240).
    tris = icosphere(2, (0.0, 0.0, 0.0), 1.0, 0.0, 1.0, 1)
    s = O.OracleScene(SceneData(tri_positions=tris))
    v = tris.reshape(-1, 3, 3)
    targets = np.concatenate([v.reshape(-1, 3), ((v + np.roll(v, 1, axis=1)) / 2).reshape(-1, 3)])
    rays = np.array([O.ray((0, 0, 0), t) for t in targets], np.float32)
    hits = s.brute(rays, np.arange(tris.shape[0], dtype=np.uint32))
    assert np.all(hits[:, 0] != 0xFFFFFFFF)
    # and from an off-centre interior point in random directions
    rng = np.random.default_rng(4)
    d = rng.standard_normal((2000, 3))
    rays = np.array([O.ray((0.1, -0.05, 0.2), x) for x in d], np.float32)
    hits = s.brute(rays, np.arange(tris.shape[0], dtype=np.uint32))
    assert np.all(hits[:, 0] != 0xFFFFFFFF)


def sphere_scene(centers, radii):
    s = SceneData()
    add_spheres(s, centers, radii)
    return O.OracleScene(s)


def test_sphere_through_centre():
    # |c - o| = 5, r = 1: the entry distance is 4; the origin is pushed along d by its transform error
    # (ray.rs:87-93) so t is 4 within a few gamma(3) * 5.
    s = sphere_scene([(0, 0, 5)], [1.0])
    ok, out = s.prim(0, O.ray((0, 0, 0), (0, 0, 1)))
    assert ok and abs(out[0] - 4.0) < 1e-5
    ok, out = s.prim(0, O.ray((0, 0, 0), (0, 0, 1)), kind=1)
    assert ok and abs(out[0] - 4.0) < 1e-5


def test_sphere_inside_returns_exit():
    s = sphere_scene([(0, 0, 0)], [2.0])
    ok, out = s.prim(0, O.ray((0, 0, 0.5), (0, 0, 1)))
    assert ok and abs(out[0] - 1.5) < 1e-5  # t0 < 0 -> t1 (sphere.rs:110-116)


def test_sphere_misses_and_tmax():
    s = sphere_scene([(0, 0, 5)], [1.0])
    assert not s.prim(0, O.ray((1.01, 0, 0), (0, 0, 1)))[0]
    assert not s.prim(0, O.ray((0, 0, 0), (0, 0, -1)))[0]
    assert not s.prim(0, O.ray((0, 0, 0), (0, 0, 1), t_max=3.9))[0]  # t0.upper_bound > t_max


def test_sphere_uv_closed_form():
    # hitting the +x side at z=0 gives phi = 0 -> u = 0 and theta = pi/2 -> v = 0.5 (sphere.rs:166-168)
    s = sphere_scene([(0, 0, 0)], [1.0])
    ok, out = s.prim(0, O.ray((5, 0, 0), (-1, 0, 0)))
    assert ok
    assert abs(out[0] - 4.0) < 1e-5
    assert abs(out[1]) < 1e-6 and abs(out[2] - 0.5) < 1e-6


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_sphere_hits_match_analytic(seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-1, 1, 3)
    r = float(rng.uniform(0.2, 1.0))
    s = sphere_scene([c], [r])
    for _ in range(300):
        o = rng.uniform(-4, 4, 3)
        d = rng.standard_normal(3)
        d /= np.linalg.norm(d)
        oc = o - c
        b = oc @ d
        disc = b * b - (oc @ oc - r * r)
        ok, out = s.prim(0, O.ray(o, d))
        if disc < -1e-4:
            assert not ok
        elif disc > 1e-4:
            t0, t1 = -b - np.sqrt(disc), -b + np.sqrt(disc)
            if t1 <= 1e-4:
                assert not ok
                continue
            if t0 > 1e-4 or t0 < -1e-4:
                assert ok
                exp = t0 if t0 > 0 else t1
                assert abs(out[0] - exp) < 5e-5 * max(1.0, abs(exp))
