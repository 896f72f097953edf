"""The oracle's DirectLightingIntegrator (direct.rs:72-113 with integrator.rs:274-328)
and WhittedIntegrator (whitted.rs:41-93), pinned by closed forms: the radiance of
a Lambertian floor point under a square area light (form factor of a parallel
rectangle), the reference semantics' missing 1/area in the light pdf (Whitted
with a double-sided light is exactly 1/area times the corrected estimate on the
same samples), the two light strategies and the path integrator at depth 1
agreeing in expectation, and an occluder's shadow."""
import numpy as np
import pytest

import cpu_oracle as O
from bonsai_amd.integrator import resolve
from bonsai_amd.scenes import Camera, SceneData, area_light, matte, quad

DIRECT, WHITTED = 1, 2
ALL, ONE = 0, 1
REF, CORRECTED = 0, 1
KD, LE, SIDE, H = 0.5, 10.0, 0.5, 1.0


def lit_floor(double_sided=False, blocker=False):
    s = 0.5 * SIDE
    floor = quad((-5, 0, -5), (-5, 0, 5), (5, 0, 5), (5, 0, -5))          # normal +y
    light = quad((-s, H, -s), (s, H, -s), (s, H, s), (-s, H, s))            # normal -y
    parts, mats = [floor, light], [0, 0, 1, 1]
    if blocker:  # a quad halfway up, wider than the light's cone from the floor points the camera sees
        b = 0.15
        parts.append(quad((-b, 0.5, -b), (-b, 0.5, b), (b, 0.5, b), (b, 0.5, -b)))
        mats += [0, 0]
    sc = SceneData(tri_positions=np.concatenate(parts).astype(np.float32))
    sc.materials = [matte(KD, KD, KD), matte(0.0, 0.0, 0.0)]
    sc.tri_material = np.array(mats, np.uint32)
    sc.lights = [area_light(2, (LE, LE, LE), double_sided), area_light(3, (LE, LE, LE), double_sided)]
    # a narrow view of the floor point under the light's centre, from below the light's plane
    sc.camera = Camera(pos=(0.0, 0.6, -1.5), look=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov=0.5, width=2, height=2)
    return sc


def render(sc, integrator, strategy=ALL, mode=CORRECTED, spp=4096, seed=3, depth=1):
    o = O.OracleScene(sc)
    o.build()
    o.set_shading(sc)
    rgb, w = o.render(sc, spp, depth, mode=mode, seed=seed, integrator=integrator, strategy=strategy)
    return resolve(rgb, w)


def form_factor_centre(side, h):
    """Point-to-parallel-square form factor under the square's centre: four corner-aligned
    quadrants of (side / 2) x (side / 2) at distance h."""
    x = y = 0.5 * side / h
    q = (x / np.sqrt(1 + x * x) * np.arctan(y / np.sqrt(1 + x * x))
         + y / np.sqrt(1 + y * y) * np.arctan(x / np.sqrt(1 + y * y))) / (2 * np.pi)
    return 4 * q


def test_form_factor_limits():
    assert form_factor_centre(1e6, 1.0) == pytest.approx(1.0, rel=1e-5)          # an infinite emitter
    small = form_factor_centre(0.01, 1.0)
    assert small == pytest.approx(0.01 ** 2 / np.pi, rel=1e-3)                     # a small one: A cos cos / (pi d^2)


@pytest.mark.parametrize("integrator,strategy", [(DIRECT, ALL), (DIRECT, ONE), (WHITTED, ALL)])
def test_floor_under_square_light(integrator, strategy):
    expected = KD * LE * form_factor_centre(SIDE, H)
    img = render(lit_floor(), integrator, strategy, spp=16384)
    # every pixel sees the floor within 0.01 of the centre, where the irradiance is flat to 1e-3
    np.testing.assert_allclose(img.mean(axis=(0, 1)), expected, rtol=0.01)


def test_reference_light_pdf_misses_one_over_area():
    sc = lit_floor(double_sided=True)
    c = render(sc, WHITTED, mode=CORRECTED, spp=64)
    r = render(sc, WHITTED, mode=REF, spp=64)
    tri_area = 0.5 * SIDE * SIDE
    np.testing.assert_allclose(r, c / tri_area, rtol=1e-5)
    # a single-sided light facing the floor: the reference evaluates emission toward +wi (away
    # from the light), so its light samples carry nothing
    assert not render(lit_floor(), WHITTED, mode=REF, spp=64).any()


def test_strategies_and_path_depth_one_agree():
    sc = lit_floor()
    a = render(sc, DIRECT, ALL, spp=16384, seed=5).mean()
    o = render(sc, DIRECT, ONE, spp=16384, seed=6).mean()
    p = render(sc, 0, spp=16384, seed=7).mean()  # path integrator, max depth 1
    assert o == pytest.approx(a, rel=0.01)
    assert p == pytest.approx(a, rel=0.02)


def test_blocker_casts_a_shadow():
    # at half height the light's cone from the floor point is 0.25 wide; the blocker is 0.3 wide
    for integrator, strategy in ((WHITTED, ALL), (DIRECT, ALL), (DIRECT, ONE)):
        assert not render(lit_floor(blocker=True), integrator, strategy, spp=256).any()
