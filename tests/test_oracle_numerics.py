"""Pins the oracle's f32 numerics (math.rs, mat4.rs, transform.rs) against
closed forms and exact arithmetic."""
from fractions import Fraction

import numpy as np
import pytest

import cpu_oracle as O


def f32_round(x: Fraction) -> np.float32:
    """Correctly rounded f32 of an exact rational (round half to even)."""
    f = np.float32(float(x))
    # float(x) rounds to f64 first; f64 -> f32 double rounding is exact for these magnitudes
    # unless the f64 value sits on an f32 tie, which the inputs below avoid.
    return f


@pytest.mark.parametrize("n", [1.0, 2.0, 3.0, 5.0, 6.0, 7.0])
def test_gamma_closed_form(n):
    # gamma(n) = n*u / (1 - n*u), u = 2^-24 (math.rs:3-5): numerator and denominator are exact in f32,
    # so the result is the correctly rounded quotient.
    u = Fraction(1, 2 ** 24)
    exact = (Fraction(n) * u) / (1 - Fraction(n) * u)
    assert np.float32(O.lib().orc_gamma(n)) == f32_round(exact)


def test_next_float_matches_nextafter():
    rng = np.random.default_rng(3)
    vals = np.concatenate([rng.standard_normal(2000).astype(np.float32) * np.float32(1e3),
                           np.array([1.0, -1.0, 1e-38, -1e-38, 3.4e38, -3.4e38, 1e-45, -1e-45], np.float32)])
    up = np.array([O.lib().orc_next_float_up(float(v)) for v in vals], np.float32)
    dn = np.array([O.lib().orc_next_float_down(float(v)) for v in vals], np.float32)
    assert np.array_equal(up, np.nextafter(vals, np.float32(np.inf)))
    assert np.array_equal(dn, np.nextafter(vals, np.float32(-np.inf)))


def test_next_float_special_cases():
    L = O.lib()
    tiny = np.float32(1.4e-45)
    assert np.float32(L.orc_next_float_up(0.0)) == tiny          # math.rs:47-48: -0 and +0 step to +denorm
    assert np.float32(L.orc_next_float_up(-0.0)) == tiny
    assert np.float32(L.orc_next_float_down(0.0)) == -tiny        # math.rs:31-32
    assert L.orc_next_float_up(float("inf")) == float("inf")
    assert L.orc_next_float_down(float("-inf")) == float("-inf")


def test_mat4_inverse_exact_cases():
    # diagonal powers of two and a permutation invert exactly
    m = np.diag(np.array([2.0, 4.0, 0.5, 1.0], np.float32)).reshape(16)
    out = np.zeros(16, np.float32)
    O.lib().orc_mat4_inverse(np.ascontiguousarray(m).ctypes.data, out.ctypes.data)
    assert np.array_equal(out.reshape(4, 4), np.diag(np.array([0.5, 0.25, 2.0, 1.0], np.float32)))
    p = np.zeros((4, 4), np.float32)
    p[0, 2] = p[1, 0] = p[2, 3] = p[3, 1] = 1.0
    O.lib().orc_mat4_inverse(np.ascontiguousarray(p.reshape(16)).ctypes.data, out.ctypes.data)
    assert np.array_equal(out.reshape(4, 4), p.T)


def test_mat4_inverse_random():
    rng = np.random.default_rng(5)
    for _ in range(50):
        m = (rng.standard_normal((4, 4)) + 4 * np.eye(4)).astype(np.float32)
        out = np.zeros(16, np.float32)
        O.lib().orc_mat4_inverse(np.ascontiguousarray(m.reshape(16)).ctypes.data, out.ctypes.data)
        np.testing.assert_allclose(out.reshape(4, 4).astype(np.float64) @ m.astype(np.float64), np.eye(4),
                                   atol=2e-5)


def test_look_at_axis_aligned():
    # looking down +z from the origin with +y up: columns are (right, up, dir) = (x, y, z)
    m, mi = O.look_at((0, 0, 0), (0, 0, 1), (0, 1, 0))
    assert np.array_equal(m.reshape(4, 4), np.eye(4, dtype=np.float32))
    assert np.array_equal(mi.reshape(4, 4), np.eye(4, dtype=np.float32))
    m, mi = O.look_at((1, 2, 3), (1, 2, 10), (0, 1, 0))
    e = np.eye(4, dtype=np.float32)
    e[:3, 3] = [1, 2, 3]
    assert np.array_equal(m.reshape(4, 4), e)
    np.testing.assert_array_equal(mi.reshape(4, 4)[:3, 3], [-1, -2, -3])


def test_look_at_degenerate_up_gives_identity():
    m, _ = O.look_at((0, 0, 0), (0, 1, 0), (0, 1, 0))  # transform.rs:163-166
    assert np.array_equal(m.reshape(4, 4), np.eye(4, dtype=np.float32))


def test_camera_center_and_fov_edge():
    # perspective.rs:117-175: the centre pixel's ray is the view direction; the ray through the middle of
    # the left film edge makes fov/2 with it (square film -> screen window [-1, 1]^2).
    m, mi = O.look_at((0, 0, 0), (0, 0, 1), (0, 1, 0))
    fov = 60.0
    rays = O.camera_rays(m, mi, fov, 3, 3)
    d = rays[4, 4:7]
    np.testing.assert_allclose(d, [0, 0, 1], atol=1e-7)
    edge = O.camera_rays(m, mi, fov, 2, 2, fx=0.0, fy=1.0)[0, 4:7]  # film point (0, 1): left edge, centre row
    ang = np.degrees(np.arccos(np.clip(edge.astype(np.float64) @ np.array([0, 0, 1.0]), -1, 1)))
    assert abs(ang - fov / 2) < 1e-4
    assert edge[0] < 0  # raster x = 0 is screen x = -1 (perspective.rs:77-83)
