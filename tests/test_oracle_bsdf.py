"""The oracle's multi-lobe BSDF and PlasticMaterial (materials/plastic.rs:24-52,
base/bsdf.rs:84-258, bxdf.rs:165-188 fresnel_dielectric, microfacet.rs:45-53
roughness_to_alpha), pinned by closed forms: the Lambertian lobe alone, the
microfacet lobe alone at normal incidence (D = 1/(pi a^2), G = 1, F = 0.04)
and at an oblique pair against the Trowbridge-Reitz D, G and the Fresnel
equations written here from the formulas, the sum / average rules across
lobes, BSDF::sample agreeing with f and pdf, and the roughness remap."""
import numpy as np
import pytest

import cpu_oracle as O
from bonsai_amd.scenes import matte, plastic

CORRECTED = 1


def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def tr_d(wh, a):
    c2 = wh[2] ** 2
    t2 = (1 - c2) / c2
    return 1.0 / (np.pi * a * a * c2 * c2 * (1 + t2 / (a * a)) ** 2)


def tr_lambda(w, a):
    c2 = w[2] ** 2
    t2 = (1 - c2) / c2
    return (-1 + np.sqrt(1 + a * a * t2)) / 2


def fresnel_dielectric(c, ei, et):
    # Fresnel equations for unpolarised light (the average of the two polarisations)
    if c <= 0:
        ei, et, c = et, ei, -c
    s = ei / et * np.sqrt(max(0.0, 1 - c * c))
    if s >= 1:
        return 1.0
    ct = np.sqrt(max(0.0, 1 - s * s))
    par = (et * c - ei * ct) / (et * c + ei * ct)
    per = (ei * c - et * ct) / (ei * c + et * ct)
    return (par * par + per * per) / 2


def test_lambertian_lobe_alone():
    m = plastic((0.5, 0.25, 0.125), (0, 0, 0), 0.3)
    wo, wi = unit([0.3, -0.2, 0.8]), unit([-0.5, 0.1, 0.6])
    r = O.bsdf_probe(m, CORRECTED, wo, wi)
    assert r["ncomp"] == 1
    np.testing.assert_allclose(r["f"], np.array([0.5, 0.25, 0.125]) / np.pi, rtol=1e-6)
    assert r["pdf"] == pytest.approx(wi[2] / np.pi, rel=1e-6)
    # identical to the matte material
    q = O.bsdf_probe(matte(0.5, 0.25, 0.125), CORRECTED, wo, wi)
    np.testing.assert_array_equal(q["f"], r["f"])


def test_microfacet_lobe_normal_incidence():
    a = 0.2
    m = plastic((0, 0, 0), (1, 1, 1), a, remap_roughness=False)
    z = np.array([0.0, 0.0, 1.0])
    r = O.bsdf_probe(m, CORRECTED, z, z)
    assert r["ncomp"] == 1
    # D(z) = 1/(pi a^2), G = 1, F = ((1.5 - 1)/(1.5 + 1))^2 = 0.04, cos_i = cos_o = 1
    np.testing.assert_allclose(r["f"], 0.04 / (np.pi * a * a * 4), rtol=2e-6)


@pytest.mark.parametrize("a", [0.1, 0.35, 0.8])
def test_microfacet_lobe_oblique_closed_form(a):
    m = plastic((0, 0, 0), (0.9, 0.7, 0.5), a, remap_roughness=False)
    wo, wi = unit([0.4, 0.1, 0.7]), unit([-0.3, -0.2, 0.9])
    r = O.bsdf_probe(m, CORRECTED, wo, wi)
    wh = unit(wo + wi)
    g = 1.0 / (1 + tr_lambda(wo, a) + tr_lambda(wi, a))
    f = tr_d(wh, a) * g * fresnel_dielectric(wi @ wh, 1.5, 1.0) / (4 * wi[2] * wo[2])
    np.testing.assert_allclose(r["f"], np.array([0.9, 0.7, 0.5]) * f, rtol=2e-5)


def test_two_lobes_sum_f_and_average_pdf():
    kd, ks, a = (0.6, 0.5, 0.4), (0.3, 0.3, 0.3), 0.25
    both = plastic(kd, ks, a, remap_roughness=False)
    lam = plastic(kd, (0, 0, 0), a, remap_roughness=False)
    mic = plastic((0, 0, 0), ks, a, remap_roughness=False)
    rng = np.random.default_rng(3)
    for _ in range(50):
        wo = unit(rng.standard_normal(3) * [1, 1, 0.2] + [0, 0, 0.6])
        wi = unit(rng.standard_normal(3) * [1, 1, 0.2] + [0, 0, 0.6])
        rb, rl, rm = (O.bsdf_probe(x, CORRECTED, wo, wi) for x in (both, lam, mic))
        assert rb["ncomp"] == 2
        np.testing.assert_allclose(rb["f"], rl["f"] + rm["f"], rtol=1e-6, atol=1e-9)
        assert rb["pdf"] == pytest.approx((rl["pdf"] + rm["pdf"]) / 2, rel=1e-6, abs=1e-9)


@pytest.mark.parametrize("mode", [0, 1])
def test_sample_agrees_with_f_and_pdf(mode):
    m = plastic((0.6, 0.5, 0.4), (0.3, 0.3, 0.3), 0.3)
    rng = np.random.default_rng(5)
    picked = {0: 0, 1: 0}
    for _ in range(300):
        wo = unit(rng.standard_normal(3) * [1, 1, 0.3] + [0, 0, 0.7])
        u0, u1 = rng.random(2)
        s = O.bsdf_probe(m, mode, wo, wo, u0, u1)
        if s["spdf"] == 0.0:
            continue
        picked[0 if u0 < 0.5 else 1] += 1
        assert bool(s["stype"] & 4) == (u0 < 0.5)  # BSDF_DIFFUSE for the first component
        e = O.bsdf_probe(m, mode, wo, s["swi"])
        np.testing.assert_allclose(s["sf"], e["f"], rtol=2e-5, atol=1e-7)
        assert s["spdf"] == pytest.approx(e["pdf"], rel=2e-5, abs=1e-7)
    assert picked[0] > 100 and picked[1] > 60  # glossy picks reflected below the horizon are dropped


def test_roughness_remap():
    # microfacet.rs:45-53: alpha(1.0) = 1.62142 (ln 1 = 0); remapped roughness r equals alpha(r) unremapped
    z = np.array([0.0, 0.0, 1.0])
    a = O.bsdf_probe(plastic((0, 0, 0), (1, 1, 1), 1.0, True), CORRECTED, z, z)
    b = O.bsdf_probe(plastic((0, 0, 0), (1, 1, 1), 1.62142, False), CORRECTED, z, z)
    np.testing.assert_allclose(a["f"], b["f"], rtol=1e-6)
    r = 0.2
    x = np.log(np.float32(r))
    alpha = 1.62142 + 0.819955 * x + 0.1734 * x * x + 0.0171201 * x ** 3 + 0.000640711 * x ** 4
    c = O.bsdf_probe(plastic((0, 0, 0), (1, 1, 1), r, True), CORRECTED, z, z)
    np.testing.assert_allclose(c["f"], 0.04 / (np.pi * alpha * alpha * 4), rtol=1e-5)


def test_black_plastic_has_no_lobes():
    r = O.bsdf_probe(plastic((0, 0, 0), (0, 0, 0), 0.3), CORRECTED, [0, 0, 1], [0, 0, 1])
    assert r["ncomp"] == 0 and r["pdf"] == 0.0 and not r["f"].any()


def test_fresnel_from_behind_the_normal():
    """base/bxdf.rs:164-187 swaps the indices for cos_i <= 0 but keeps the sign of cos_i: with wo and
    wi below the surface the reference mode evaluates that signed form (a pole near cos_i = -0.5547),
    the corrected mode the Fresnel equations for light from the other side (|cos_i|)."""
    a = 0.3
    m = plastic((0, 0, 0), (1, 1, 1), a, remap_roughness=False)
    wo, wi = unit([0.4, 0.1, 0.7]), unit([-0.3, -0.2, 0.9])
    front = O.bsdf_probe(m, CORRECTED, wo, wi)["f"]
    flip = np.array([1.0, 1.0, -1.0])
    back_c = O.bsdf_probe(m, CORRECTED, wo * flip, wi * flip)["f"]
    np.testing.assert_allclose(back_c * fresnel_dielectric(wi @ unit(wo + wi), 1.5, 1.0),
                               front * fresnel_dielectric(-(wi @ unit(wo + wi)), 1.5, 1.0), rtol=2e-5)
    back_r = O.bsdf_probe(m, 0, wo * flip, wi * flip)["f"]
    c = -(wi @ unit(wo + wi))  # the signed cosine the reference passes
    ei, et = 1.0, 1.5
    st = ei / et * np.sqrt(1 - c * c)
    ct = np.sqrt(1 - st * st)
    par = (et * c - ei * ct) / (et * c + ei * ct)
    per = (ei * c - et * ct) / (ei * c + et * ct)
    signed = (par * par + per * per) / 2
    np.testing.assert_allclose(back_r / back_c, signed / fresnel_dielectric(c, 1.5, 1.0), rtol=2e-4)
    assert signed > 1.0  # not a reflectance: the reference's estimator gains energy there
