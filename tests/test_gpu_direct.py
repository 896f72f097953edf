"""DirectLightingIntegrator (direct.rs:72-113, both light strategies) and
WhittedIntegrator (whitted.rs:41-93) on the GPU against the oracle's
restatements on the same sample streams, in both semantics, with single-lobe
and plastic materials; the closed-form floor under a square light; and the
film's independence from how samples are split into passes."""
import numpy as np
import pytest

import cpu_oracle as O
from bonsai_amd import GpuBvh
from bonsai_amd.integrator import (MODE_CORRECTED, MODE_REF, UNIFORM_SAMPLE_ALL, UNIFORM_SAMPLE_ONE,
                                   DirectLightingIntegrator, Film, WhittedIntegrator, rel_mse, resolve)
from bonsai_amd.scenes import cornell_box, path_tracer_scene
from test_oracle_direct import KD, LE, SIDE, H, form_factor_centre, lit_floor

pytestmark = pytest.mark.gpu

KINDS = [("direct-all", 1, UNIFORM_SAMPLE_ALL), ("direct-one", 1, UNIFORM_SAMPLE_ONE), ("whitted", 2, 0)]


def gpu_integrator(ctx, kind, cam, spp, mode, seed):
    _, integ, strategy = kind
    if integ == 2:
        return WhittedIntegrator(ctx, cam, spp, mode=mode, seed=seed)
    return DirectLightingIntegrator(ctx, cam, spp, strategy=strategy, mode=mode, seed=seed)


def oracle_image(scene, bvh, kind, spp, mode, seed):
    _, integ, strategy = kind
    o = O.OracleScene(scene)
    nodes, slots = bvh.bvh()
    o.use_bvh(nodes, slots)
    o.set_shading(scene)
    rgb, w = o.render(scene, spp, 5, mode=mode, seed=seed, integrator=integ, strategy=strategy)
    return resolve(rgb, w), w


def gpu_film(ctx, integ, bvh, cam, **kw):
    film = Film(ctx, cam.width, cam.height)
    st = integ.render_into(bvh, film, **kw)
    rgb, w = film.sums()
    img = film.write_image()
    film.close()
    return img, rgb, w, st


@pytest.mark.parametrize("kind", KINDS, ids=[k[0] for k in KINDS])
@pytest.mark.parametrize("mode", [MODE_REF, MODE_CORRECTED])
@pytest.mark.parametrize("with_plastic", [False, True])
def test_cornell_matches_oracle_same_streams(ctx, kind, mode, with_plastic):
    s = cornell_box(double_sided_light=True, width=64, height=64, with_plastic=with_plastic)
    bvh = GpuBvh(ctx, s)
    img, _, w, st = gpu_film(ctx, gpu_integrator(ctx, kind, s.camera, 16, mode, 17), bvh, s.camera)
    ref, wref = oracle_image(s, bvh, kind, 16, mode, 17)
    assert np.array_equal(w, wref.astype(np.float32))
    err = rel_mse(img, ref)
    print(f"{kind[0]} mode {mode} plastic {with_plastic}: relMSE {err:.2e}, rays {st.rays}, shadow {st.shadow_rays}")
    assert err < 1e-4, err
    assert st.samples == 64 * 64 * 16 and st.rays > st.samples


@pytest.mark.parametrize("kind", KINDS, ids=[k[0] for k in KINDS])
def test_many_lights_scene_matches_oracle(ctx, kind):
    """The path-tracer scene kind (8 light triangles, spheres, matte and conductor materials)."""
    s = path_tracer_scene(64, 36, small=True)
    bvh = GpuBvh(ctx, s)
    for mode in (MODE_REF, MODE_CORRECTED):
        img, _, w, _ = gpu_film(ctx, gpu_integrator(ctx, kind, s.camera, 4, mode, 23), bvh, s.camera)
        ref, wref = oracle_image(s, bvh, kind, 4, mode, 23)
        assert np.array_equal(w, wref.astype(np.float32))
        finite = np.isfinite(ref) & np.isfinite(img)
        rel = np.abs(img - ref) / np.maximum(np.abs(ref), 1e-3)
        assert np.quantile(rel[finite], 0.99) < 1e-3, (kind[0], mode)


@pytest.mark.parametrize("kind", KINDS, ids=[k[0] for k in KINDS])
def test_floor_under_square_light_closed_form(ctx, kind):
    s = lit_floor()
    bvh = GpuBvh(ctx, s)
    img, _, _, _ = gpu_film(ctx, gpu_integrator(ctx, kind, s.camera, 16384, MODE_CORRECTED, 3), bvh, s.camera)
    np.testing.assert_allclose(img.mean(axis=(0, 1)), KD * LE * form_factor_centre(SIDE, H), rtol=0.01)


@pytest.mark.parametrize("kind", KINDS, ids=[k[0] for k in KINDS])
def test_passes_and_windows_do_not_change_the_film(ctx, kind):
    s = cornell_box(double_sided_light=True, width=48, height=40, with_plastic=True)
    bvh = GpuBvh(ctx, s)
    integ = gpu_integrator(ctx, kind, s.camera, 8, MODE_CORRECTED, 29)
    _, rgb, w, _ = gpu_film(ctx, integ, bvh, s.camera)
    # one sample per pass
    _, rgb1, w1, _ = gpu_film(ctx, integ, bvh, s.camera, paths_per_pass=1)
    np.testing.assert_array_equal(rgb1, rgb)
    np.testing.assert_array_equal(w1, w)
    # the samples split over two calls into one film
    film = Film(ctx, 48, 40)
    integ.render_into(bvh, film, spp=3, first_sample=0)
    integ.render_into(bvh, film, spp=5, first_sample=3)
    rgb2, w2 = film.sums()
    film.close()
    np.testing.assert_array_equal(rgb2, rgb)
    np.testing.assert_array_equal(w2, w)
    # a tile-aligned window gets exactly the full render's sums
    film = Film(ctx, 48, 40)
    integ.render_into(bvh, film, window=(16, 16, 48, 32))
    rgb3, w3 = film.sums()
    film.close()
    np.testing.assert_array_equal(rgb3[16:32, 16:48], rgb[16:32, 16:48])
    np.testing.assert_array_equal(w3[16:32, 16:48], w[16:32, 16:48])


def test_ray_counts(ctx):
    s = cornell_box(double_sided_light=True, width=32, height=32)
    bvh = GpuBvh(ctx, s)
    n = 32 * 32 * 2
    for kind in KINDS:
        _, _, _, st = gpu_film(ctx, gpu_integrator(ctx, kind, s.camera, 2, MODE_CORRECTED, 31), bvh, s.camera)
        per_sample_max = 1 + (2 if kind[2] == UNIFORM_SAMPLE_ONE and kind[1] == 1 else 2 * 2)
        assert n < st.rays <= n * per_sample_max, kind[0]
        assert 0 < st.shadow_rays < st.rays
