"""Area lights on spheres (Sphere::sample and pdf_from_it, sphere.rs:328-377)
on the GPU: the path, direct-lighting and Whitted integrators against the
oracle on the same sample streams in both semantics, the furnace closed form
inside an emitting sphere (plain and guided), and scene batches holding a
sphere light equal to per-scene renders bit for bit."""
import numpy as np
import pytest

from bonsai_amd import GpuBvh
from bonsai_amd.integrator import (MODE_CORRECTED, MODE_REF, BatchRenderer, Film, GpuPathIntegrator,
                                   GuidedPathIntegrator, rel_mse, resolve)
from bonsai_amd.scenes import Camera, SceneData, add_spheres, area_light, cornell_box, matte
from test_gpu_direct import KINDS, gpu_film, gpu_integrator
from test_gpu_direct import oracle_image as oracle_direct
from test_gpu_render import oracle_image

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [MODE_REF, MODE_CORRECTED])
@pytest.mark.parametrize("double_sided", [False, True])
def test_path_with_sphere_light_matches_oracle(ctx, mode, double_sided):
    s = cornell_box(double_sided_light=double_sided, width=64, height=64, sphere_light=True)
    bvh = GpuBvh(ctx, s)
    gi = GpuPathIntegrator(ctx, s.camera, spp=16, max_depth=5, mode=mode, seed=13)
    film = Film(ctx, 64, 64)
    gi.render_into(bvh, film)
    rgb, w = film.sums()
    img = resolve(rgb.astype(np.float64), w.astype(np.float64))
    ref, wref = oracle_image(s, 16, 5, mode, seed=13, bvh=bvh)
    assert np.array_equal(w, wref.astype(np.float32))
    err = rel_mse(img, ref)
    print(f"sphere light, mode {mode}, double-sided {double_sided}: relMSE {err:.2e}, mean {img.mean():.4f}")
    assert err < 1e-4, err
    # the sphere light adds light: the image differs from the scene without it
    plain = cornell_box(double_sided_light=double_sided, width=64, height=64)
    img2 = GpuPathIntegrator(ctx, plain.camera, spp=16, max_depth=5, mode=mode, seed=13).render(GpuBvh(ctx, plain))
    assert rel_mse(img2, ref) > 1e-3


@pytest.mark.parametrize("kind", KINDS, ids=[k[0] for k in KINDS])
@pytest.mark.parametrize("mode", [MODE_REF, MODE_CORRECTED])
def test_direct_with_sphere_light_matches_oracle(ctx, kind, mode):
    s = cornell_box(double_sided_light=True, width=64, height=64, sphere_light=True)
    bvh = GpuBvh(ctx, s)
    img, _, w, _ = gpu_film(ctx, gpu_integrator(ctx, kind, s.camera, 16, mode, 19), bvh, s.camera)
    ref, wref = oracle_direct(s, bvh, kind, 16, mode, 19)
    assert np.array_equal(w, wref.astype(np.float32))
    err = rel_mse(img, ref)
    print(f"{kind[0]} with a sphere light, mode {mode}: relMSE {err:.2e}")
    assert err < 1e-4, err


def emitting_sphere_furnace(res=32):
    s = SceneData()
    add_spheres(s, [(0.0, 0.0, 0.0)], [0.8], [0])
    s.materials = [matte(0.5, 0.5, 0.5)]
    s.lights = [area_light(0, (1.0, 1.0, 1.0), True)]
    s.camera = Camera(pos=(0.05, 0.02, 0.01), look=(1.0, 0.3, 0.2), up=(0.0, 1.0, 0.0), fov=60.0, width=res,
                      height=res)
    return s


def test_furnace_inside_an_emitting_sphere(ctx):
    """L = sum_{k<=d} 0.5^k inside a double-sided emitting sphere of albedo 0.5, corrected semantics,
    plain and guided (one-sample MIS with a trained and an untrained network)."""
    import sys
    from pathlib import Path
    sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
    import guide_mlp as G
    from bonsai_amd.guide import GuideMLP

    s = emitting_sphere_furnace()
    bvh = GpuBvh(ctx, s)
    for depth in (1, 5):
        exact = sum(0.5 ** k for k in range(depth + 1))
        img = GpuPathIntegrator(ctx, s.camera, spp=64, max_depth=depth, mode=MODE_CORRECTED).render(bvh)
        assert abs(img.mean() - exact) < 0.01 * exact, (depth, img.mean(), exact)
        guide = GuideMLP(ctx, G.init_params(3).astype(np.float32), 65536)
        gi = GuidedPathIntegrator(ctx, s.camera, spp=64, max_depth=depth, guide=guide, train=True, train_paths=4096)
        film = Film(ctx, 32, 32)
        for k in range(8):
            gi.render_into(bvh, film, spp=8, first_sample=8 * k)
        g = film.image()
        assert abs(g.mean() - exact) < 0.02 * exact, (depth, g.mean(), exact)


def test_batch_with_sphere_light_equals_per_scene(ctx):
    scenes = [cornell_box(double_sided_light=True, width=48, height=40, sphere_light=True),
              cornell_box(double_sided_light=False, width=32, height=32),
              cornell_box(double_sided_light=False, width=40, height=48, sphere_light=True)]
    bvhs = [GpuBvh(ctx, s) for s in scenes]
    items = [(b, s.camera, 50 + k) for k, (b, s) in enumerate(zip(bvhs, scenes))]
    for mode, schedule in ((MODE_REF, 0), (MODE_CORRECTED, 0), (MODE_REF, 1), (MODE_CORRECTED, 1)):
        br = BatchRenderer(ctx, items)
        br.render(8, 5, mode=mode, autotune=False, schedule=schedule)
        for k, (b, cam, seed) in enumerate(items):
            film = Film(ctx, cam.width, cam.height)
            GpuPathIntegrator(ctx, cam, spp=8, max_depth=5, mode=mode, seed=seed).render_into(b, film)
            rgb, w = film.sums()
            brgb, bw = br.sums(k)
            np.testing.assert_array_equal(brgb, rgb)
            np.testing.assert_array_equal(bw, w)
        br.close()
