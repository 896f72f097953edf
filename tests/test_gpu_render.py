"""GPU path integrator against the oracle's restatement of PathIntegrator
(integrators/path.rs, base/integrator.rs) on the same sample streams."""
import numpy as np
import pytest

import cpu_oracle as O
from bonsai_amd import GpuBvh
from bonsai_amd.integrator import MODE_CORRECTED, MODE_REF, Film, GpuPathIntegrator, rel_mse, resolve
from bonsai_amd.scenes import cornell_box, path_tracer_scene
from test_oracle_integrator import furnace

pytestmark = pytest.mark.gpu


def oracle_image(scene, spp, depth, mode, seed=1, bvh=None):
    o = O.OracleScene(scene)
    if bvh is not None:
        nodes, slots = bvh.bvh()
        o.use_bvh(nodes, slots)
    else:
        o.build()
    o.set_shading(scene)
    rgb, w = o.render(scene, spp, depth, mode=mode, seed=seed)
    return resolve(rgb, w), w


@pytest.mark.parametrize("mode", [MODE_REF, MODE_CORRECTED])
@pytest.mark.parametrize("double_sided", [False, True])
def test_cornell_matches_oracle_same_streams(ctx, mode, double_sided):
    s = cornell_box(double_sided_light=double_sided, width=96, height=96)
    bvh = GpuBvh(ctx, s)
    gi = GpuPathIntegrator(ctx, s.camera, spp=16, max_depth=5, mode=mode, seed=7)
    film = Film(ctx, 96, 96)
    gi.render_into(bvh, film)
    rgb, w = film.sums()
    img = resolve(rgb.astype(np.float64), w.astype(np.float64))
    ref, wref = oracle_image(s, 16, 5, mode, seed=7, bvh=bvh)
    assert np.array_equal(w, wref.astype(np.float32))
    err = rel_mse(img, ref)
    # same sample streams: only paths whose directions differ by an ulp of a device sin/cos/acos diverge
    assert err < 1e-4, err


@pytest.mark.parametrize("mode", [MODE_REF, MODE_CORRECTED])
def test_plastic_matches_oracle_same_streams(ctx, mode):
    """Two-lobe plastic (materials/plastic.rs:24-52) on a sphere and the floor, beside single-lobe
    matte and conductor materials: the GPU's two-lobe shading kernel against the oracle. The light is
    double-sided so that the reference semantics (emission toward +wi) light the surfaces as well."""
    s = cornell_box(double_sided_light=True, width=96, height=96, with_plastic=True)
    bvh = GpuBvh(ctx, s)
    gi = GpuPathIntegrator(ctx, s.camera, spp=16, max_depth=5, mode=mode, seed=21)
    film = Film(ctx, 96, 96)
    gi.render_into(bvh, film)
    rgb, w = film.sums()
    img = resolve(rgb.astype(np.float64), w.astype(np.float64))
    ref, wref = oracle_image(s, 16, 5, mode, seed=21, bvh=bvh)
    assert np.array_equal(w, wref.astype(np.float32))
    err = rel_mse(img, ref)
    assert err < 1e-4, err
    # the plastic changes the image: the same scene without it differs well beyond that
    plain = cornell_box(double_sided_light=True, width=96, height=96)
    img2 = GpuPathIntegrator(ctx, plain.camera, spp=16, max_depth=5, mode=mode, seed=21).render(GpuBvh(ctx, plain))
    assert rel_mse(img2, ref) > 1e-2


@pytest.mark.parametrize("mode", [MODE_REF, MODE_CORRECTED])
def test_path_tracer_scene_matches_oracle_same_streams(ctx, mode):
    """The plain path-tracer configuration's scene kind (displaced icospheres, 256 spheres, matte and
    conductor materials, double-sided quad lights), three subdivision levels lower, at 80 x 45."""
    s = path_tracer_scene(80, 45, small=True)
    bvh = GpuBvh(ctx, s)
    gi = GpuPathIntegrator(ctx, s.camera, spp=4, max_depth=8, mode=mode, seed=9)
    film = Film(ctx, 80, 45)
    gi.render_into(bvh, film)
    rgb, w = film.sums()
    img = resolve(rgb.astype(np.float64), w.astype(np.float64))
    ref, wref = oracle_image(s, 4, 8, mode, seed=9, bvh=bvh)
    assert np.array_equal(w, wref.astype(np.float32))
    finite = np.isfinite(ref) & np.isfinite(img)
    print(f"path tracer scene mode {mode}: image mean GPU {img[finite].mean():.4g} oracle {ref[finite].mean():.4g}, "
          f"max GPU {img.max():.4g} oracle {ref.max():.4g}")
    # same sample streams: the per-pixel values agree except where a path's direction differs by
    # an ulp of a device transcendental; compare the bulk of pixels
    rel = np.abs(img - ref) / np.maximum(np.abs(ref), 1e-3)
    assert np.quantile(rel[finite], 0.99) < 1e-3


def test_corrected_furnace_on_gpu(ctx):
    s = furnace(res=32)
    bvh = GpuBvh(ctx, s)
    for depth in (1, 5):
        gi = GpuPathIntegrator(ctx, s.camera, spp=64, max_depth=depth, mode=MODE_CORRECTED)
        img = gi.render(bvh)
        exact = sum(0.5 ** k for k in range(depth + 1))
        assert abs(img.mean() - exact) < 0.02 * exact, (depth, img.mean(), exact)


def test_passes_and_windows_do_not_change_the_film(ctx):
    s = cornell_box(width=64, height=48)
    bvh = GpuBvh(ctx, s)
    gi = GpuPathIntegrator(ctx, s.camera, spp=8, max_depth=5, mode=MODE_CORRECTED, seed=3)
    a = Film(ctx, 64, 48)
    gi.render_into(bvh, a)
    b = Film(ctx, 64, 48)
    gi.render_into(bvh, b, spp=3, first_sample=0, paths_per_pass=64 * 48)
    gi.render_into(bvh, b, spp=5, first_sample=3, paths_per_pass=2 * 64 * 48)
    ra, wa = a.sums()
    rb, wb = b.sums()
    assert np.array_equal(ra, rb) and np.array_equal(wa, wb)
    # a tile-aligned window gets exactly the full render's sums
    c = Film(ctx, 64, 48)
    gi.render_into(bvh, c, window=(16, 16, 64, 32))
    rc, wc = c.sums()
    assert np.array_equal(rc[16:32, 16:64], ra[16:32, 16:64]) and np.array_equal(wc[16:32, 16:64], wa[16:32, 16:64])


def test_launch_shape_does_not_change_the_image(ctx):
    s = cornell_box(width=48, height=48)
    bvh = GpuBvh(ctx, s)
    gi = GpuPathIntegrator(ctx, s.camera, spp=4, max_depth=5, mode=MODE_REF, seed=5)
    saved = ctx.launch()
    imgs = []
    try:
        for block, stack in [(64, 8), (128, 16), (256, 32)]:
            ctx.set_launch(block, stack, 4)
            imgs.append(gi.render(bvh))
    finally:
        ctx.set_launch(*saved)
    assert all(np.array_equal(imgs[0], x) for x in imgs[1:])


@pytest.mark.parametrize("mode", [MODE_REF, MODE_CORRECTED])
def test_converges_to_the_oracle_at_4096_spp(ctx, mode):
    """The Cornell box (10 wall triangles, a matte and a conductor sphere, a 0.25 x 0.25 light of
    radiance 17, depth 5) at 48 x 48: the GPU render and the oracle's, each at 4096 samples per
    pixel with independent sample streams (different seeds), agree within relMSE 1e-3 — the two
    converge to the same image, whatever the streams. (Reference semantics with a double-sided
    light: with the one-sided one only the emitter is lit there.)"""
    s = cornell_box(double_sided_light=mode == MODE_REF, width=48, height=48)
    bvh = GpuBvh(ctx, s)
    img = GpuPathIntegrator(ctx, s.camera, spp=4096, max_depth=5, mode=mode, seed=101).render(bvh)
    ref, _ = oracle_image(s, 4096, 5, mode, seed=202, bvh=bvh)
    err = rel_mse(img, ref)
    print(f"mode {mode}: relMSE GPU 4096 spp vs oracle 4096 spp (independent seeds) {err:.2e}, "
          f"means {img.mean():.4f} / {ref.mean():.4f}")
    assert err < 1e-3
    assert abs(img.mean() / ref.mean() - 1.0) < 0.01


@pytest.mark.parametrize("size", [(512, 256), (1024, 600)])
@pytest.mark.parametrize("mode", [MODE_REF, MODE_CORRECTED])
def test_compaction_does_not_change_the_film(ctx, mode, size):
    """Tracing only the rays that can hit (compacted on the device, for launches of 1M slots or
    more that do not refill) instead of every path slot gives bit-identical film sums: on the
    path-tracer scene kind (spheres and triangles, 8 light
    triangles, roulette ending paths) and, in the corrected semantics, for the guided estimator
    with a frozen network. 1024 x 600 x 8 spp = 4,915,200 slots gives 19,200 block counts, so
    the one-block scan runs both its eight-wide and its single-count loops. Variant 5 (refill from
    an atomic work queue, persistent, on the compacted index list or on every slot) gives the same
    sums too."""
    import sys
    from pathlib import Path
    from bonsai_amd.guide import GuideMLP
    from bonsai_amd.integrator import GuidedPathIntegrator
    sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
    import guide_mlp as G

    w, h = size
    s = path_tracer_scene(w, h, small=True)  # 512 x 256 x 8 spp = 2^20 path slots: compacted launches
    bvh = GpuBvh(ctx, s)
    guide = GuideMLP(ctx, G.init_params(5).astype(np.float32), 65536)
    saved, saved_c = ctx.launch(), ctx.compaction()
    out = {}
    configs = [(False, saved[3]), (True, saved[3]), (True, 5), (False, 5)]
    try:
        for compact, variant in configs:
            ctx.set_launch(saved[0], saved[1], saved[2], variant, compact=compact)
            integs = [GpuPathIntegrator(ctx, s.camera, spp=8, max_depth=8, mode=mode, seed=19)]
            if mode == MODE_CORRECTED:
                integs.append(GuidedPathIntegrator(ctx, s.camera, spp=8, max_depth=8, guide=guide, seed=23,
                                                   train=False, rr_threshold=0.0))
            sums = []
            for gi in integs:
                film = Film(ctx, w, h)
                st = gi.render_into(bvh, film)
                sums.append(film.sums() + (st.rays,))
                film.close()
            out[(compact, variant)] = sums
    finally:
        ctx.set_launch(*saved, compact=saved_c)
    for key in configs[1:]:
        for (rgb0, w0, r0), (rgb1, w1, r1) in zip(out[configs[0]], out[key]):
            np.testing.assert_array_equal(rgb1, rgb0)
            np.testing.assert_array_equal(w1, w0)
            assert r1 == r0


@pytest.mark.parametrize("value", [float("nan"), -2.0, float("inf")])
@pytest.mark.parametrize("both", [False, True])
def test_sample_guards_match_the_oracle(ctx, value, both):
    """integrator.rs:92-117 on the GPU: a light emitting NaN, negative or infinite radiance. Samples of
    NaN radiance, luminance below -1e-5 or infinite luminance go to the film as black with their
    filter weight; the film's sums equal the oracle's and hold no NaN or infinity."""
    from test_oracle_integrator import poisoned_cornell
    s = poisoned_cornell(value, both=both, res=48)
    bvh = GpuBvh(ctx, s)
    for mode in (MODE_REF, MODE_CORRECTED):
        film = Film(ctx, 48, 48)
        GpuPathIntegrator(ctx, s.camera, spp=8, max_depth=5, mode=mode, seed=3).render_into(bvh, film)
        rgb, w = film.sums()
        film.close()
        ref, wref = oracle_image(s, 8, 5, mode, seed=3, bvh=bvh)
        assert np.all(np.isfinite(rgb)) and np.all(w == 8.0)
        assert np.array_equal(w, wref.astype(np.float32))
        img = resolve(rgb.astype(np.float64), w.astype(np.float64))
        assert rel_mse(img, ref) < 1e-4, (value, both, mode)
        if both and value != -2.0:
            assert np.all(rgb == 0.0)


def test_profiling_counts_and_times_traversal_without_changing_the_image(ctx):
    """bg_set_profiling: the render's traversal launches count their box and primitive tests and are
    timed by HIP event pairs; the film is the same bit for bit."""
    s = cornell_box(width=64, height=64)
    bvh = GpuBvh(ctx, s)
    gi = GpuPathIntegrator(ctx, s.camera, spp=8, max_depth=5, mode=MODE_CORRECTED, seed=4)
    films = []
    stats = []
    for flags in ((False, False), (True, True), (False, True)):
        ctx.set_profiling(*flags)
        try:
            film = Film(ctx, 64, 64)
            stats.append(gi.render_into(bvh, film))
            films.append(film.sums())
            film.close()
        finally:
            ctx.set_profiling()
    for rgb, w in films[1:]:
        np.testing.assert_array_equal(rgb, films[0][0])
        np.testing.assert_array_equal(w, films[0][1])
    plain, both, timed = stats
    assert plain.nodes_visited == 0 and plain.traverse_ms == 0.0
    assert both.nodes_visited > both.rays > 0 and both.tris_tested > 0 and both.spheres_tested > 0
    assert both.rays == plain.rays
    assert 0.0 < timed.traverse_ms <= timed.kernel_ms and timed.nodes_visited == 0
