"""Scene batches: every item's film from one batched wavefront loop (hipGraph
replay, autotuned or fixed launch shapes, work-queue traversal) equals what
the single-scene integrator renders for that item alone, bit for bit."""
import numpy as np
import pytest

from bonsai_amd import GpuBvh
from bonsai_amd.integrator import (MODE_CORRECTED, MODE_REF, SCHEDULE_PERSISTENT, SCHEDULE_WAVEFRONT, BatchRenderer,
                                   Film, GpuPathIntegrator)
from bonsai_amd.scenes import Camera, batch_scene, cornell_box, path_tracer_scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def items(ctx):
    scenes = [batch_scene(i, scale=0.02) for i in range(4)]
    scenes.append(cornell_box(width=40, height=30))
    scenes.append(cornell_box(width=36, height=28, with_plastic=True))  # two-lobe shading for the whole batch
    scenes.append(path_tracer_scene(48, 27, small=True))
    bvhs = [GpuBvh(ctx, s) for s in scenes]
    out = [(bvh, s.camera, 100 + k) for k, (bvh, s) in enumerate(zip(bvhs, scenes))]
    # the same scene twice, from another viewpoint and with another seed
    out.append((bvhs[0], Camera(pos=(5.0, 6.0, -7.0), look=(0.0, 0.5, 0.0), up=(0.0, 1.0, 0.0), fov=50.0,
                                width=33, height=21), 999))
    yield out


def per_item(ctx, items, spp, depth, mode):
    res = []
    for bvh, cam, seed in items:
        gi = GpuPathIntegrator(ctx, cam, spp=spp, max_depth=depth, mode=mode, seed=seed)
        f = Film(ctx, cam.width, cam.height)
        gi.render_into(bvh, f)
        res.append(f.sums())
        f.close()
    return res


@pytest.mark.parametrize("mode", [MODE_REF, MODE_CORRECTED])
@pytest.mark.parametrize("graph", [True, False])
def test_batch_equals_per_scene(ctx, items, mode, graph):
    spp, depth = 4, 6
    ref = per_item(ctx, items, spp, depth, mode)
    br = BatchRenderer(ctx, items)
    # paths_per_pass small enough to force several replays of the captured pass
    total = sum(c.width * c.height for _, c, _ in items)
    st = br.render(spp, depth, mode=mode, use_graph=graph, autotune=True, paths_per_pass=2 * total)
    assert st.samples == total * spp and st.rays > 0
    for k in range(len(items)):
        rgb, w = br.sums(k)
        np.testing.assert_array_equal(w, ref[k][1])
        np.testing.assert_array_equal(rgb, ref[k][0])
    block, stack, bpc, variant = br.launch()
    assert block in (128, 256) and stack in (8, 16) and bpc in (8, 16, 32, 64) and variant in (0, 2, 4, 5)
    br.close()


@pytest.mark.parametrize("mode", [MODE_REF, MODE_CORRECTED])
@pytest.mark.parametrize("graph", [True, False])
def test_persistent_schedule_equals_per_scene(ctx, items, mode, graph):
    # one persistent path kernel per pass (lanes take path slots from an atomic counter): the same films
    # and ray counts as the wavefront schedule and as each item rendered alone
    spp, depth = 4, 6
    ref = per_item(ctx, items, spp, depth, mode)
    total = sum(c.width * c.height for _, c, _ in items)
    br = BatchRenderer(ctx, items)
    wf = br.render(spp, depth, mode=mode, use_graph=graph, autotune=False, paths_per_pass=2 * total,
                   schedule=SCHEDULE_WAVEFRONT)
    st = br.render(spp, depth, mode=mode, use_graph=graph, autotune=False, paths_per_pass=2 * total,
                   schedule=SCHEDULE_PERSISTENT)
    assert st.samples == total * spp and st.rays == wf.rays and st.shadow_rays == wf.shadow_rays
    for k in range(len(items)):
        rgb, w = br.sums(k)
        np.testing.assert_array_equal(w, ref[k][1])
        np.testing.assert_array_equal(rgb, ref[k][0])
    br.close()


def test_unknown_schedule_is_rejected(ctx, items):
    br = BatchRenderer(ctx, items[:2])
    with pytest.raises(RuntimeError, match="schedule"):
        br.render(1, 2, schedule=2)
    br.close()


@pytest.mark.parametrize("variant", [0, 2, 4, 5])
def test_fixed_shapes_and_work_queue(ctx, items, variant):
    b0, s0, c0, v0 = ctx.launch()
    ctx.set_launch(128, 8, 4, variant)
    try:
        ref = per_item(ctx, items, 2, 4, MODE_CORRECTED)
        br = BatchRenderer(ctx, items)
        br.render(2, 4, mode=MODE_CORRECTED, autotune=False)
        assert br.launch() == (128, 8, 4, variant)
        for k in range(len(items)):
            rgb, w = br.sums(k)
            np.testing.assert_array_equal(rgb, ref[k][0])
            np.testing.assert_array_equal(w, ref[k][1])
        br.close()
    finally:
        ctx.set_launch(b0, s0, c0, v0)


def test_measured_batch_render(ctx, items):
    """With the context's measurement on (bg_set_profiling), a batch render counts its traversal
    launches' tests and times them (passes run as direct launches instead of graph replays); the
    films are the same as without it."""
    br = BatchRenderer(ctx, items[:3])
    st0 = br.render(2, 4, autotune=False)
    ref = [br.sums(k) for k in range(3)]
    assert st0.nodes_visited == 0 and st0.traverse_ms == 0.0
    try:
        ctx.set_profiling(count=True)
        st1 = br.render(2, 4, autotune=False)
        ctx.set_profiling(traversal_time=True)
        st2 = br.render(2, 4, autotune=False)
    finally:
        ctx.set_profiling()
    assert st1.nodes_visited > st1.rays > 0 and st1.tris_tested > 0 and st1.rays == st0.rays
    assert st2.traverse_ms > 0.0 and st2.traverse_ms <= st2.kernel_ms and st2.rays == st0.rays
    for k in range(3):
        rgb, w = br.sums(k)
        np.testing.assert_array_equal(rgb, ref[k][0])
        np.testing.assert_array_equal(w, ref[k][1])
    br.close()


def test_batch_rerender_clears_the_film(ctx, items):
    br = BatchRenderer(ctx, items[:3])
    br.render(2, 3, autotune=False)
    a = [br.sums(k) for k in range(3)]
    br.render(2, 3, autotune=False)
    for k in range(3):
        rgb, w = br.sums(k)
        np.testing.assert_array_equal(rgb, a[k][0])
        np.testing.assert_array_equal(w, a[k][1])


def test_refill_traversal_with_per_ray_roots(ctx):
    """Triangle-only scenes batched with one block per CU and 64-thread blocks: the first bounces
    give every lane at least four rays, so the traversal runs variant 4's refill kernel with per-ray
    root records. Each item's film equals the single-scene render (grid-stride traversal)."""
    import dataclasses
    scenes = [batch_scene(i, scale=0.02) for i in range(3)]
    bvhs = [GpuBvh(ctx, s) for s in scenes]
    items = [(b, dataclasses.replace(s.camera, width=160, height=160), 200 + k)
             for k, (b, s) in enumerate(zip(bvhs, scenes))]
    ref = per_item(ctx, items, 2, 4, MODE_CORRECTED)
    saved = ctx.launch()
    ctx.set_launch(64, 8, 1, 4)
    try:
        br = BatchRenderer(ctx, items)
        total = sum(c.width * c.height for _, c, _ in items)
        br.render(2, 4, mode=MODE_CORRECTED, autotune=False, paths_per_pass=2 * total)
        assert br.launch() == (64, 8, 1, 4)
        for k in range(len(items)):
            rgb, w = br.sums(k)
            np.testing.assert_array_equal(w, ref[k][1])
            np.testing.assert_array_equal(rgb, ref[k][0])
        br.close()
    finally:
        ctx.set_launch(*saved)


@pytest.mark.parametrize("variant", [0, 2, 5])
def test_compacted_batch_equals_per_scene(ctx, variant):
    """A batch whose passes cover over 1M path slots and hold a scene with spheres (so the launches do
    not refill) compacts its traversal launches, per-ray root records included; each item's film
    equals its single-scene render (whose smaller launches do not compact). Variants 2 and 5 take the
    compacted rays from the atomic work queue, reading their count from the device (5 refilling
    finished lanes from the runs it takes)."""
    scenes = [path_tracer_scene(512, 256, small=True), cornell_box(width=512, height=256, with_plastic=True),
              batch_scene(1, scale=0.02)]
    bvhs = [GpuBvh(ctx, s) for s in scenes]
    items = [(b, s.camera, 300 + k) for k, (b, s) in enumerate(zip(bvhs, scenes))]
    ref = per_item(ctx, items, 4, 5, MODE_CORRECTED)
    saved, saved_c = ctx.launch(), ctx.compaction()
    ctx.set_launch(256, 8, 32, variant, compact=True)
    try:
        br = BatchRenderer(ctx, items)
        total = sum(c.width * c.height for _, c, _ in items)
        assert total * 4 >= 1 << 20
        br.render(4, 5, mode=MODE_CORRECTED, autotune=False, paths_per_pass=4 * total)
        for k in range(len(items)):
            rgb, w = br.sums(k)
            np.testing.assert_array_equal(w, ref[k][1])
            np.testing.assert_array_equal(rgb, ref[k][0])
        br.close()
    finally:
        ctx.set_launch(*saved, compact=saved_c)
