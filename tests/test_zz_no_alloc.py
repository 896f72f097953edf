"""bg_intersect_device_noalloc (static scratch, no device allocation) against
bg_intersect_device on the same device-resident rays: identical hits and box /
primitive test counts, on two calls in a row, the second with fewer rays."""
import ctypes as C

import numpy as np
import pytest

from bonsai_amd import GpuBvh
from bonsai_amd import _native as N
from bonsai_amd.scenes import SceneData, intersection_scene, incoherent_rays, path_tracer_scene
from hipmem import DeviceArray

pytestmark = pytest.mark.gpu


def _run(ctx, bvh, fn, d_rays, n, any_hit):
    d_hits = DeviceArray(max(n, 1) * 16)
    st = N.bg_stats()
    flags = (N.BG_ANY if any_hit else N.BG_CLOSEST) | N.BG_COUNT
    N.check(getattr(N.lib(), fn)(ctx.h, bvh.h, d_rays.ptr, n, d_hits.ptr, flags, C.byref(st)), ctx.h)
    hits = d_hits.to_numpy(np.uint32, (n, 4))
    d_hits.free()
    return hits, (st.rays, st.nodes_visited, st.tris_tested, st.spheres_tested)


@pytest.mark.parametrize("kind", ["triangles", "spheres"])
def test_noalloc_equals_device_entry_point(ctx, kind):
    if kind == "triangles":
        s = intersection_scene(level=5)
    else:
        s = path_tracer_scene(64, 36, small=True)
    bvh = GpuBvh(ctx, s)
    lo, hi = bvh.bounds()
    rays = incoherent_rays(400_000, tuple(lo), tuple(hi), 11)
    d_rays = DeviceArray.from_numpy(rays)
    print(f"noalloc {kind}: {bvh.info.n_prims} primitives, BVH depth {bvh.info.max_depth}", flush=True)
    # two calls in a row on the shared scratch, the second smaller (and the other query kind)
    for n, any_hit in ((400_000, False), (70_001, True)):
        want, wst = _run(ctx, bvh, "bg_intersect_device", d_rays, n, any_hit)
        got, gst = _run(ctx, bvh, "bg_intersect_device_noalloc", d_rays, n, any_hit)
        np.testing.assert_array_equal(got, want)
        assert gst == wst, (gst, wst)
        assert (want[:, 0] != N.BG_MISS).any()
    d_rays.free()


def test_noalloc_empty_scene_and_zero_rays(ctx):
    empty = SceneData()
    bvh = GpuBvh(ctx, empty)
    rays = incoherent_rays(1000, (-1, -1, -1), (1, 1, 1), 3)
    d_rays = DeviceArray.from_numpy(rays)
    got, st = _run(ctx, bvh, "bg_intersect_device_noalloc", d_rays, 1000, False)
    assert (got[:, 0] == N.BG_MISS).all() and st[0] == 1000
    got, st = _run(ctx, bvh, "bg_intersect_device_noalloc", d_rays, 0, False)
    assert got.shape == (0, 4)
    d_rays.free()
