"""GPU traversal parity: the HIP closest-hit / any-hit kernels through the C ABI
against the oracle's reference-order traversal on the same flattened BVH.
Primitive ids, t and triangle barycentrics must be bit-exact; sphere (u, v)
come from atan2/acos and are compared at 1e-6 relative."""
import numpy as np
import pytest

import cpu_oracle as O
from bonsai_amd import GpuBvh
from bonsai_amd.scenes import (SceneData, add_spheres, cornell_box, icosphere, incoherent_rays,
                               intersection_camera, intersection_scene)

pytestmark = pytest.mark.gpu
MISS = 0xFFFFFFFF


def oracle_for(bvh: GpuBvh, scene: SceneData) -> O.OracleScene:
    o = O.OracleScene(scene)
    nodes, slots = bvh.bvh()
    o.use_bvh(nodes, slots)
    return o


def same_bits(a, b):
    """Bitwise equal f32 words, except that any NaN equals any NaN (x86 and gfx950 produce different
    default NaN payloads for the same invalid operation)."""
    fa, fb = a.view(np.float32), b.view(np.float32)
    return np.all((a == b) | (np.isnan(fa) & np.isnan(fb)))


def assert_hits_equal(gpu_hits, ref, n_tris, any_hit=False):
    g = O.hits_as_words(gpu_hits)
    assert np.array_equal(g[:, 0], ref[:, 0]), f"prim ids differ on {np.sum(g[:, 0] != ref[:, 0])} rays"
    hit = g[:, 0] != MISS
    assert same_bits(g[hit, 1], ref[hit, 1]), "t differs"
    if any_hit:
        return
    tri = hit & (g[:, 0] < n_tris)
    assert same_bits(g[tri, 2:4], ref[tri, 2:4]), "triangle barycentrics differ"
    sph = hit & (g[:, 0] >= n_tris)
    if sph.any():
        gf = g[sph, 2:4].view(np.float32)
        rf = ref[sph, 2:4].view(np.float32)
        np.testing.assert_allclose(gf, rf, rtol=1e-6, atol=1e-6)


def run_both(ctx, scene, rays):
    bvh = GpuBvh(ctx, scene)
    o = oracle_for(bvh, scene)
    for any_hit in (False, True):
        g = bvh.intersect_test(rays) if any_hit else bvh.intersect(rays)
        r = o.intersect(rays, any_hit=any_hit)
        assert_hits_equal(g, r, scene.n_tris, any_hit)
    return bvh, o


def box_rays(n, lo, hi, seed, aim_lo=None, aim_hi=None):
    rng = np.random.default_rng(seed)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = rng.uniform(lo, hi, (n, 3))
    r[:, 3] = np.inf
    if aim_lo is None:
        r[:, 4:7] = rng.standard_normal((n, 3))
    else:
        r[:, 4:7] = rng.uniform(aim_lo, aim_hi, (n, 3)) - r[:, 0:3]
    return r


def test_cornell_geometry(ctx):
    s = cornell_box()
    rays = box_rays(200_000, 0.01, 0.99, 1)
    bvh, o = run_both(ctx, s, rays)
    g = O.hits_as_words(bvh.intersect(rays))
    assert (g[:, 0] != MISS).mean() > 0.8  # five walls: only rays leaving through the open front miss


def test_icosphere_incoherent(ctx):
    s = SceneData(tri_positions=icosphere(6, (0.0, 0.0, 0.0), 1.0, 0.15, 4.0, 7))
    lo, hi = s.bounds()
    rays = incoherent_rays(300_000, lo, hi, 5)
    run_both(ctx, s, rays)


def test_icosphere_coherent_camera(ctx):
    s = SceneData(tri_positions=icosphere(6, (0.0, 0.0, 0.0), 1.0, 0.15, 4.0, 7))
    cam = intersection_camera(384, 384)
    rays = ctx.camera_rays(cam)
    m, mi = cam.matrices()
    ref = O.camera_rays(m, mi, cam.fov, cam.width, cam.height)
    assert np.array_equal(rays.view(np.uint32), ref.view(np.uint32)), "camera rays differ"
    run_both(ctx, s, rays)


def test_spheres_and_triangles_mixed(ctx):
    rng = np.random.default_rng(8)
    n = 3000
    c = rng.uniform(-2, 2, (n, 1, 3))
    tris = (c + rng.uniform(-0.2, 0.2, (n, 3, 3))).astype(np.float32).reshape(n, 9)
    s = SceneData(tri_positions=tris)
    add_spheres(s, rng.uniform(-2, 2, (300, 3)), rng.uniform(0.01, 0.5, 300))
    rays = box_rays(200_000, -3, 3, 9, -2, 2)
    bvh, _ = run_both(ctx, s, rays)
    g = O.hits_as_words(bvh.intersect(rays))
    assert (g[:, 0] >= s.n_tris).sum() > 1000 and ((g[:, 0] < s.n_tris) & (g[:, 0] != MISS)).sum() > 1000


def test_edge_case_rays(ctx):
    s = cornell_box()
    bvh = GpuBvh(ctx, s)
    o = oracle_for(bvh, s)
    rays = [
        O.ray((0.5, 0.5, 0.5), (1, 0, 0)), O.ray((0.5, 0.5, 0.5), (0, -1, 0)),   # axis-aligned (0 components)
        O.ray((0.5, 0.5, 0.5), (0, 0, 1)), O.ray((0.5, 0.5, 0.5), (-0.0, 0, -1)),
        O.ray((0.5, 0.5, 0.5), (0, 0, 0)),                                        # zero direction
        O.ray((0.5, 0.5, 0.5), (np.nan, 0, 1)),                                   # NaN direction
        O.ray((0.5, 0.5, 0.5), (1, 1, 1), t_max=0.0),                             # empty interval
        O.ray((0.5, 0.5, 0.5), (1, 1, 1), t_max=-1.0),
        O.ray((0.5, 0.5, 0.5), (1, 1, 1), t_max=0.1),                             # stops short of walls
        O.ray((0.0, 0.5, 0.5), (1, 0, 0)),                                        # origin on a wall plane
        O.ray((0.5, 0.0, 0.5), (0, 1, 0)),
        O.ray((0.5, 0.5, 0.5), (1, 0, 1)),                                        # towards an edge of two walls
        O.ray((0.5, 0.5, 0.5), (1, 1, 1)),                                        # towards a corner
        O.ray((0.5, 0.5, -5), (0, 0, 1)),                                         # from outside through the open front
        O.ray((2, 2, 2), (1, 1, 1)),                                              # outside, away
        O.ray((0.5, 0.5, 0.5), (np.inf, 0, 0)),
    ]
    rays = np.array(rays, np.float32)
    for any_hit in (False, True):
        g = bvh.intersect_test(rays) if any_hit else bvh.intersect(rays)
        assert_hits_equal(g, o.intersect(rays, any_hit=any_hit), s.n_tris, any_hit)


def test_empty_scene_and_empty_batch(ctx):
    bvh = GpuBvh(ctx, SceneData())
    rays = box_rays(100, -1, 1, 3)
    assert np.all(O.hits_as_words(bvh.intersect(rays))[:, 0] == MISS)
    assert np.all(O.hits_as_words(bvh.intersect_test(rays))[:, 0] == MISS)
    s = cornell_box()
    b2 = GpuBvh(ctx, s)
    assert b2.intersect(np.zeros((0, 8), np.float32)).shape[0] == 0


def deep_scene(n=120):
    """Planar triangles at x = 2^k: the centroid buckets peel a few off per level, giving a 30-level tree,
    deeper than the 8- and 16-entry LDS stacks, so the spill path runs."""
    tris = [[2.0 ** k, 0, 0, 2.0 ** k, 1, 0, 2.0 ** k, 0, 1] for k in range(n)]
    return SceneData(tri_positions=np.asarray(tris, np.float32))


def deep_rays(n, seed):
    """Rays along +x and -x through the triangles' (y, z) span, reaching the deepest leaves."""
    rng = np.random.default_rng(seed)
    r = np.zeros((n, 8), np.float32)
    r[:, 0] = np.where(rng.random(n) < 0.5, -1.0, 2.0 ** 121)
    r[:, 1:3] = rng.uniform(0.02, 0.45, (n, 2))
    r[:, 3] = np.inf
    r[:, 4] = np.where(r[:, 0] < 0, 1.0, -1.0)
    r[:, 5:7] = rng.uniform(-1e-3, 1e-3, (n, 2))
    return r


@pytest.mark.parametrize("block,stack,variant", [(64, 8, 0), (128, 16, 0), (256, 32, 0), (128, 8, 0), (64, 32, 0),
                                                  (64, 8, 1), (128, 16, 1), (256, 32, 1), (64, 8, 3), (256, 16, 3),
                                                  (128, 16, 2), (64, 8, 4), (256, 8, 4), (128, 32, 4),
                                                  (256, 8, 5), (64, 16, 5)])
def test_launch_shapes_do_not_change_results(ctx, block, stack, variant):
    saved = ctx.launch()
    try:
        ctx.set_launch(block, stack, 4, variant)
        s = deep_scene()
        bvh = GpuBvh(ctx, s)
        rays = np.concatenate([deep_rays(20_000, 17), box_rays(20_000, -0.5, 1.5, 17, 0, 1)])
        o = oracle_for(bvh, s)
        for any_hit in (False, True):
            g = bvh.intersect_test(rays) if any_hit else bvh.intersect(rays)
            assert_hits_equal(g, o.intersect(rays, any_hit=any_hit), s.n_tris, any_hit)
        s2 = SceneData(tri_positions=icosphere(4, (0.0, 0.0, 0.0), 1.0, 0.15, 4.0, 3))
        lo, hi = s2.bounds()
        run_both(ctx, s2, incoherent_rays(50_000, lo, hi, 2))
    finally:
        ctx.set_launch(*saved)


@pytest.mark.parametrize("block,stack", [(64, 8), (256, 8), (128, 16)])
def test_refill_variant_with_many_rays_per_lane(ctx, block, stack):
    """Variant 4 refills finished lanes from each wavefront's ray sequence once every lane has at
    least four rays (triangle scenes); with one block per CU every lane traces several rays (and
    the ray count is not a multiple of a chunk), so refills run mid-wavefront. Hits and visit counts
    equal the oracle's. (Sized for up to 256 CUs: 300,003 rays give each of 1,024 wavefronts at
    least 256.)"""
    saved = ctx.launch()
    ctx.set_launch(block, stack, 1, 4)
    try:
        s = deep_scene()
        rays = np.concatenate([deep_rays(150_001, 5), box_rays(150_002, -0.5, 1.5, 6, 0, 1)])
        rng = np.random.default_rng(3)
        rays = rays[rng.permutation(len(rays))]
        bvh = GpuBvh(ctx, s)
        o = oracle_for(bvh, s)
        for any_hit in (False, True):
            g = bvh.intersect_test(rays) if any_hit else bvh.intersect(rays)
            assert_hits_equal(g, o.intersect(rays, any_hit=any_hit), s.n_tris, any_hit)
            st = bvh.count(rays, any_hit)
            _, cnt = o.intersect(rays, any_hit=any_hit, counters=True)
            assert (st.nodes_visited, st.tris_tested) == tuple(int(x) for x in cnt.astype(np.uint64).sum(0)[:2])
        s2 = SceneData(tri_positions=icosphere(4, (0.0, 0.0, 0.0), 1.0, 0.15, 4.0, 3))
        lo, hi = s2.bounds()
        run_both(ctx, s2, incoherent_rays(300_003, lo, hi, 2))
        # a scene with spheres runs variant 4 as 0: the same results
        add_spheres(s2, rng.uniform(-1, 1, (50, 3)), rng.uniform(0.01, 0.2, 50))
        run_both(ctx, s2, incoherent_rays(100_000, lo, hi, 4))
    finally:
        ctx.set_launch(*saved)


@pytest.mark.parametrize("block,stack", [(64, 8), (256, 8), (128, 16)])
def test_queue_refill_variant(ctx, block, stack):
    """Variant 5: persistent wavefronts take runs of 256 rays from an atomic counter and refill
    finished lanes from them. With one block per CU every wavefront takes many runs, the ray count
    is not a multiple of a run, and the scene has spheres in one case (the refill kernel with the
    sphere test): hits equal the oracle's, and the counter is zeroed by each launch (a second
    query of the same rays gives the same hits)."""
    saved = ctx.launch()
    ctx.set_launch(block, stack, 1, 5)
    try:
        s = deep_scene()
        rays = np.concatenate([deep_rays(150_001, 7), box_rays(150_002, -0.5, 1.5, 8, 0, 1)])
        rays = rays[np.random.default_rng(9).permutation(len(rays))]
        bvh = GpuBvh(ctx, s)
        o = oracle_for(bvh, s)
        for any_hit in (False, True):
            ref = o.intersect(rays, any_hit=any_hit)
            for _ in range(2):
                g = bvh.intersect_test(rays) if any_hit else bvh.intersect(rays)
                assert_hits_equal(g, ref, s.n_tris, any_hit)
        s2 = SceneData(tri_positions=icosphere(4, (0.0, 0.0, 0.0), 1.0, 0.15, 4.0, 3))
        lo, hi = s2.bounds()
        rng = np.random.default_rng(11)
        add_spheres(s2, rng.uniform(-1, 1, (50, 3)), rng.uniform(0.01, 0.2, 50))
        run_both(ctx, s2, incoherent_rays(300_003, lo, hi, 5))
    finally:
        ctx.set_launch(*saved)


@pytest.mark.parametrize("variant", [0, 1, 3, 4])
@pytest.mark.parametrize("any_hit", [False, True])
def test_visit_counts_equal_oracle(ctx, any_hit, variant):
    """The kernel's box/primitive test counts (the basis of the algorithmic-bytes figure) equal the oracle's
    reference-order counts summed over the ray set, which holds only if every lane replays the reference visit
    sequence: for the pair-record kernel (variant 0, which counts a far child's box test when it pops it,
    where the reference makes it) and the node-record kernels (1, 3)."""
    saved = ctx.launch()
    ctx.set_launch(saved[0], saved[1], saved[2], variant)
    rng = np.random.default_rng(8)
    n = 2000
    c = rng.uniform(-2, 2, (n, 1, 3))
    tris = (c + rng.uniform(-0.2, 0.2, (n, 3, 3))).astype(np.float32).reshape(n, 9)
    s = SceneData(tri_positions=tris)
    add_spheres(s, rng.uniform(-2, 2, (100, 3)), rng.uniform(0.01, 0.5, 100))
    rays = box_rays(100_000, -3, 3, 19, -2, 2)
    bvh = GpuBvh(ctx, s)
    o = oracle_for(bvh, s)
    try:
        st = bvh.count(rays, any_hit)
    finally:
        ctx.set_launch(*saved)
    _, cnt = o.intersect(rays, any_hit=any_hit, counters=True)
    tot = cnt.astype(np.uint64).sum(0)
    assert (st.nodes_visited, st.tris_tested, st.spheres_tested) == tuple(int(x) for x in tot)


def grid_scene_and_rays(n_tris, n_rays, seed):
    """Triangles with vertices on a 0.5 grid (so every box plane is a grid value) and rays whose origins
    are grid points and whose direction components come from {+0, -0, +-0.5, +-1, +-2}: plane
    distances (plane - origin) * (1 / 0) are 0 * inf = NaN or +-inf on most box tests, the cases
    where the slab test's comparisons decide by NaN ordering."""
    rng = np.random.default_rng(seed)
    c = rng.integers(-4, 5, (n_tris, 1, 3)) * 0.5
    tris = (c + rng.integers(-1, 2, (n_tris, 3, 3)) * 0.5).astype(np.float32).reshape(n_tris, 9)
    r = np.zeros((n_rays, 8), np.float32)
    r[:, 0:3] = rng.integers(-6, 7, (n_rays, 3)) * 0.5
    comps = np.array([0.0, -0.0, 0.5, -0.5, 1.0, -1.0, 2.0, -2.0], np.float32)
    d = comps[rng.integers(0, len(comps), (n_rays, 3))]
    d[np.all(d == 0, axis=1), 2] = 1.0
    r[:, 4:7] = d
    r[:, 3] = np.inf
    return SceneData(tri_positions=tris), r


@pytest.mark.parametrize("variant", [0, 1, 4])
def test_axis_aligned_rays_on_box_planes(ctx, variant):
    s, rays = grid_scene_and_rays(3000, 100_000, 29)
    saved = ctx.launch()
    ctx.set_launch(saved[0], saved[1], saved[2], variant)
    try:
        bvh = GpuBvh(ctx, s)
        o = oracle_for(bvh, s)
        for any_hit in (False, True):
            g = bvh.intersect_test(rays) if any_hit else bvh.intersect(rays)
            ref = o.intersect(rays, any_hit=any_hit)
            assert_hits_equal(g, ref, s.n_tris, any_hit)
            st = bvh.count(rays, any_hit)
            _, cnt = o.intersect(rays, any_hit=any_hit, counters=True)
            assert (st.nodes_visited, st.tris_tested) == tuple(int(x) for x in cnt.astype(np.uint64).sum(0)[:2])
        hit = O.hits_as_words(bvh.intersect(rays))[:, 0] != MISS
        assert 0.05 < hit.mean() < 0.95
    finally:
        ctx.set_launch(*saved)


def test_degenerate_triangles_are_never_hit(ctx):
    """The reference rejects every hit on a degenerate triangle (triangle.rs:241-268); the traversal gets
    the same result from the collapsed slot vertices the scene build stores for it. Collinear, repeated-
    vertex and tiny triangles (whose squared cross product underflows, so the check calls them degenerate
    although the ray meets them) sit in front of a backdrop; every ray aimed at them must reach the
    backdrop, with visit counts equal to the oracle's."""
    back = np.array([[-5, -5, 5, 5, -5, 5, 5, 5, 5], [-5, -5, 5, 5, 5, 5, -5, 5, 5]], np.float32)
    degen = np.array([[0, 0, 0, 1, 1, 0, 2, 2, 0],            # collinear
                      [0, 1, 0, 0, 1, 0, 1, 0, 0],            # repeated vertex
                      [1e-20, 0, 0, 0, 1e-20, 0, 0, 0, 0],    # tiny: |cross|^2 underflows in f32
                      [0.5, 0.5, 0, 0.5, 0.5, 0, 0.5, 0.5, 0]], np.float32)  # a point
    s = SceneData(tri_positions=np.concatenate([back, degen]))
    aims = np.array([[1.0, 1.0], [0.2, 0.6], [3e-21, 3e-21], [0.5, 0.5], [1.5, 1.5], [0.1, 0.9]], np.float32)
    rng = np.random.default_rng(12)
    aims = np.concatenate([aims, rng.uniform(-0.5, 2.5, (500, 2)).astype(np.float32)])
    rays = np.zeros((len(aims), 8), np.float32)
    rays[:, 0:2] = aims
    rays[:, 2] = -1.0
    rays[:, 3] = np.inf
    rays[:, 6] = 1.0
    bvh, o = run_both(ctx, s, rays)
    hits = O.hits_as_words(bvh.intersect(rays))
    assert set(hits[:, 0].tolist()) <= {0, 1}  # the backdrop only
    st = bvh.count(rays, False)
    _, cnt = o.intersect(rays, any_hit=False, counters=True)
    assert (st.nodes_visited, st.tris_tested) == tuple(int(x) for x in cnt.astype(np.uint64).sum(0)[:2])


def test_deep_tree_really_is_deep(ctx):
    bvh = GpuBvh(ctx, deep_scene())
    assert bvh.info.max_depth > 20
    o = oracle_for(bvh, deep_scene())
    _, cnt = o.intersect(deep_rays(2000, 3), counters=True)
    assert cnt[:, 0].max() > 20  # rays really walk down the deep path


def test_full_size_intersection_benchmark_subset(ctx):
    """The 2^24-ray, 1.31M-triangle workload (coherent and incoherent sets): every ray on the GPU, random
    subsets of closest and any hits (the any-hit id is the first hit in the reference's traversal
    order) against the oracle, and a size-independent property on all rays (any-hit reports a hit
    exactly when closest-hit does)."""
    s = intersection_scene(8)
    assert s.n_tris == 1_310_720
    bvh = GpuBvh(ctx, s)
    o = oracle_for(bvh, s)
    cam = intersection_camera(4096, 4096)
    rays = ctx.camera_rays(cam)
    closest = O.hits_as_words(bvh.intersect(rays))
    anyh = O.hits_as_words(bvh.intersect_test(rays))
    assert np.array_equal(closest[:, 0] != MISS, anyh[:, 0] != MISS)
    rng = np.random.default_rng(0)
    idx = rng.choice(rays.shape[0], 40_000, replace=False)
    ref = o.intersect(rays[idx])
    assert np.array_equal(closest[idx, 0], ref[:, 0]) and same_bits(closest[idx, 1:], ref[:, 1:])
    ref = o.intersect(rays[idx], any_hit=True)
    assert np.array_equal(anyh[idx, 0], ref[:, 0]) and same_bits(anyh[idx, 1:], ref[:, 1:])
    del rays, closest, anyh
    lo, hi = s.bounds()
    inc = incoherent_rays(1 << 24, lo, hi, 11)
    g = O.hits_as_words(bvh.intersect(inc))
    ga = O.hits_as_words(bvh.intersect_test(inc))
    assert np.array_equal(g[:, 0] != MISS, ga[:, 0] != MISS)
    idx = rng.choice(inc.shape[0], 20_000, replace=False)
    ref = o.intersect(inc[idx])
    assert np.array_equal(g[idx, 0], ref[:, 0]) and same_bits(g[idx, 1:], ref[:, 1:])
    ref = o.intersect(inc[idx], any_hit=True)
    assert np.array_equal(ga[idx, 0], ref[:, 0]) and same_bits(ga[idx, 1:], ref[:, 1:])
