"""BVH build and traversal: structural invariants of the oracle's build
(bvh.rs:91-269), equality of the library's host builder with it, and the
oracle's reference-order traversal against brute force (bvh.rs:281-423)."""
import ctypes as C

import numpy as np
import pytest

import cpu_oracle as O
from bonsai_amd import _native as N
from bonsai_amd.scenes import SceneData, add_spheres, cornell_box, icosphere

MISS = 0xFFFFFFFF


def random_scene(seed, n_tris, n_sph=0, spread=4.0, size=0.3):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-spread, spread, (n_tris, 1, 3))
    tris = (c + rng.uniform(-size, size, (n_tris, 3, 3))).astype(np.float32).reshape(n_tris, 9)
    s = SceneData(tri_positions=tris)
    if n_sph:
        add_spheres(s, rng.uniform(-spread, spread, (n_sph, 3)), rng.uniform(0.05, 0.5, n_sph))
    return s


def product_bvh(scene):
    desc, keep = scene.desc()
    nn, npr, depth = C.c_uint64(0), C.c_uint64(0), C.c_uint32(0)
    L = N.lib()
    N.check(L.bg_build_bvh(C.byref(desc), None, C.byref(nn), None, C.byref(npr), C.byref(depth)))
    nodes = np.zeros((nn.value, 8), np.uint32)
    slots = np.zeros(npr.value, np.uint32)
    N.check(L.bg_build_bvh(C.byref(desc), nodes.ctypes.data, C.byref(nn), slots.ctypes.data, C.byref(npr),
                           C.byref(depth)))
    return nodes, slots, depth.value


def check_structure(nodes, slots, n_prims, prim_bounds=None):
    f = nodes.view(np.float32)
    assert sorted(slots.tolist()) == list(range(n_prims))  # each primitive in exactly one leaf
    n = nodes.shape[0]
    covered = np.zeros(n_prims, np.int32)
    for i in range(n):
        cnt = nodes[i, 7] >> 3
        lo, hi = f[i, 0:3], f[i, 3:6]
        assert np.all(lo <= hi)
        if cnt:
            off = nodes[i, 6]
            covered[off:off + cnt] += 1
            if prim_bounds is not None:
                pb = prim_bounds[slots[off:off + cnt]]
                assert np.all(pb[:, 0:3] >= lo) and np.all(pb[:, 3:6] <= hi)
        else:
            a, b = i + 1, nodes[i, 6]
            assert b > a and b < n  # DFS layout: first child follows, second child after its subtree
            for ch in (a, b):
                assert np.all(f[ch, 0:3] >= lo) and np.all(f[ch, 3:6] <= hi)
    assert np.all(covered == 1)


def tri_bounds(tris):
    v = tris.reshape(-1, 3, 3)
    return np.concatenate([v.min(1), v.max(1)], axis=1)


@pytest.mark.parametrize("seed,n", [(1, 1), (2, 2), (3, 3), (4, 17), (5, 300), (6, 3000)])
def test_oracle_build_invariants(seed, n):
    s = random_scene(seed, n)
    o = O.OracleScene(s)
    nodes, slots = o.build()
    check_structure(nodes, slots, n, tri_bounds(s.tri_positions))


def test_coincident_centroids_make_one_leaf():
    # all centroids equal -> leaf regardless of size (bvh.rs:127-137), even above 256 primitives
    tri = np.array([[-1, -1, 0, 1, -1, 0, 0, 2, 0]], np.float32)
    s = SceneData(tri_positions=np.repeat(tri, 300, axis=0))
    nodes, slots = O.OracleScene(s).build()
    assert nodes.shape[0] == 1 and nodes[0, 7] >> 3 == 300
    assert slots.tolist() == list(range(300))


def test_leaf_size_bound():
    s = random_scene(9, 5000, spread=1.0, size=0.01)
    nodes, _ = O.OracleScene(s).build()
    assert int((nodes[:, 7] >> 3).max()) <= 256


@pytest.mark.parametrize("make", [
    lambda: random_scene(21, 1),
    lambda: random_scene(22, 50),
    lambda: random_scene(23, 2000, 40),
    lambda: random_scene(24, 20000, 0, spread=1.0, size=0.05),
    lambda: cornell_box(),
    lambda: SceneData(tri_positions=icosphere(5, (0.0, 0.0, 0.0), 1.0, 0.15, 4.0, 7)),
])
def test_library_build_equals_oracle_build(make):
    s = make()
    o = O.OracleScene(s)
    on, os_ = o.build()
    pn, ps, depth = product_bvh(s)
    assert np.array_equal(on, pn)
    assert np.array_equal(os_, ps)
    assert depth == o.height()


def test_library_build_empty_scene():
    pn, ps, depth = product_bvh(SceneData())
    assert pn.shape[0] == 0 and ps.shape[0] == 0


@pytest.mark.parametrize("seed", [31, 32, 33])
def test_traversal_matches_brute_force(seed):
    s = random_scene(seed, 400, 20, spread=2.0, size=0.5)
    o = O.OracleScene(s)
    nodes, slots = o.build()
    rng = np.random.default_rng(seed)
    n = 3000
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = rng.uniform(-5, 5, (n, 3))
    rays[:, 3] = np.inf
    rays[:, 4:7] = rng.uniform(-2, 2, (n, 3)) - rays[:, 0:3]  # aimed at the primitives' region
    bvh = o.intersect(rays)
    brute = o.brute(rays, slots)
    assert np.array_equal(bvh[:, 0], brute[:, 0])
    assert np.array_equal(bvh[:, 1], brute[:, 1])  # t bit-exact
    anyh = o.intersect(rays, any_hit=True)
    assert np.array_equal(anyh[:, 0] != MISS, bvh[:, 0] != MISS)
    assert (bvh[:, 0] != MISS).mean() > 0.2


def test_counters_are_consistent():
    s = random_scene(41, 1000)
    o = O.OracleScene(s)
    o.build()
    rng = np.random.default_rng(1)
    rays = np.zeros((500, 8), np.float32)
    rays[:, 0:3] = rng.uniform(-5, 5, (500, 3))
    rays[:, 3] = np.inf
    rays[:, 4:7] = rng.standard_normal((500, 3))
    hits, cnt = o.intersect(rays, counters=True)
    assert np.all(cnt[:, 0] >= 1)                    # the root box is always tested
    assert np.all(cnt[hits[:, 0] != MISS, 1] >= 1)   # a hit needs a primitive test
