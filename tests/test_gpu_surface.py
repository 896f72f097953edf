"""bg_surface_interactions (the Interaction GeometricPrimitive::intersect fills,
primitives/geometric.rs:27-38) on the GPU against the oracle's triangle.rs /
sphere.rs restatement: point, its error bounds, t, geometric and shading
normals and shading dpdu of every hit within 1e-6 relative to the vector's
largest component, uv (parameters on [0, 1]) within 1e-6 + 1e-6 relative as the
hits' (u, v) are tested; triangles bit-exact in every field. On triangles of
both orientations and full, partial, rotated and reversed spheres."""
import numpy as np
import pytest

import cpu_oracle as O
from bonsai_amd import GpuBvh
from bonsai_amd._native import BonsaiError
from bonsai_amd.scenes import SceneData
from test_gpu_intersect import box_rays, oracle_for

pytestmark = pytest.mark.gpu

FIELDS = [("p", slice(1, 4)), ("p_error", slice(4, 7)), ("n", slice(7, 10)), ("ns", slice(10, 13)),
          ("dpdu", slice(13, 16)), ("uv", slice(16, 18))]


def rot_z(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, -s, 0, 0], [s, c, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]], np.float64)


def surface_scene(seed=4):
    rng = np.random.default_rng(seed)
    n = 1500
    c = rng.uniform(-2, 2, (n, 1, 3))
    s = SceneData(tri_positions=(c + rng.uniform(-0.25, 0.25, (n, 3, 3))).astype(np.float32).reshape(n, 9))
    s.tri_reverse = (rng.random(n) < 0.5).astype(np.uint8)
    o2w, w2o, prm = [], [], []
    for k in range(60):
        t = np.eye(4)
        t[:3, 3] = rng.uniform(-2, 2, 3)
        r = rot_z(rng.uniform(0, 2 * np.pi)) if k % 3 == 0 else np.eye(4)
        m = t @ r
        o2w.append(m)
        w2o.append(np.linalg.inv(m))
        rad = rng.uniform(0.1, 0.5)
        prm.append([rad, -rad, rad, 360.0] if k % 2 == 0 else [rad, -0.6 * rad, 0.7 * rad, 270.0])
    s.sphere_o2w = np.asarray(o2w, np.float32).reshape(-1, 16)
    s.sphere_w2o = np.asarray(w2o, np.float32).reshape(-1, 16)
    s.sphere_params = np.asarray(prm, np.float32)
    s.sphere_reverse = (np.arange(60) % 4 == 1).astype(np.uint8)
    return s


def test_surface_interactions_match_oracle(ctx):
    s = surface_scene()
    bvh = GpuBvh(ctx, s)
    o = oracle_for(bvh, s)
    rays = box_rays(40_000, -3, 3, 12, -2, 2)
    hits = bvh.intersect(rays)
    sf = bvh.surfaces(rays, hits)
    words = O.hits_as_words(hits)
    hit = words[:, 0] != 0xFFFFFFFF
    assert hit.sum() > 10_000 and (words[hit, 0] >= s.n_tris).sum() > 1000
    # misses: every field NaN
    for name, _ in FIELDS:
        assert np.isnan(sf[name][~hit]).all()
    exact_tri = 0
    worst = {name: 0.0 for name, _ in FIELDS}
    for i in np.flatnonzero(hit):
        prim = int(words[i, 0])
        ok, ref = o.surface(prim, rays[i])
        assert ok
        assert sf["t"][i] == ref[0]
        same = True
        for name, sl in FIELDS:
            g, r = np.asarray(sf[name][i], np.float64), ref[sl].astype(np.float64)
            if name == "uv":  # parameters on [0, 1]: |g - r| <= 1e-6 + 1e-6 |r| as a ratio of that bound
                err = (np.abs(g - r) / (1.0 + np.abs(r))).max()
            else:
                err = np.abs(g - r).max() / max(np.abs(r).max(), 1e-30)
            worst[name] = max(worst[name], err)
            same &= bool(np.array_equal(g, r))
            if prim < s.n_tris:
                assert err == 0.0, (name, prim, g, r)
        exact_tri += same and prim < s.n_tris
    print("worst relative error per field:", {k: f"{v:.1e}" for k, v in worst.items()},
          f"triangle hits bit-exact: {exact_tri}")
    for name, v in worst.items():
        assert v <= 1e-6, (name, v)


def test_surface_interactions_reject_unknown_primitives(ctx):
    s = surface_scene()
    bvh = GpuBvh(ctx, s)
    rays = box_rays(4, -3, 3, 13, -2, 2)
    hits = bvh.intersect(rays)
    hits["prim_id"][0] = s.n_tris + s.n_spheres
    with pytest.raises(BonsaiError):
        bvh.surfaces(rays, hits)
