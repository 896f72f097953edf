"""A scene ingested through OBJ::read (io/obj.rs:16-124, via bg_obj_read) on the
GPU path: the loaded, transformed mesh in a Cornell box, traversed and
rendered on the GPU, against the oracle on the same triangles."""
import numpy as np
import pytest

import cpu_oracle as O
from bonsai_amd import GpuBvh
from bonsai_amd.integrator import MODE_CORRECTED, MODE_REF, Film, GpuPathIntegrator, rel_mse, resolve
from bonsai_amd.scenes import cornell_box, icosphere, incoherent_rays, load_obj
from test_gpu_intersect import assert_hits_equal

pytestmark = pytest.mark.gpu


def write_obj(path, tris):
    """Triangles (n, 9) as an OBJ file with shared vertices and 'f a/b c/d e/f' faces."""
    verts, index, faces = [], {}, []
    for t in np.asarray(tris, np.float32).reshape(-1, 3, 3):
        f = []
        for v in t:
            key = tuple(float(x) for x in v)
            if key not in index:
                index[key] = len(verts) + 1
                verts.append(v)
            f.append(index[key])
        faces.append(f)
    lines = [f"v {np.float32(v[0])} {np.float32(v[1])} {np.float32(v[2])}" for v in verts]
    lines += ["vt 0 0", "vt 1 0", "vt 1 1"]
    lines += [f"f {a}/1 {b}/2 {c}/3" for a, b, c in faces]
    path.write_text("\n".join(lines) + "\n")


@pytest.fixture(scope="module")
def obj_scene(tmp_path_factory):
    path = tmp_path_factory.mktemp("obj") / "blob.obj"
    write_obj(path, icosphere(3, (0.0, 0.0, 0.0), 1.0, 0.2, 3.0, 11))
    # object_to_world: scale 0.22 about the origin, then translate into the box
    m = np.eye(4, dtype=np.float32)
    m[0, 0] = m[1, 1] = m[2, 2] = 0.22
    m[0:3, 3] = (0.62, 0.3, 0.4)
    mesh = load_obj(str(path), m)
    s = cornell_box(double_sided_light=True, width=64, height=64)
    n_box = s.n_tris
    s.tri_positions = np.concatenate([s.tri_positions, mesh["positions"]]).astype(np.float32)
    s.tri_material = np.concatenate([s.tri_material, np.full(len(mesh["positions"]), 4, np.uint32)])
    return s, n_box, mesh


def test_obj_mesh_is_where_the_transform_puts_it(obj_scene):
    _, _, mesh = obj_scene
    p = mesh["positions"].reshape(-1, 3)
    c = p.mean(0)
    np.testing.assert_allclose(c, (0.62, 0.3, 0.4), atol=0.03)
    r = np.linalg.norm(p - (0.62, 0.3, 0.4), axis=1)
    assert 0.22 * 0.75 < r.min() and r.max() < 0.22 * 1.25


def test_obj_scene_hits_bit_exact(ctx, obj_scene):
    s, _, _ = obj_scene
    bvh = GpuBvh(ctx, s)
    o = O.OracleScene(s)
    nodes, slots = bvh.bvh()
    o.use_bvh(nodes, slots)
    rays = incoherent_rays(50_000, np.zeros(3, np.float32), np.ones(3, np.float32), 5)
    for any_hit in (False, True):
        g = bvh.intersect_test(rays) if any_hit else bvh.intersect(rays)
        # ids, t and triangle barycentrics bit-exact; sphere (u, v) to 1e-6 (atan2 / acos)
        assert_hits_equal(g, o.intersect(rays, any_hit=any_hit), s.n_tris, any_hit)


@pytest.mark.parametrize("mode", [MODE_REF, MODE_CORRECTED])
def test_obj_scene_render_matches_oracle(ctx, obj_scene, mode):
    s, _, _ = obj_scene
    bvh = GpuBvh(ctx, s)
    gi = GpuPathIntegrator(ctx, s.camera, spp=8, max_depth=5, mode=mode, seed=41)
    film = Film(ctx, 64, 64)
    gi.render_into(bvh, film)
    rgb, w = film.sums()
    img = resolve(rgb.astype(np.float64), w.astype(np.float64))
    o = O.OracleScene(s)
    nodes, slots = bvh.bvh()
    o.use_bvh(nodes, slots)
    o.set_shading(s)
    ref_rgb, ref_w = o.render(s, 8, 5, mode=mode, seed=41)
    assert np.array_equal(w, ref_w.astype(np.float32))
    assert rel_mse(img, resolve(ref_rgb, ref_w)) < 1e-4
    # the mesh is in view: the same render without it differs
    plain = cornell_box(double_sided_light=True, width=64, height=64)
    img2 = GpuPathIntegrator(ctx, plain.camera, spp=8, max_depth=5, mode=mode, seed=41).render(GpuBvh(ctx, plain))
    assert rel_mse(img2, img) > 1e-3
