"""OBJ ingestion (io/obj.rs:16-124, TriangleMesh::new triangle.rs:48-90): the
library's bg_obj_read against closed-form triangles and against the oracle's
restatement, including the reference's quirks (quads taken as consecutive
triples, 1-based indices, transforms applied to positions and normals)."""
import sys
from pathlib import Path

import numpy as np
import pytest

from bonsai_amd import _native as N
from bonsai_amd.scenes import SceneData, load_obj

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import obj as O  # noqa: E402

CUBE_CORNER = """# two triangles, every face-vertex syntax the reference accepts
o corner
v 0 0 0
v 1 0 0
v 0 1 0
v 0 0 1
vt 0 0
vt 1 0
vt 0 1
vn 0 0 -1
vn 0 -1 0
g faces
s off
f 1/1/1 3/3/1 2/2/1
f 1/1/2 2/2/2 4/3/2
"""


def write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_every_syntax_closed_form(tmp_path):
    path = write(tmp_path, "corner.obj", CUBE_CORNER)
    r = load_obj(path)
    np.testing.assert_array_equal(r["positions"], np.array([[0, 0, 0, 0, 1, 0, 1, 0, 0],
                                                            [0, 0, 0, 1, 0, 0, 0, 0, 1]], np.float32))
    np.testing.assert_array_equal(r["normals"], np.array([[0, 0, -1] * 3, [0, -1, 0] * 3], np.float32))
    np.testing.assert_array_equal(r["uvs"], np.array([[0, 0, 0, 1, 1, 0], [0, 0, 1, 0, 0, 1]], np.float32))
    for k, v in zip(("positions", "normals", "uvs"), O.read(path)):
        np.testing.assert_array_equal(r[k], v)


@pytest.mark.parametrize("face", ["f 1 2 3", "f 1/1 2/2 3/3", "f 1//1 2//1 3//1", "f +1 +2 +3"])
def test_single_triangle_forms(tmp_path, face):
    text = "v 1 2 3\nv 4 5 6\nv 7 8 9.5\nvt 0 0\nvt 1 0\nvt 0 1\nvn 0 0 1\n" + face + "\n"
    path = write(tmp_path, "t.obj", text)
    r = load_obj(path)
    np.testing.assert_array_equal(r["positions"], np.array([[1, 2, 3, 4, 5, 6, 7, 8, 9.5]], np.float32))
    assert (r["normals"] is not None) == ("//" in face)
    assert (r["uvs"] is not None) == ("/" in face and "//" not in face)
    for k, v in zip(("positions", "normals", "uvs"), O.read(path)):
        if v is None:
            assert r[k] is None
        else:
            np.testing.assert_array_equal(r[k], v)


def test_quads_are_consecutive_triples(tmp_path):
    # the reference does not fan polygons: indices 1 2 3 4 | 1 3 4 ... group as (1,2,3) (4,1,3) (4, ...)
    text = "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nf 1 2 3 4\nf 1 3 4 2 1\n"
    path = write(tmp_path, "q.obj", text)
    r = load_obj(path)
    v = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], np.float32)
    idx = [1, 2, 3, 4, 1, 3, 4, 2, 1]
    np.testing.assert_array_equal(r["positions"], v[np.array(idx) - 1].reshape(3, 9))


def test_transform_points_and_normals(tmp_path):
    path = write(tmp_path, "corner.obj", CUBE_CORNER)
    m = np.array([[2, 0, 0, 1], [0, 3, 0, -2], [0, 0, 4, 0.5], [0, 0, 0, 1]], np.float32)
    r = load_obj(path, m)
    p = O.read(path)[0].reshape(-1, 3)
    exp = p * np.array([2, 3, 4], np.float32) + np.array([1, -2, 0.5], np.float32)
    np.testing.assert_array_equal(r["positions"].reshape(-1, 3), exp.astype(np.float32))
    # normals by the inverse transpose: (0,0,-1) -> (0,0,-1/4), (0,-1,0) -> (0,-1/3,0)
    np.testing.assert_allclose(r["normals"][0, :3], [0, 0, -0.25], rtol=1e-7)
    np.testing.assert_allclose(r["normals"][1, :3], [0, -1 / 3, 0], rtol=1e-6)
    for k, v in zip(("positions", "normals"), O.read(path, m)):
        np.testing.assert_array_equal(r[k], v)


def test_random_mesh_matches_oracle(tmp_path):
    rng = np.random.default_rng(4)
    v = rng.uniform(-5, 5, (200, 3)).astype(np.float32)
    faces = rng.integers(1, 201, (300, 3))
    text = "".join(f"v {a!r} {b!r} {c!r}\n" for a, b, c in v.tolist())
    text += "".join(f"f {a} {b} {c}\n" for a, b, c in faces.tolist())
    path = write(tmp_path, "r.obj", text)
    ang = 0.7
    m = np.array([[np.cos(ang), -np.sin(ang), 0, 3], [np.sin(ang), np.cos(ang), 0, 0], [0, 0, 1, -1],
                  [0, 0, 0, 1]], np.float32)
    r = load_obj(path, m)
    np.testing.assert_array_equal(r["positions"], O.read(path, m)[0])
    s = SceneData(tri_positions=r["positions"])
    assert s.n_tris == 300


@pytest.mark.parametrize("text,code", [("v 0 0 0\nf 1 2 3\n", N.BG_ERR_INVALID),
                                       ("v 0 0\n", N.BG_ERR_INVALID),
                                       ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 0 1 2\n", N.BG_ERR_INVALID),
                                       ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf -1 -2 -3\n", N.BG_ERR_INVALID),
                                       ("v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 1\nf 1//1 2 3\n", N.BG_ERR_UNSUPPORTED)])
def test_malformed_files_fail_like_the_reference(tmp_path, text, code):
    path = write(tmp_path, "bad.obj", text)
    with pytest.raises(N.BonsaiError) as e:
        load_obj(path)
    assert e.value.code == code
    if code == N.BG_ERR_INVALID:
        with pytest.raises(O.ObjError):
            O.read(path)


def test_missing_file(tmp_path):
    with pytest.raises(N.BonsaiError):
        load_obj(str(tmp_path / "none.obj"))
