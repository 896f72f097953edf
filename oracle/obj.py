"""Oracle (test infrastructure only): restatement of the reference's OBJ loader
(io/obj.rs:16-124 OBJ::read) and of TriangleMesh::new (shapes/triangle.rs:48-90)
in NumPy float32 arithmetic, for checking the library's bg_obj_read.

Pinned by tests/test_obj.py: hand-written files whose triangles are known in
closed form (every face-vertex syntax, comments, quads split as consecutive
triples, translate / scale transforms), and malformed files the reference
would fail on."""
from __future__ import annotations

import numpy as np


class ObjError(ValueError):
    pass


def _index(tok: str) -> int:
    t = tok[1:] if tok.startswith("+") else tok
    if not t or not t.isdigit():
        raise ObjError(f"bad index {tok!r}")
    return int(t)


def read(path, object_to_world=None, world_to_object=None):
    """(positions (n, 9), normals (n, 9) or None, uvs (n, 6) or None), float32."""
    verts, normals, uvs = [], [], []
    vi, ni, ti = [], [], []
    with open(path) as f:
        for raw in f:
            line = raw.strip()
            if line.startswith("f"):                      # obj.rs:33-58
                for v in line.split()[1:]:
                    if "//" in v:
                        p = v.split("//")
                        if len(p) != 2:
                            raise ObjError(v)
                        vi.append(_index(p[0]))
                        ni.append(_index(p[1]))
                    elif "/" in v:
                        p = v.split("/")
                        if len(p) not in (2, 3):
                            raise ObjError(v)
                        vi.append(_index(p[0]))
                        ti.append(_index(p[1]))
                        if len(p) == 3:
                            ni.append(_index(p[2]))
                    else:
                        vi.append(_index(v))
            if line.startswith("v"):                      # obj.rs:61-77
                try:
                    p = [np.float32(x) for x in line.split()[1:]]
                except ValueError as e:
                    raise ObjError(line) from e
                if line.startswith("vt"):
                    if len(p) < 2:
                        raise ObjError(line)
                    uvs.append(p[:2])
                elif line.startswith("vn"):
                    if len(p) < 3:
                        raise ObjError(line)
                    normals.append(p[:3])
                else:
                    if len(p) < 3:
                        raise ObjError(line)
                    verts.append(p[:3])
    for idx, arr in ((vi, verts), (ni, normals), (ti, uvs)):
        if any(k < 1 or k > len(arr) for k in idx):
            raise ObjError("index out of range")
    n = len(vi) // 3                                       # obj.rs:111-123
    pos = np.array([verts[k - 1] for k in vi[:3 * n]], np.float32).reshape(n, 9)
    nrm = np.array([normals[k - 1] for k in ni[:3 * n]], np.float32).reshape(n, 9) if ni else None
    uv = np.array([uvs[k - 1] for k in ti[:3 * n]], np.float32).reshape(n, 6) if ti else None
    m = np.eye(4, dtype=np.float32) if object_to_world is None else np.asarray(object_to_world, np.float32)
    if not np.array_equal(m, np.eye(4, dtype=np.float32)):  # triangle.rs:49-57
        mi = (np.linalg.inv(m.astype(np.float64)).astype(np.float32) if world_to_object is None
              else np.asarray(world_to_object, np.float32))
        p = pos.reshape(-1, 3)
        x, y, z = p[:, 0], p[:, 1], p[:, 2]
        # point3.rs:51-69, float32 products summed left to right
        xp = m[0, 0] * x + m[0, 1] * y + m[0, 2] * z + m[0, 3]
        yp = m[1, 0] * x + m[1, 1] * y + m[1, 2] * z + m[1, 3]
        zp = m[2, 0] * x + m[2, 1] * y + m[2, 2] * z + m[2, 3]
        wp = m[3, 0] * x + m[3, 1] * y + m[3, 2] * z + m[3, 3]
        one = wp == np.float32(1.0)
        iw = np.where(one, np.float32(1.0), np.float32(1.0) / np.where(one, np.float32(1.0), wp))
        out = np.stack([np.where(one, xp, xp * iw), np.where(one, yp, yp * iw), np.where(one, zp, zp * iw)], 1)
        pos = out.astype(np.float32).reshape(n, 9)
        if nrm is not None:                                # normal.rs:21-30
            q = nrm.reshape(-1, 3)
            x, y, z = q[:, 0], q[:, 1], q[:, 2]
            out = np.stack([mi[0, 0] * x + mi[1, 0] * y + mi[2, 0] * z,
                            mi[0, 1] * x + mi[1, 1] * y + mi[2, 1] * z,
                            mi[0, 2] * x + mi[1, 2] * y + mi[2, 2] * z], 1)
            nrm = out.astype(np.float32).reshape(n, 9)
    return pos, nrm, uv
