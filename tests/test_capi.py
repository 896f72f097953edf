"""The C-ABI library loads without a GPU and exports every function its headers declare."""
import ctypes as C
import re
from pathlib import Path

import numpy as np

from bonsai_amd import _native as N

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[\w\s\*]+?\b(bgx?_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_headers_declare_entry_points():
    names = declared_functions()
    assert {"bg_create", "bg_scene_create", "bg_intersect", "bg_intersect_device", "bg_camera_rays",
            "bg_build_bvh", "bg_look_at", "bgx_icosphere"} <= names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(str(N.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # every declared function is bound with a signature in the Python binding
    assert declared_functions() <= set(N.SIGNATURES)


def test_abi_version_and_errors_without_device():
    L = N.lib()
    assert L.bg_abi_version() == 4
    cnt = C.c_int(-1)
    assert L.bg_device_count(C.byref(cnt)) == 0 and cnt.value >= 0
    assert L.bg_device_count(None) == N.BG_ERR_INVALID
    h = C.c_void_p()
    assert L.bg_create(10_000, C.byref(h)) == N.BG_ERR_INVALID  # no such device
    assert L.bg_scene_create(None, None, None) == N.BG_ERR_INVALID
    assert L.bg_intersect(None, None, None, 0, None, 0, None) == N.BG_ERR_INVALID


def test_look_at_matches_oracle():
    import cpu_oracle as O
    rng = np.random.default_rng(2)
    for _ in range(50):
        pos, look = rng.uniform(-5, 5, 3).astype(np.float32), rng.uniform(-5, 5, 3).astype(np.float32)
        up = rng.standard_normal(3).astype(np.float32)
        m = np.zeros(16, np.float32)
        mi = np.zeros(16, np.float32)
        assert N.lib().bg_look_at(pos.ctypes.data, look.ctypes.data, up.ctypes.data, m.ctypes.data,
                                  mi.ctypes.data) == 0
        om, omi = O.look_at(pos, look, up)
        assert np.array_equal(m.view(np.uint32), om.view(np.uint32))
        assert np.array_equal(mi.view(np.uint32), omi.view(np.uint32))


def test_generators_are_deterministic():
    from bonsai_amd.scenes import icosphere, incoherent_rays
    a = icosphere(3, (0.0, 0.0, 0.0), 1.0, 0.15, 4.0, 7)
    b = icosphere(3, (0.0, 0.0, 0.0), 1.0, 0.15, 4.0, 7)
    assert a.shape == (1280, 9) and np.array_equal(a, b)
    c = icosphere(3, (0.0, 0.0, 0.0), 1.0, 0.15, 4.0, 8)
    assert not np.array_equal(a, c)
    r = incoherent_rays(10000, (-1, -1, -1), (1, 1, 1), 3)
    assert np.all(np.abs(r[:, :3]) <= 1)
    np.testing.assert_allclose(np.linalg.norm(r[:, 4:7], axis=1), 1, atol=1e-6)
    assert abs(r[:, 4:7].mean(0)).max() < 0.03  # isotropic
