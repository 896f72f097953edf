"""ctypes wrapper of the CPU oracle (oracle/liboracle.so), for tests only."""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
ORACLE_DIR = ROOT / "oracle"
ORACLE_SO = ORACLE_DIR / "liboracle.so"

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not ORACLE_SO.exists():
            subprocess.run(["make", "-C", str(ORACLE_DIR)], check=True, capture_output=True)
        L = C.CDLL(str(ORACLE_SO))
        P = C.c_void_p
        sig = {
            "orc_scene_new": (P, [P, C.c_uint64, P, P, P, P, P, C.c_uint64]),
            "orc_scene_free": (None, [P]),
            "orc_bvh_build": (C.c_uint64, [P]),
            "orc_bvh_height": (C.c_uint32, [P]),
            "orc_bvh_export": (None, [P, P, P]),
            "orc_bvh_import": (None, [P, P, C.c_uint64, P, C.c_uint64]),
            "orc_intersect": (None, [P, P, C.c_uint64, C.c_uint32, P, P, C.c_int]),
            "orc_intersect_brute": (None, [P, P, C.c_uint64, P, P]),
            "orc_prim_intersect": (C.c_int, [P, C.c_uint32, P, C.c_uint32, P]),
            "orc_surface": (C.c_int, [P, C.c_uint32, P, P]),
            "orc_camera_rays": (None, [P, P, C.c_float, C.c_float, C.c_float, C.c_uint32, C.c_uint32,
                                       C.c_float, C.c_float, P]),
            "orc_gamma": (C.c_float, [C.c_float]),
            "orc_next_float_up": (C.c_float, [C.c_float]),
            "orc_next_float_down": (C.c_float, [C.c_float]),
            "orc_look_at": (None, [P, P, P, P, P]),
            "orc_translate_scale": (None, [P, C.c_float, C.c_float, C.c_float, P, P]),
            "orc_mat4_inverse": (None, [P, P]),
            "orc_scene_set_shading": (None, [P, C.c_uint32, P, P, P, C.c_uint32, P, P, P]),
            "orc_render": (None, [P, P, P] + [C.c_float] * 7 + [C.c_uint32] * 4 + [C.c_float, C.c_uint64] +
                           [C.c_uint32] * 7 + [C.c_int, P, P]),
            "orc_bsdf_probe": (None, [C.c_uint32, P, C.c_int, P, P, C.c_float, C.c_float, P]),
            "orc_light_sample": (None, [P, C.c_uint32, C.c_int, C.c_uint64, P, P, P, P, P, P]),
            "orc_light_pdf": (None, [P, C.c_uint32, C.c_int, C.c_uint64, P, P, P, P, P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


class OracleScene:
    """Scene + BVH held by the oracle."""

    def __init__(self, scene):
        self._keep = []

        def arr(a, dt, shape=None):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dt)
            self._keep.append(a)
            return a

        tp = arr(scene.tri_positions, np.float32)
        tf = arr(scene.tri_reverse, np.uint8)
        so = arr(scene.sphere_o2w, np.float32)
        sw = arr(scene.sphere_w2o, np.float32)
        sp = arr(scene.sphere_params, np.float32)
        sr = arr(scene.sphere_reverse, np.uint8)
        self.n_prims = scene.n_tris + scene.n_spheres
        self.h = lib().orc_scene_new(_ptr(tp), scene.n_tris, _ptr(tf), _ptr(so), _ptr(sw), _ptr(sp), _ptr(sr),
                                     scene.n_spheres)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_scene_free(self.h)
            self.h = None

    def build(self) -> tuple[np.ndarray, np.ndarray]:
        n = lib().orc_bvh_build(self.h)
        nodes = np.empty((n, 8), np.uint32)
        ordered = np.empty(self.n_prims, np.uint32)
        lib().orc_bvh_export(self.h, nodes.ctypes.data, ordered.ctypes.data)
        return nodes, ordered

    def height(self) -> int:
        return int(lib().orc_bvh_height(self.h))

    def use_bvh(self, nodes: np.ndarray, slots: np.ndarray) -> None:
        nodes = np.ascontiguousarray(nodes, np.uint32)
        slots = np.ascontiguousarray(slots, np.uint32)
        self._keep += [nodes, slots]
        lib().orc_bvh_import(self.h, nodes.ctypes.data, nodes.shape[0], slots.ctypes.data, slots.shape[0])

    def intersect(self, rays: np.ndarray, any_hit: bool = False, counters: bool = False,
                  threads: int | None = None):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0]
        hits = np.empty((n, 4), np.uint32)
        cnt = np.empty((n, 3), np.uint32) if counters else None
        th = threads if threads is not None else min(8, os.cpu_count() or 1)
        lib().orc_intersect(self.h, rays.ctypes.data, n, 1 if any_hit else 0, hits.ctypes.data, _ptr(cnt), th)
        return (hits, cnt) if counters else hits

    def brute(self, rays: np.ndarray, order: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        order = np.ascontiguousarray(order, np.uint32)
        hits = np.empty((rays.shape[0], 4), np.uint32)
        lib().orc_intersect_brute(self.h, rays.ctypes.data, rays.shape[0], order.ctypes.data, hits.ctypes.data)
        return hits

    def set_shading(self, scene) -> None:
        mats = scene.materials
        mt = np.array([m.type for m in mats], np.uint32)
        mp = np.array([[*m.reflectance, m.alpha, *m.eta, *m.k] for m in mats], np.float32).reshape(-1, 10)
        n_prims = scene.n_tris + scene.n_spheres
        pm = np.zeros(n_prims, np.uint32)
        if scene.tri_material is not None:
            pm[:scene.n_tris] = scene.tri_material
        if scene.sphere_material is not None:
            pm[scene.n_tris:] = scene.sphere_material
        lp = np.array([l.prim_id for l in scene.lights], np.uint32)
        lL = np.array([list(l.radiance) for l in scene.lights], np.float32).reshape(-1, 3)
        lds = np.array([l.double_sided for l in scene.lights], np.uint8)
        self._keep += [mt, mp, pm, lp, lL, lds]
        lib().orc_scene_set_shading(self.h, len(mats), mt.ctypes.data, mp.ctypes.data, pm.ctypes.data, len(lp),
                                    lp.ctypes.data, lL.ctypes.data, lds.ctypes.data)

    def render(self, scene, spp: int, max_depth: int, mode: int = 0, seed: int = 1, rr: float = 1.0,
               window=None, threads: int | None = None, integrator: int = 0, strategy: int = 0):
        """Returns (rgb_sum (h, w, 3) float64, weight_sum (h, w)) for the window (x0, y0, x1, y1).
        integrator: 0 path, 1 direct lighting (strategy 0 all lights, 1 one light), 2 Whitted."""
        cam = scene.camera
        m, mi = cam.matrices()
        x0, y0, x1, y1 = window or (0, 0, cam.width, cam.height)
        rgb = np.zeros((y1 - y0, x1 - x0, 3), np.float64)
        w = np.zeros((y1 - y0, x1 - x0), np.float64)
        th = threads if threads is not None else len(os.sched_getaffinity(0))
        lib().orc_render(self.h, m.ctypes.data, mi.ctypes.data, cam.fov, cam.near, cam.far, cam.lens_radius,
                         cam.focal_distance, cam.shutter_open, cam.shutter_close, cam.width, cam.height, spp,
                         max_depth, rr, seed, mode, x0, y0, x1, y1, integrator, strategy, th, rgb.ctypes.data,
                         w.ctypes.data)
        return rgb, w

    def light_sample(self, light: int, mode: int, p, perr, n, u0, u1) -> np.ndarray:
        """DiffuseAreaLight::sample_point from points (p, perr, n): (m, 10) = light point, its normal,
        solid-angle pdf, radiance toward the point (set_shading first)."""
        a = [np.ascontiguousarray(x, np.float32) for x in (p, perr, n, u0, u1)]
        m = a[3].size
        out = np.empty((m, 10), np.float32)
        self._keep += a
        lib().orc_light_sample(self.h, light, mode, m, *[x.ctypes.data for x in a], out.ctypes.data)
        return out

    def light_pdf(self, light: int, mode: int, p, perr, n, wi) -> np.ndarray:
        a = [np.ascontiguousarray(x, np.float32).reshape(-1, 3) for x in (p, perr, n, wi)]
        out = np.empty(a[0].shape[0], np.float32)
        lib().orc_light_pdf(self.h, light, mode, a[0].shape[0], *[x.ctypes.data for x in a], out.ctypes.data)
        return out

    def surface(self, prim_id: int, ray):
        """(ok, 18 floats: t, p, p_error, n, ns, shading dpdu, u, v) of primitive prim_id on a ray."""
        r = np.ascontiguousarray(ray, np.float32).reshape(8)
        out = np.full(18, np.nan, np.float32)
        ok = lib().orc_surface(self.h, prim_id, r.ctypes.data, out.ctypes.data)
        return bool(ok), out

    def prim(self, prim_id: int, ray, kind: int = 0):
        r = np.ascontiguousarray(ray, np.float32).reshape(8)
        out = np.zeros(4, np.float32)
        ok = lib().orc_prim_intersect(self.h, prim_id, r.ctypes.data, kind, out.ctypes.data)
        return bool(ok), out


def camera_rays(m, mi, fov, width, height, fx=0.5, fy=0.5, near=0.01, far=1000.0) -> np.ndarray:
    m = np.ascontiguousarray(m, np.float32)
    mi = np.ascontiguousarray(mi, np.float32)
    out = np.empty((width * height, 8), np.float32)
    lib().orc_camera_rays(m.ctypes.data, mi.ctypes.data, fov, near, far, width, height, fx, fy, out.ctypes.data)
    return out


def look_at(pos, look, up):
    m = np.zeros(16, np.float32)
    mi = np.zeros(16, np.float32)
    a = [np.asarray(x, np.float32) for x in (pos, look, up)]
    lib().orc_look_at(a[0].ctypes.data, a[1].ctypes.data, a[2].ctypes.data, m.ctypes.data, mi.ctypes.data)
    return m, mi


def hits_as_words(h) -> np.ndarray:
    """Product hits (structured) -> (n, 4) uint32 words for bitwise comparison."""
    return np.ascontiguousarray(h).view(np.uint32).reshape(-1, 4)


def ray(o, d, t_max=np.inf, time=0.0) -> np.ndarray:
    return np.array([o[0], o[1], o[2], t_max, d[0], d[1], d[2], time], np.float32)


def bsdf_probe(material, mode, wo, wi, u0=0.5, u1=0.5) -> dict:
    """The oracle BSDF of a bg_material in the frame n = z, dpdu = x: f and pdf of (wo, wi), and one
    BSDF::sample with (u0, u1)."""
    params = np.array([*material.reflectance, material.alpha, *material.eta, *material.k], np.float32)
    a = np.ascontiguousarray(wo, np.float32)
    b = np.ascontiguousarray(wi, np.float32)
    out = np.zeros(13, np.float32)
    lib().orc_bsdf_probe(material.type, params.ctypes.data, mode, a.ctypes.data, b.ctypes.data, u0, u1,
                         out.ctypes.data)
    return {"f": out[0:3].copy(), "pdf": float(out[3]), "swi": out[4:7].copy(), "sf": out[7:10].copy(),
            "spdf": float(out[10]), "stype": int(out[11]), "ncomp": int(out[12])}
