"""Host-side mirror of the reference's hot-path trait surface over the C ABI.

``GpuBvh`` stands where ``accelerators::BVH`` stands in the reference
(src/accelerators/bvh.rs:20-23, 271-423): it is built from the scene's
primitives, reports ``bounds()``, and answers ``intersect`` (closest hit,
bvh.rs:281-352) and ``intersect_test`` (any hit, bvh.rs:354-423) — here for
whole batches of rays at once, on the GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .scenes import Camera, SceneData

HIT_DTYPE = np.dtype([("prim_id", "<u4"), ("t", "<f4"), ("b1", "<f4"), ("b2", "<f4")])
SURFACE_DTYPE = np.dtype([("p", "<f4", 3), ("t", "<f4"), ("p_error", "<f4", 3), ("n", "<f4", 3), ("ns", "<f4", 3),
                          ("dpdu", "<f4", 3), ("uv", "<f4", 2), ("pad", "<f4", 2)])


class Context:
    """A device context (one HIP stream). Raises if no GPU is present."""

    def __init__(self, device: int = 0):
        L = N.lib()
        cnt = C.c_int(0)
        N.check(L.bg_device_count(C.byref(cnt)))
        if cnt.value <= device:
            raise RuntimeError(f"no HIP device {device} (found {cnt.value}); the GPU path has no CPU fallback")
        h = C.c_void_p()
        N.check(L.bg_create(device, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            N.lib().bg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self) -> int:
        return int(N.lib().bg_stream(self.h) or 0)

    def set_launch(self, block: int, stack: int, blocks_per_cu: int, variant: int = 0, compact=None) -> None:
        """Traversal launch shape; ``compact`` (None: unchanged) sets whether path-integrator passes
        trace only the rays that can hit, compacted on the device."""
        if compact is None:
            compact = self.compaction()
        cfg = N.bg_launch_config(block, stack, blocks_per_cu, variant, 1 if compact else 0)
        N.check(N.lib().bg_set_launch_config(self.h, C.byref(cfg)), self.h)

    def set_profiling(self, count: bool = False, traversal_time: bool = False) -> None:
        """Render calls count their traversal launches' box / primitive tests (stats.nodes_visited,
        tris_tested, spheres_tested) and / or time them (stats.traverse_ms)."""
        flags = (N.BG_PROFILE_COUNT if count else 0) | (N.BG_PROFILE_TRAVERSAL_TIME if traversal_time else 0)
        N.check(N.lib().bg_set_profiling(self.h, flags), self.h)

    def compaction(self) -> bool:
        cfg = N.bg_launch_config()
        N.check(N.lib().bg_get_launch_config(self.h, C.byref(cfg)), self.h)
        return bool(cfg.compact)

    def launch(self) -> tuple[int, int, int, int]:
        cfg = N.bg_launch_config()
        N.check(N.lib().bg_get_launch_config(self.h, C.byref(cfg)), self.h)
        return cfg.block, cfg.stack, cfg.blocks_per_cu, cfg.variant

    def camera_rays(self, cam: Camera, fx: float = 0.5, fy: float = 0.5) -> np.ndarray:
        d = cam.desc()
        out = np.empty((cam.width * cam.height, 8), np.float32)
        N.check(N.lib().bg_camera_rays(self.h, C.byref(d), fx, fy, out.ctypes.data), self.h)
        return out

    def camera_rays_device(self, cam: Camera, d_rays: int, fx: float = 0.5, fy: float = 0.5) -> None:
        d = cam.desc()
        N.check(N.lib().bg_camera_rays_device(self.h, C.byref(d), fx, fy, d_rays), self.h)


class GpuBvh:
    """Drop-in for the reference BVH aggregate, answering batched queries on the GPU."""

    def __init__(self, ctx: Context, scene: SceneData):
        self.ctx = ctx
        self.scene = scene
        desc, keep = scene.desc()
        h = C.c_void_p()
        N.check(N.lib().bg_scene_create(ctx.h, C.byref(desc), C.byref(h)), ctx.h)
        del keep
        self.h = h
        info = N.bg_scene_info()
        N.check(N.lib().bg_scene_info_get(self.h, C.byref(info)))
        self.info = info

    def close(self):
        if self.h:
            N.lib().bg_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def bvh(self) -> tuple[np.ndarray, np.ndarray]:
        """Flattened nodes (n, 8) as raw 32-bit words and the slot -> primitive map."""
        nodes = np.empty((self.info.n_nodes, 8), np.uint32)
        slots = np.empty(self.info.n_prims, np.uint32)
        N.check(N.lib().bg_scene_bvh(self.h, nodes.ctypes.data, slots.ctypes.data))
        return nodes, slots

    def bounds(self) -> tuple[np.ndarray, np.ndarray]:
        """Root box (Primitive::bounds, bvh.rs:272-279)."""
        if self.info.n_nodes == 0:
            return (np.full(3, np.finfo(np.float32).max, np.float32),
                    np.full(3, np.finfo(np.float32).min, np.float32))
        n, _ = self.bvh()
        f = n[0].view(np.float32)
        return f[0:3].copy(), f[3:6].copy()

    def _query(self, rays: np.ndarray, flags: int) -> tuple[np.ndarray, N.bg_stats]:
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        hits = np.empty(rays.shape[0], HIT_DTYPE)
        st = N.bg_stats()
        N.check(N.lib().bg_intersect(self.ctx.h, self.h, rays.ctypes.data, rays.shape[0], hits.ctypes.data, flags,
                                     C.byref(st)), self.ctx.h)
        return hits, st

    def intersect(self, rays: np.ndarray) -> np.ndarray:
        """Closest hit per ray (prim_id = BG_MISS on a miss)."""
        return self._query(rays, N.BG_CLOSEST)[0]

    def surfaces(self, rays: np.ndarray, hits: np.ndarray) -> np.ndarray:
        """The surface interaction of each hit (as GeometricPrimitive::intersect fills it): point, its
        error bounds, t, geometric and shading normals, shading dpdu and uv; NaN for misses."""
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        hits = np.ascontiguousarray(hits).view(HIT_DTYPE).reshape(-1)
        if hits.shape[0] != rays.shape[0]:
            raise ValueError("one hit per ray")
        out = np.empty(rays.shape[0], SURFACE_DTYPE)
        N.check(N.lib().bg_surface_interactions(self.ctx.h, self.h, rays.ctypes.data, hits.ctypes.data, rays.shape[0],
                                                out.ctypes.data), self.ctx.h)
        return out

    def intersect_test(self, rays: np.ndarray) -> np.ndarray:
        """Any hit per ray: prim_id of the first primitive that reported a hit."""
        return self._query(rays, N.BG_ANY)[0]

    def intersect_device(self, d_rays: int, n: int, d_hits: int, any_hit: bool = False,
                         count: bool = False) -> N.bg_stats:
        """Queries on device buffers (e.g. torch tensors' data_ptr()); returns kernel timing and, with
        ``count``, the number of box and primitive tests."""
        st = N.bg_stats()
        flags = (N.BG_ANY if any_hit else N.BG_CLOSEST) | (N.BG_COUNT if count else 0)
        N.check(N.lib().bg_intersect_device(self.ctx.h, self.h, d_rays, n, d_hits, flags, C.byref(st)), self.ctx.h)
        return st

    def count(self, rays: np.ndarray, any_hit: bool = False) -> N.bg_stats:
        """Host-buffer query with box/primitive test counting."""
        return self._query(rays, (N.BG_ANY if any_hit else N.BG_CLOSEST) | N.BG_COUNT)[1]
