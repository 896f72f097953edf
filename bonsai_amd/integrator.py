"""GpuPathIntegrator: stands where the reference's PathIntegrator stands
(src/integrators/path.rs:19-40; render = SamplerIntegrator::render,
src/base/integrator.rs:35-133), rendering on the GPU into a Film whose
sums can be read back or resolved like Film::write_image (film.rs:156-202).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .gpu import Context, GpuBvh
from .scenes import Camera

MODE_REF = N.BG_MODE_REF
MODE_CORRECTED = N.BG_MODE_CORRECTED
SCHEDULE_WAVEFRONT = N.BG_SCHEDULE_WAVEFRONT
SCHEDULE_PERSISTENT = N.BG_SCHEDULE_PERSISTENT

_XYZ = np.array([[0.412453, 0.357580, 0.180423], [0.212671, 0.715160, 0.072169], [0.019334, 0.119193, 0.950227]])
_RGB = np.array([[3.240479, -1.537150, -0.498535], [-0.969256, 1.875991, 0.041556], [0.055648, -0.204043, 1.057311]])


def resolve(rgb_sum: np.ndarray, weight: np.ndarray) -> np.ndarray:
    """Film::write_image: RGB sums -> XYZ -> RGB, divided by the weight sum, clamped at 0."""
    xyz = rgb_sum @ _XYZ.T
    rgb = xyz @ _RGB.T
    w = weight[..., None]
    out = np.where(w != 0, np.maximum(rgb / np.where(w != 0, w, 1), 0), rgb)
    return out


class Film:
    def __init__(self, ctx: Context, width: int, height: int, f64: bool = False):
        self.ctx = ctx
        self.width, self.height = width, height
        self.f64 = f64
        h = C.c_void_p()
        N.check(N.lib().bg_film_create(ctx.h, width, height, N.BG_FILM_F64 if f64 else 0, C.byref(h)), ctx.h)
        self.h = h

    def close(self):
        if self.h:
            N.lib().bg_film_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def clear(self):
        N.check(N.lib().bg_film_clear(self.h), self.ctx.h)

    def sums(self, f64: bool | None = None) -> tuple[np.ndarray, np.ndarray]:
        """(rgb_sum (h, w, 3), weight_sum (h, w)); FP64 sums when the film keeps them."""
        use64 = self.f64 if f64 is None else f64
        if use64:
            a = np.empty((self.height, self.width, 4), np.float64)
            N.check(N.lib().bg_film_read(self.h, None, a.ctypes.data), self.ctx.h)
        else:
            a = np.empty((self.height, self.width, 4), np.float32)
            N.check(N.lib().bg_film_read(self.h, a.ctypes.data, None), self.ctx.h)
        return a[..., :3], a[..., 3]

    def image(self) -> np.ndarray:
        rgb, w = self.sums()
        return resolve(rgb.astype(np.float64), w.astype(np.float64))

    def load(self, rgb_sum: np.ndarray, weight: np.ndarray) -> None:
        """Replaces the film's sums (e.g. to resume from sums saved with `sums`)."""
        a = np.empty((self.height, self.width, 4), np.float32)
        a[..., :3] = rgb_sum
        a[..., 3] = weight
        N.check(N.lib().bg_film_load(self.h, a.ctypes.data), self.ctx.h)

    def write_image(self, path: str | None = None, splat_scale: float = 1.0, scale: float = 1.0) -> np.ndarray:
        """Film::write_image (film.rs:156-202) resolved on the GPU: the (h, w, 3) float32 image, also
        written to `path` as OpenEXR (Image::write, io/image.rs:47-53) when given."""
        out = np.empty((self.height, self.width, 3), np.float32)
        p = None if path is None else str(path).encode()
        N.check(N.lib().bg_film_write_image(self.h, float(splat_scale), float(scale), out.ctypes.data, p), self.ctx.h)
        return out


class GpuPathIntegrator:
    """PathIntegrator::new(camera, sampler, max_depth, rr_threshold) with an independent
    (sampled_dimensions = 0) sampler of `spp` samples per pixel."""

    def __init__(self, ctx: Context, camera: Camera, spp: int, max_depth: int, rr_threshold: float = 1.0,
                 mode: int = MODE_REF, seed: int = 1):
        self.ctx = ctx
        self.camera = camera
        self.spp = spp
        self.max_depth = max_depth
        self.rr_threshold = rr_threshold
        self.mode = mode
        self.seed = seed

    def desc(self, spp=None, first_sample=0, window=None, paths_per_pass=0) -> N.bg_render_desc:
        d = N.bg_render_desc()
        d.camera = self.camera.desc()
        d.spp = self.spp if spp is None else spp
        d.first_sample = first_sample
        d.max_depth = self.max_depth
        d.rr_threshold = self.rr_threshold
        d.mode = self.mode
        d.seed = self.seed
        if window is not None:
            d.x0, d.y0, d.x1, d.y1 = window
        d.paths_per_pass = paths_per_pass
        return d

    def render_into(self, bvh: GpuBvh, film: Film, spp=None, first_sample=0, window=None,
                    paths_per_pass=0) -> N.bg_stats:
        d = self.desc(spp, first_sample, window, paths_per_pass)
        st = N.bg_stats()
        N.check(N.lib().bg_render_film(self.ctx.h, bvh.h, C.byref(d), film.h, C.byref(st)), self.ctx.h)
        return st

    def render(self, bvh: GpuBvh) -> np.ndarray:
        """Integrator::render: the resolved (h, w, 3) image."""
        d = self.desc()
        out = np.empty((self.camera.height, self.camera.width, 3), np.float32)
        st = N.bg_stats()
        N.check(N.lib().bg_render(self.ctx.h, bvh.h, C.byref(d), out.ctypes.data, C.byref(st)), self.ctx.h)
        self.last_stats = st
        return out


UNIFORM_SAMPLE_ALL = N.BG_LIGHT_SAMPLE_ALL
UNIFORM_SAMPLE_ONE = N.BG_LIGHT_SAMPLE_ONE


class _FilmRenderer(GpuPathIntegrator):
    """A SamplerIntegrator whose render call fills a Film (render_into) and resolves it (render)."""

    def _render_film(self, bvh, d, film, st):
        raise NotImplementedError

    def render_into(self, bvh: GpuBvh, film: Film, spp=None, first_sample=0, window=None,
                    paths_per_pass=0) -> N.bg_stats:
        d = self.desc(spp, first_sample, window, paths_per_pass)
        st = N.bg_stats()
        self._render_film(bvh, d, film, st)
        return st

    def render(self, bvh: GpuBvh) -> np.ndarray:
        film = Film(self.ctx, self.camera.width, self.camera.height)
        self.last_stats = self.render_into(bvh, film)
        img = film.write_image()
        film.close()
        return img


class DirectLightingIntegrator(_FilmRenderer):
    """DirectLightingIntegrator::new(camera, sampler, scene, max_depth, strategy) (direct.rs:28-53)
    on the GPU: emission plus estimate_direct for every light (UNIFORM_SAMPLE_ALL) or one chosen
    light (UNIFORM_SAMPLE_ONE) at the camera ray's hit."""

    def __init__(self, ctx: Context, camera: Camera, spp: int, max_depth: int = 5,
                 strategy: int = UNIFORM_SAMPLE_ALL, mode: int = MODE_REF, seed: int = 1):
        super().__init__(ctx, camera, spp, max_depth, 1.0, mode, seed)
        self.direct = N.bg_direct_desc(N.BG_INTEGRATOR_DIRECT, strategy)

    def _render_film(self, bvh, d, film, st):
        N.check(N.lib().bg_render_film_direct(self.ctx.h, bvh.h, C.byref(d), C.byref(self.direct), film.h,
                                              C.byref(st)), self.ctx.h)


class WhittedIntegrator(_FilmRenderer):
    """WhittedIntegrator::new(camera, sampler, max_depth) (whitted.rs:16-24) on the GPU: emission
    plus one light-sampled, shadow-tested contribution per light at the camera ray's hit."""

    def __init__(self, ctx: Context, camera: Camera, spp: int, max_depth: int = 5, mode: int = MODE_REF,
                 seed: int = 1):
        super().__init__(ctx, camera, spp, max_depth, 1.0, mode, seed)
        self.direct = N.bg_direct_desc(N.BG_INTEGRATOR_WHITTED, 0)

    def _render_film(self, bvh, d, film, st):
        N.check(N.lib().bg_render_film_direct(self.ctx.h, bvh.h, C.byref(d), C.byref(self.direct), film.h,
                                              C.byref(st)), self.ctx.h)


class GuidedPathIntegrator(GpuPathIntegrator):
    """The corrected path integrator with a learned directional proposal: at every vertex the
    continuation is drawn from the guiding network's vMF mixture with probability ``select_prob``,
    else from the BSDF, and weighted by the one-sample MIS density, so the image converges to the
    plain integrator's for any network state. With ``train`` the network is trained online after
    every pass on that pass's own path vertices (no counterpart in the reference)."""

    def __init__(self, ctx: Context, camera: Camera, spp: int, max_depth: int, guide, rr_threshold: float = 1.0,
                 seed: int = 1, select_prob: float = 0.5, train: bool = True, train_paths: int = 1 << 20,
                 batch: int = 16384, steps: int = 8, lr: float = 1e-2, replay: int = 1 << 20, lobes: int = 0,
                 bounces: int = 0, rr_start: int = 0):
        super().__init__(ctx, camera, spp, max_depth, rr_threshold, MODE_CORRECTED, seed)
        self.guide = guide
        self.guiding = N.bg_guiding_desc(select_prob, 1 if train else 0, train_paths, batch, steps, lr, replay,
                                         lobes, bounces, rr_start)

    def render_into(self, bvh: GpuBvh, film: Film, spp=None, first_sample=0, window=None,
                    paths_per_pass=0) -> N.bg_stats:
        d = self.desc(spp, first_sample, window, paths_per_pass)
        st = N.bg_stats()
        N.check(N.lib().bg_render_film_guided(self.ctx.h, bvh.h, C.byref(d), self.guide.h, C.byref(self.guiding),
                                              film.h, C.byref(st)), self.ctx.h)
        return st

    def records(self, max_records: int = 1 << 22) -> tuple[np.ndarray, np.ndarray]:
        """(inputs (n, 12), targets (n, 4)) of the last trained pass."""
        n = C.c_uint64(0)
        N.check(N.lib().bg_guide_records(self.ctx.h, None, None, 0, C.byref(n)), self.ctx.h)
        m = min(n.value, max_records)
        inp = np.empty((m, 12), np.float32)
        tgt = np.empty((m, 4), np.float32)
        N.check(N.lib().bg_guide_records(self.ctx.h, inp.ctypes.data, tgt.ctypes.data, m, C.byref(n)), self.ctx.h)
        return inp, tgt

    def render(self, bvh: GpuBvh) -> np.ndarray:
        film = Film(self.ctx, self.camera.width, self.camera.height)
        self.last_stats = self.render_into(bvh, film)
        img = film.write_image()
        film.close()
        return img


def rel_mse(img: np.ndarray, ref: np.ndarray) -> float:
    """mean over pixels and channels of (I - R)^2 / (R^2 + 0.01), in FP64."""
    i = img.astype(np.float64)
    r = ref.astype(np.float64)
    return float(np.mean((i - r) ** 2 / (r * r + 0.01)))


class BatchRenderer:
    """Many (scene, camera, seed) items rendered by one wavefront loop with a hipGraph-captured pass
    and an autotuned traversal launch shape. Each item's film equals what GpuPathIntegrator with the
    same camera, seed, spp, depth and mode renders (no counterpart in the reference)."""

    def __init__(self, ctx: Context, items: list[tuple[GpuBvh, Camera, int]]):
        self.ctx = ctx
        self.items = items
        arr = (N.bg_batch_item * len(items))()
        for k, (bvh, cam, seed) in enumerate(items):
            arr[k].scene = bvh.h
            arr[k].camera = cam.desc()
            arr[k].seed = seed
        h = C.c_void_p()
        N.check(N.lib().bg_batch_create(ctx.h, arr, len(items), C.byref(h)), ctx.h)
        self.h = h

    def close(self):
        if self.h:
            N.lib().bg_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def render(self, spp: int, max_depth: int, mode: int = MODE_CORRECTED, rr_threshold: float = 1.0,
               use_graph: bool = True, autotune: bool = True, paths_per_pass: int = 0,
               schedule: int = 0) -> N.bg_stats:
        """schedule: SCHEDULE_WAVEFRONT (a kernel per stage of each pass) or SCHEDULE_PERSISTENT (each
        pass as one persistent path kernel whose lanes take path slots from an atomic counter)."""
        d = N.bg_batch_desc(spp, max_depth, rr_threshold, mode, paths_per_pass, 1 if use_graph else 0,
                            1 if autotune else 0, schedule)
        st = N.bg_stats()
        N.check(N.lib().bg_render_batch(self.ctx.h, self.h, C.byref(d), C.byref(st)), self.ctx.h)
        return st

    def sums(self, item: int) -> tuple[np.ndarray, np.ndarray]:
        cam = self.items[item][1]
        a = np.empty((cam.height, cam.width, 4), np.float32)
        N.check(N.lib().bg_batch_film_read(self.h, item, a.ctypes.data), self.ctx.h)
        return a[..., :3], a[..., 3]

    def launch(self) -> tuple[int, int, int, int]:
        cfg = N.bg_launch_config()
        N.check(N.lib().bg_batch_launch_config(self.h, C.byref(cfg)), self.ctx.h)
        return cfg.block, cfg.stack, cfg.blocks_per_cu, cfg.variant
