"""Film::write_image (film.rs:156-202) resolved on the GPU against the oracle's
float32 restatement, bit for bit, on planted sums (negative, NaN, infinite,
zero-weight, huge) and on rendered films; the OpenEXR file the film writes
(Image::write, io/image.rs:47-53) decoded by the oracle's reader."""
import ctypes as C
import sys
from pathlib import Path

import numpy as np
import pytest

from bonsai_amd import GpuBvh
from bonsai_amd import _native as N
from bonsai_amd.integrator import MODE_CORRECTED, MODE_REF, Film, GpuPathIntegrator
from bonsai_amd.scenes import cornell_box

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import film as F  # noqa: E402

pytestmark = pytest.mark.gpu


def same_bits(a, b):
    """Equal bit patterns, except that any NaN matches any NaN."""
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    assert a.shape == b.shape
    na, nb = np.isnan(a), np.isnan(b)
    assert np.array_equal(na, nb)
    assert np.array_equal(a[~na].view(np.uint32), b[~nb].view(np.uint32))


def planted(h, w, seed):
    rng = np.random.default_rng(seed)
    s = np.empty((h, w, 4), np.float32)
    s[..., :3] = rng.standard_normal((h, w, 3)) * rng.choice([1e-3, 1.0, 1e3, 1e30], (h, w, 1))
    s[..., 3] = rng.choice([0.0, 1.0, 3.0, 1024.0, 1e-20], (h, w))
    flat = s.reshape(-1, 4)
    flat[0] = [np.nan, 1.0, 2.0, 4.0]
    flat[1] = [np.inf, 0.0, 0.0, 2.0]
    flat[2] = [-np.inf, 1.0, 1.0, 0.0]
    flat[3] = [-0.0, -0.0, -0.0, 1.0]
    flat[4] = [3e38, 3e38, 3e38, 1.0]
    return s


@pytest.mark.parametrize("splat_scale,scale", [(1.0, 1.0), (0.5, 2.5), (1.0, -1.0), (np.inf, 1.0)])
def test_resolve_matches_oracle_on_planted_sums(ctx, splat_scale, scale):
    h, w = 37, 53
    s = planted(h, w, 11)
    film = Film(ctx, w, h)
    film.load(s[..., :3], s[..., 3])
    rgb, wt = film.sums()
    same_bits(rgb, s[..., :3])
    same_bits(wt, s[..., 3])
    img = film.write_image(splat_scale=splat_scale, scale=scale)
    same_bits(img, F.write_image(s, splat_scale, scale))
    film.close()


@pytest.mark.parametrize("mode", [MODE_REF, MODE_CORRECTED])
def test_rendered_film_resolves_and_writes_exr(ctx, tmp_path, mode):
    s = cornell_box(double_sided_light=True, width=64, height=48)
    bvh = GpuBvh(ctx, s)
    gi = GpuPathIntegrator(ctx, s.camera, spp=8, max_depth=4, mode=mode, seed=5)
    film = Film(ctx, 64, 48)
    gi.render_into(bvh, film)
    rgb, wt = film.sums()
    sums = np.concatenate([rgb, wt[..., None]], axis=-1)
    path = tmp_path / "result.exr"
    img = film.write_image(str(path))
    same_bits(img, F.write_image(sums))
    e = F.read_exr(path.read_bytes())
    assert (e["width"], e["height"]) == (64, 48)
    same_bits(np.stack([e["pixels"][c] for c in "RGB"], axis=-1), img)
    # bg_render (a fresh film, the same render, write_image(1.0)) gives the same image
    same_bits(gi.render(bvh), img)
    film.close()


def test_device_resolve_matches_host_resolve(ctx):
    from hipmem import DeviceArray

    h, w = 64, 96
    s = planted(h, w, 12)
    film = Film(ctx, w, h)
    film.load(s[..., :3], s[..., 3])
    d = DeviceArray(h * w * 3 * 4)
    try:
        ms = C.c_double(-1.0)
        N.check(N.lib().bg_film_resolve_device(film.h, 1.0, 1.0, C.c_void_p(d.ptr), C.byref(ms)), ctx.h)
        assert ms.value >= 0.0
        same_bits(d.to_numpy(np.float32, (h, w, 3)), film.write_image())
    finally:
        d.free()
        film.close()


def test_write_image_needs_a_destination(ctx):
    film = Film(ctx, 4, 4)
    assert N.lib().bg_film_write_image(film.h, 1.0, 1.0, None, None) == N.BG_ERR_INVALID
    with pytest.raises(N.BonsaiError):
        film.write_image("/nonexistent-dir/result.exr")
    film.close()
