"""Guided path integration on the GPU: the one-sample MIS estimator stays
unbiased for any network state (furnace closed form, agreement with the plain
integrator), select_prob 0 leaves the plain integrator's sample streams
untouched, and the online-trained network lowers the error at equal samples on
the door-gap interior."""
import sys
from pathlib import Path

import numpy as np
import pytest

from bonsai_amd import GpuBvh
from bonsai_amd.guide import GuideMLP
from bonsai_amd.integrator import MODE_CORRECTED, Film, GpuPathIntegrator, GuidedPathIntegrator, rel_mse
from bonsai_amd.scenes import cornell_box, door_gap_room
from test_oracle_integrator import furnace

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import guide_mlp as G  # noqa: E402

pytestmark = pytest.mark.gpu


def new_guide(ctx, seed=3, cap=65536):
    return GuideMLP(ctx, G.init_params(seed).astype(np.float32), cap)


@pytest.mark.parametrize("select_prob,train", [(0.5, True), (1.0, False)])
def test_guided_furnace_closed_form(ctx, select_prob, train):
    s = furnace(res=32)
    bvh = GpuBvh(ctx, s)
    for depth in (1, 5):
        gi = GuidedPathIntegrator(ctx, s.camera, spp=64, max_depth=depth, guide=new_guide(ctx), select_prob=select_prob,
                                  train=train, train_paths=4096)
        film = Film(ctx, 32, 32)
        for k in range(8):  # 8 passes of 8 spp: the network changes between passes when training
            gi.render_into(bvh, film, spp=8, first_sample=8 * k)
        img = film.image()
        exact = sum(0.5 ** k for k in range(depth + 1))
        assert abs(img.mean() - exact) < 0.02 * exact, (depth, img.mean(), exact)


def test_select_prob_zero_matches_plain_streams(ctx):
    s = cornell_box(width=64, height=64)
    bvh = GpuBvh(ctx, s)
    plain = GpuPathIntegrator(ctx, s.camera, spp=8, max_depth=5, mode=MODE_CORRECTED, seed=4)
    a = Film(ctx, 64, 64)
    plain.render_into(bvh, a)
    gi = GuidedPathIntegrator(ctx, s.camera, spp=8, max_depth=5, guide=new_guide(ctx), seed=4, select_prob=0.0,
                              train=True, train_paths=1024)
    b = Film(ctx, 64, 64)
    st = gi.render_into(bvh, b)
    assert st.train_steps > 0
    ra, wa = a.sums()
    rb, wb = b.sums()
    np.testing.assert_array_equal(wa, wb)
    np.testing.assert_array_equal(ra, rb)


def test_guided_agrees_with_plain_on_cornell(ctx):
    s = cornell_box(width=48, height=48)
    bvh = GpuBvh(ctx, s)
    ref = GpuPathIntegrator(ctx, s.camera, spp=4096, max_depth=5, mode=MODE_CORRECTED, seed=11).render(bvh)
    gi = GuidedPathIntegrator(ctx, s.camera, spp=1024, max_depth=5, guide=new_guide(ctx), seed=12,
                              train_paths=8192)
    film = Film(ctx, 48, 48)
    for k in range(16):
        gi.render_into(bvh, film, spp=64, first_sample=64 * k)
    img = film.image()
    plain = GpuPathIntegrator(ctx, s.camera, spp=1024, max_depth=5, mode=MODE_CORRECTED, seed=13).render(bvh)
    e_g, e_p = rel_mse(img, ref), rel_mse(plain, ref)
    print(f"cornell 1024 spp relMSE vs 4096-spp plain: guided {e_g:.2e}, plain {e_p:.2e}")
    assert abs(img.mean() / ref.mean() - 1) < 0.01
    assert e_g < 3 * e_p + 1e-4


def test_trained_guide_lowers_error_on_door_gap(ctx):
    """Equal samples per pixel: plain vs guided (trained online during the same render) against a
    high-sample plain reference on the door-gap interior."""
    W, H, depth = 96, 54, 10
    s = door_gap_room(W, H, cell=0.25)
    bvh = GpuBvh(ctx, s)
    ref = GpuPathIntegrator(ctx, s.camera, spp=8192, max_depth=depth, mode=MODE_CORRECTED, seed=21).render(bvh)
    spp, passes = 256, 16
    plain = GpuPathIntegrator(ctx, s.camera, spp=spp, max_depth=depth, mode=MODE_CORRECTED, seed=22).render(bvh)
    gi = GuidedPathIntegrator(ctx, s.camera, spp=spp, max_depth=depth, guide=new_guide(ctx), seed=22)
    film = Film(ctx, W, H)
    per = spp // passes
    for k in range(passes):
        gi.render_into(bvh, film, spp=per, first_sample=per * k)
    img = film.image()
    e_p, e_g = rel_mse(plain, ref), rel_mse(img, ref)
    print(f"door gap {W}x{H} {spp} spp relMSE: plain {e_p:.4f}, guided {e_g:.4f} (ratio {e_p / e_g:.2f})")
    assert np.isfinite(img).all()
    assert e_g < 0.8 * e_p
