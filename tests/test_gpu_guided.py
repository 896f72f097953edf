"""Guided path integration on the GPU: the one-sample MIS estimator stays
unbiased for any network state (furnace closed form, agreement with the plain
integrator), select_prob 0 never reads the network, and the online-trained
network lowers the error at equal samples on the door-gap interior."""
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


@pytest.mark.parametrize("select_prob,train,bounces", [(0.5, True, 0), (1.0, False, 0), (1.0, True, 2)])
def test_guided_furnace_closed_form(ctx, select_prob, train, bounces):
    """Unbiased for any network state, whether every vertex is guided or only the first `bounces`."""
    s = furnace(res=32)
    bvh = GpuBvh(ctx, s)
    for depth in (1, 5):
        gi = GuidedPathIntegrator(ctx, s.camera, spp=64, max_depth=depth, guide=new_guide(ctx), select_prob=select_prob,
                                  train=train, train_paths=4096, bounces=bounces)
        film = Film(ctx, 32, 32)
        for k in range(8):  # 8 passes of 8 spp: the network changes between passes when training
            gi.render_into(bvh, film, spp=8, first_sample=8 * k)
        img = film.image()
        exact = sum(0.5 ** k for k in range(depth + 1))
        assert abs(img.mean() - exact) < 0.02 * exact, (depth, img.mean(), exact)


def test_select_prob_zero_ignores_the_network(ctx):
    """With select_prob 0 the guided estimator (light sampling + MIS-weighted continuation emission)
    never reads the network: different parameters and training on or off give the same film."""
    s = cornell_box(width=64, height=64)
    bvh = GpuBvh(ctx, s)
    films = []
    for seed, train in ((3, True), (9, False)):
        gi = GuidedPathIntegrator(ctx, s.camera, spp=8, max_depth=5, guide=new_guide(ctx, seed=seed), seed=4,
                                  select_prob=0.0, train=train, train_paths=1024)
        f = Film(ctx, 64, 64)
        for k in range(4):
            st = gi.render_into(bvh, f, spp=2, first_sample=2 * k)
            assert (st.train_steps > 0) == train
        films.append(f.sums())
    np.testing.assert_array_equal(films[0][0], films[1][0])
    np.testing.assert_array_equal(films[0][1], films[1][1])


def test_roulette_start(ctx):
    """rr_start sets the first bounce Russian roulette may end a guided path at: 0 is the plain
    path's bounce 4 (the same film as 4), and a start past the last bounce never draws, which is the
    film of no roulette; a start of 2 ends paths earlier (fewer rays, another film)."""
    s = cornell_box(width=64, height=64)
    bvh = GpuBvh(ctx, s)
    g = new_guide(ctx)

    def film(rr, start):
        gi = GuidedPathIntegrator(ctx, s.camera, spp=8, max_depth=6, guide=g, seed=4, train=False,
                                  rr_threshold=rr, rr_start=start)
        f = Film(ctx, 64, 64)
        st = gi.render_into(bvh, f)
        out = f.sums() + (st.rays,)
        f.close()
        return out

    a, b = film(1.0, 0), film(1.0, 4)
    np.testing.assert_array_equal(a[0], b[0])
    assert a[2] == b[2]
    c, d = film(1.0, 7), film(0.0, 0)
    np.testing.assert_array_equal(c[0], d[0])
    assert c[2] == d[2]
    e = film(1.0, 2)
    assert e[2] < a[2] < d[2]


@pytest.mark.parametrize("with_plastic", [False, True])
def test_guided_agrees_with_plain_on_cornell(ctx, with_plastic):
    s = cornell_box(width=48, height=48, with_plastic=with_plastic)
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
    """Equal samples per pixel on the door-gap interior (480 x 270, depth 10): the estimator without
    the network (select_prob 0) against the same estimator sampling the network's mixture with
    probability 0.5 while the network trains online, both scored against an 8192-sample render.
    Every kernel is deterministic, so the comparison is reproducible."""
    W, H, depth, spp = 480, 270, 10, 1024
    s = door_gap_room(W, H)
    bvh = GpuBvh(ctx, s)
    ref = GpuPathIntegrator(ctx, s.camera, spp=8192, max_depth=depth, mode=MODE_CORRECTED, seed=21).render(bvh)
    errs = {}
    for sel in (0.0, 0.5):
        gi = GuidedPathIntegrator(ctx, s.camera, spp=spp, max_depth=depth, guide=new_guide(ctx), seed=22,
                                  select_prob=sel, steps=16, train=sel > 0)
        film = Film(ctx, W, H)
        for k in range(0, spp, 16):
            gi.render_into(bvh, film, spp=16, first_sample=k)
        img = film.image()
        assert np.isfinite(img).all()
        errs[sel] = rel_mse(img, ref)
    print(f"door gap {W}x{H} {spp} spp relMSE: without network {errs[0.0]:.4f}, guided {errs[0.5]:.4f} "
          f"(ratio {errs[0.0] / errs[0.5]:.2f})")
    assert errs[0.5] < 0.9 * errs[0.0]


def paired_rel_var(a, b, r=3):
    """Per-sample-count relative variance from two independent renders: per pixel (A - B)^2 / 2 over
    the squared (2r+1)^2-box mean of both renders + 0.01 (unbiased for the variance of one render)."""
    m = 0.5 * (a + b).astype(np.float64)
    for axis in (0, 1):
        c = np.cumsum(np.pad(m, [(r + 1, r) if k == axis else (0, 0) for k in range(m.ndim)], mode="edge"), axis=axis)
        n = m.shape[axis]
        m = (np.take(c, np.arange(2 * r + 1, 2 * r + 1 + n), axis=axis) - np.take(c, np.arange(n), axis=axis)) / (2 * r + 1)
    return float(np.mean(0.5 * (a - b) ** 2 / (m * m + 0.01)))


def test_trained_guide_cuts_variance_without_roulette(ctx):
    """The tuned regime on the door-gap interior (480 x 270, depth 10, no Russian roulette, the
    network guiding the first 4 vertices): after a warm-up render that trains the network, the
    frozen-network estimator's per-sample variance (two independent renders, equal samples) is at
    least 2x below the same estimator without the network, and below the plain path's with the
    reference's default roulette. Deterministic kernels: the figures are reproducible."""
    W, H, depth, spp = 480, 270, 10, 64
    s = door_gap_room(W, H)
    bvh = GpuBvh(ctx, s)
    g = new_guide(ctx)

    def pair(make):
        imgs = []
        for seed in (41, 42):
            film = Film(ctx, W, H)
            gi = make(seed)
            for k in range(0, spp, 16):
                gi.render_into(bvh, film, spp=16, first_sample=k)
            imgs.append(film.image())
            film.close()
        assert all(np.isfinite(i).all() for i in imgs)
        return paired_rel_var(*imgs) * spp

    def guided(sel, train):
        return lambda seed: GuidedPathIntegrator(ctx, s.camera, spp=spp, max_depth=depth, guide=g, seed=seed,
                                                 select_prob=sel, steps=32, train=train, bounces=4, rr_threshold=0.0)

    v_plain = pair(lambda seed: GpuPathIntegrator(ctx, s.camera, spp=spp, max_depth=depth, mode=MODE_CORRECTED,
                                                  seed=seed))
    v_free = pair(guided(0.0, False))
    warm = GuidedPathIntegrator(ctx, s.camera, spp=2 * spp, max_depth=depth, guide=g, seed=40, select_prob=0.5,
                                steps=32, train=True, bounces=4, rr_threshold=0.0)
    film = Film(ctx, W, H)
    for k in range(0, 2 * spp, 16):
        st = warm.render_into(bvh, film, spp=16, first_sample=(1 << 20) + k)
        assert st.train_steps > 0
    film.close()
    v_guided = pair(guided(0.5, False))
    print(f"door gap {W}x{H}: per-sample relative variance plain (default roulette) {v_plain:.1f}, "
          f"without network {v_free:.1f}, guided {v_guided:.1f} (ratio {v_free / v_guided:.2f})")
    assert v_guided * 2.0 < v_free
    assert v_guided < v_plain


def test_growing_path_buffers_keeps_the_guiding_buffers(ctx):
    """Growing the path-state buffers (a plain render with more path slots than any call before)
    keeps the guiding buffers a guided render allocated: the next guided render of the same size
    allocates nothing more on the device."""
    from hipmem import free_bytes
    s = cornell_box(width=64, height=64)
    bvh = GpuBvh(ctx, s)
    guide = new_guide(ctx)
    gi = GuidedPathIntegrator(ctx, s.camera, spp=4, max_depth=4, guide=guide, train=True, train_paths=4096)
    film = Film(ctx, 64, 64)
    gi.render_into(bvh, film)
    # more path slots than any render before: 768 x 768 x 64 spp passes hold the 16M-slot default
    big = cornell_box(width=768, height=768)
    GpuPathIntegrator(ctx, big.camera, spp=64, max_depth=1, mode=MODE_CORRECTED).render_into(
        GpuBvh(ctx, big), Film(ctx, 768, 768), spp=64)
    before = free_bytes()
    gi.render_into(bvh, film, first_sample=4)
    after = free_bytes()
    assert after == before, (before, after)


def test_snapshot_and_restore_resume_bit_for_bit(ctx):
    """A guided render trained for two calls, snapshotted, then continued for two more, equals a
    fresh guide restored from the snapshot and given the same two calls: film sums, parameters
    and Adam moments bit for bit."""
    from bonsai_amd.guide import GuideMLP as GM
    s = cornell_box(width=64, height=64)
    bvh = GpuBvh(ctx, s)

    def integ(g):
        return GuidedPathIntegrator(ctx, s.camera, spp=2, max_depth=5, guide=g, train=True, train_paths=8192,
                                    steps=4, batch=4096)

    g1 = GM(ctx, 7, 65536)
    warm = Film(ctx, 64, 64)
    for k in range(2):
        integ(g1).render_into(bvh, warm, spp=2, first_sample=2 * k)
    snap = g1.snapshot()

    def continue_from(g):
        film = Film(ctx, 64, 64)
        for k in range(2, 4):
            integ(g).render_into(bvh, film, spp=2, first_sample=2 * k)
        return film.sums(), g.params()

    (a_rgb, a_w), a_par = continue_from(g1)
    g2 = GM(ctx, 99, 65536)
    g2.restore(snap)
    (b_rgb, b_w), b_par = continue_from(g2)
    np.testing.assert_array_equal(b_rgb, a_rgb)
    np.testing.assert_array_equal(b_w, a_w)
    for x, y in zip(a_par, b_par):
        np.testing.assert_array_equal(x, y)
    # without the snapshot the same calls give another film
    g3 = GM(ctx, 99, 65536)
    (c_rgb, _), _ = continue_from(g3)
    assert not np.array_equal(c_rgb, a_rgb)
    with pytest.raises(Exception):
        g3.restore(b"not a snapshot" * 10)
