"""bg_guide_init_params (host only): He-uniform weights U(-sqrt(6/64), sqrt(6/64))
from the counter-based uniform of (seed, parameter index), zero biases."""
import numpy as np
from scipy.stats import chi2

from bonsai_amd import _native as N
from bonsai_amd.guide import init_params

LAYERS = [(64, 64)] * 4 + [(40, 64)]


def split(p):
    out, k = [], 0
    for no, ni in LAYERS:
        w = p[k:k + no * ni]
        k += no * ni
        b = p[k:k + no]
        k += no
        out.append((w, b))
    assert k == N.BG_GUIDE_PARAMS
    return out


def test_init_params_layout_and_distribution():
    p = init_params(5)
    assert p.shape == (N.BG_GUIDE_PARAMS,) and p.dtype == np.float32
    assert np.array_equal(p, init_params(5))
    assert not np.array_equal(p, init_params(6))
    lim = np.float32(np.sqrt(np.float32(6.0) / np.float32(64.0)))
    for w, b in split(p):
        assert np.all(b == 0.0)
        assert np.all(np.abs(w) <= lim)
        # uniform on [-lim, lim]: chi-square over 16 equal cells
        h, _ = np.histogram(w, bins=16, range=(-lim, lim))
        e = w.size / 16
        assert chi2.sf(((h - e) ** 2 / e).sum(), 15) > 1e-3
    # each weight is (2 u - 1) lim with u the 24-bit counter uniform of (seed, parameter index)
    k = np.arange(N.BG_GUIDE_PARAMS)
    u = np.array([N.lib().bgx_uniform(5, int(i)) for i in k[:300]], np.float32)
    np.testing.assert_array_equal(p[:300], (np.float32(2.0) * u - np.float32(1.0)) * lim)


def test_init_params_rejects_null():
    assert N.lib().bg_guide_init_params(1, None) == N.BG_ERR_INVALID
