"""Pins of single oracle functions by what the mathematics fixes: closed forms,
invariants and distribution tests, each chosen so that a plausible one-token
mistake in the restatement fails it (tests/test_oracle_mutants.py plants such
mistakes in copies of the oracle and checks that these pins catch them).

Every pin takes a loaded oracle library (``bind(ctypes.CDLL(path))``) and raises
AssertionError when the library disagrees with the mathematics:

- conductor Fresnel (bxdf.rs:190-217): the normal-incidence closed form
  ((eta-1)^2 + k^2) / ((eta+1)^2 + k^2), exactly 1 at grazing incidence, and
  the Fresnel equations of a complex refractive index eta + ik evaluated with
  complex arithmetic at random angles;
- Trowbridge-Reitz (microfacet.rs): Lambda against its isotropic closed form
  (-1 + sqrt(1 + alpha^2 tan^2 theta)) / 2; in the corrected semantics the
  visible-normal density D G1 |wo.wh| / cos(theta_o) integrates to 1 over the
  visible microfacets (Smith's normalisation, which holds only with
  G1 = 1 / (1 + Lambda) and the exact D); the reference's G1 = Lambda (Q6) makes
  that density exactly Lambda (1 + Lambda) times the corrected one;
- sample11 / sample (microfacet.rs:55-153): at normal incidence the slopes
  invert the slope distribution's radial CDF r^2 / (1 + r^2) = u1 at angle
  2 pi u2; the reference semantics always keep the smaller root slope_x_1 (Q7),
  so their slope_x never exceeds the corrected one and equals it exactly where
  the corrected branch takes that root too; the corrected sampler's normals
  follow the visible-normal density (chi-square over 256 cells); sampling
  from -wo gives exactly -wh;
- offset_ray_origin / spawn_ray / spawn_ray_to_it (point3.rs:31-49,
  interaction.rs:115-143): the origin moves along n away from the error box
  towards the side of the direction, every moved component by at least the
  offset (rounded away from p) and by at most one ulp more, unmoved components
  stay bit-equal; t_max is f32(1 - 1e-4) (the reference) or infinity
  (corrected); spawn_ray_to_it's target is offset back towards the origin and
  the segment ends on the origin's side of the target surface's plane.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

F32 = np.float32
P = C.c_void_p
SIGNATURES = {
    "orc_fresnel_conductor": (None, [C.c_uint64, P, P, P, P]),
    "orc_tr_eval": (None, [C.c_int, C.c_float, C.c_uint64, P, P, P]),
    "orc_sample11": (None, [C.c_int, C.c_uint64, P, P, P, P]),
    "orc_tr_sample": (None, [C.c_int, C.c_float, P, C.c_uint64, P, P, P]),
    "orc_spawn_ray": (None, [C.c_int, C.c_uint64, P, P, P, P, P]),
    "orc_spawn_ray_to": (None, [C.c_uint64, P, P, P, P, P, P, P]),
}
REF, CORRECTED = 0, 1
T_MAX_SHADOW = F32(1.0) - F32(1e-4)  # 1.0 - 0.0001 in f32 (interaction.rs:120)


def bind(L: C.CDLL) -> C.CDLL:
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


def _f(a) -> np.ndarray:
    return np.ascontiguousarray(a, F32)


# ------------------------------------------------------------------ wrappers
def fresnel_conductor(L, cos_i, eta, k):
    cos_i, eta, k = _f(cos_i), _f(eta), _f(k)
    out = np.empty((cos_i.size, 3), F32)
    L.orc_fresnel_conductor(cos_i.size, cos_i.ctypes.data, eta.ctypes.data, k.ctypes.data, out.ctypes.data)
    return out


def tr_eval(L, mode, alpha, wo, wh):
    wo, wh = _f(wo).reshape(-1, 3), _f(wh).reshape(-1, 3)
    out = np.empty((wo.shape[0], 4), F32)
    L.orc_tr_eval(mode, float(alpha), wo.shape[0], wo.ctypes.data, wh.ctypes.data, out.ctypes.data)
    return out


def sample11(L, mode, cos_t, u1, u2):
    cos_t, u1, u2 = _f(cos_t), _f(u1), _f(u2)
    out = np.empty((cos_t.size, 2), F32)
    L.orc_sample11(mode, cos_t.size, cos_t.ctypes.data, u1.ctypes.data, u2.ctypes.data, out.ctypes.data)
    return out


def tr_sample(L, mode, alpha, wo, u1, u2):
    wo, u1, u2 = _f(wo), _f(u1), _f(u2)
    out = np.empty((u1.size, 3), F32)
    L.orc_tr_sample(mode, float(alpha), wo.ctypes.data, u1.size, u1.ctypes.data, u2.ctypes.data, out.ctypes.data)
    return out


def spawn_ray(L, mode, p, perr, n, d):
    p, perr, n, d = _f(p), _f(perr), _f(n), _f(d)
    out = np.empty((p.shape[0], 8), F32)
    L.orc_spawn_ray(mode, p.shape[0], p.ctypes.data, perr.ctypes.data, n.ctypes.data, d.ctypes.data, out.ctypes.data)
    return out


def spawn_ray_to(L, p0, e0, n0, p1, e1, n1):
    a = [_f(x) for x in (p0, e0, n0, p1, e1, n1)]
    out = np.empty((a[0].shape[0], 11), F32)
    L.orc_spawn_ray_to(a[0].shape[0], *[x.ctypes.data for x in a], out.ctypes.data)
    return out


# ------------------------------------------------------------------- helpers
def _dirs(theta, phi):
    return np.stack([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)], -1)


def _lambda_closed(alpha, cos_t):
    tan2 = (1.0 - cos_t * cos_t) / (cos_t * cos_t)
    return (-1.0 + np.sqrt(1.0 + alpha * alpha * tan2)) / 2.0


def _u_of(wh, alpha):
    """The coordinate u = tan^2 / (alpha^2 + tan^2) of a normal's polar angle."""
    z = wh[:, 2].astype(np.float64)
    tan2 = np.maximum(1.0 - z * z, 0.0) / (z * z)
    return tan2 / (alpha * alpha + tan2)


def _vndf_cells(L, mode, alpha, wo, nu, nphi, sub):
    """Integral of the oracle's visible-normal pdf (masked to wo.wh > 0) over nu x nphi cells of
    (u, phi), by sub x sub midpoints per cell. Within a cell the points are midpoints of
    w = 1 - sqrt(1 - u), which takes away the integrable 1 / sqrt(1 - u) growth of the density
    towards the horizon (FP64 jacobians)."""
    edges_u = np.arange(nu + 1) / nu
    edges_w = 1.0 - np.sqrt(1.0 - edges_u)
    t = (np.arange(sub) + 0.5) / sub
    w = (edges_w[:-1, None] + (edges_w[1:] - edges_w[:-1])[:, None] * t[None, :]).ravel()
    dw = np.repeat((edges_w[1:] - edges_w[:-1]) / sub, sub)
    u = 1.0 - (1.0 - w) ** 2
    dudw = 2.0 * (1.0 - w)
    kp = nphi * sub
    phi = (np.arange(kp) + 0.5) / kp * 2.0 * np.pi
    tan = alpha * np.sqrt(u / (1.0 - u))
    theta = np.arctan(tan)
    # d u / d theta for u = tan^2 / (alpha^2 + tan^2)
    dudt = 2.0 * tan * alpha * alpha / (alpha * alpha + tan * tan) ** 2 * (1.0 + tan * tan)
    jac = np.sin(theta) / dudt * dudw * dw
    TH, PH = np.meshgrid(theta, phi, indexing="ij")
    wh = _dirs(TH.ravel(), PH.ravel())
    whf = wh.astype(F32)
    wob = np.broadcast_to(_f(wo), whf.shape)
    pdf = tr_eval(L, mode, alpha, wob, whf)[:, 3].astype(np.float64)
    vis = whf.astype(np.float64) @ _f(wo).astype(np.float64) > 0.0
    val = (pdf * vis).reshape(nu * sub, kp) * jac[:, None] * (2.0 * np.pi / kp)
    return val.reshape(nu, sub, nphi, sub).sum(axis=(1, 3))


# ---------------------------------------------------------------------- pins
def pin_fresnel_conductor_normal_incidence(L):
    rng = np.random.default_rng(1)
    eta = rng.uniform(0.05, 4.0, (2000, 3)).astype(F32)
    k = rng.uniform(0.0, 6.0, (2000, 3)).astype(F32)
    got = fresnel_conductor(L, np.ones(2000), eta, k).astype(np.float64)
    e, kk = eta.astype(np.float64), k.astype(np.float64)
    want = ((e - 1) ** 2 + kk ** 2) / ((e + 1) ** 2 + kk ** 2)
    np.testing.assert_allclose(got, want, rtol=2e-5, atol=1e-7)


def pin_fresnel_conductor_grazing(L):
    rng = np.random.default_rng(2)
    eta = rng.uniform(0.05, 4.0, (500, 3))
    k = rng.uniform(0.0, 6.0, (500, 3))
    got = fresnel_conductor(L, np.zeros(500), eta, k)
    assert np.all(got == 1.0), got[got != 1.0][:5]


def pin_fresnel_conductor_complex_index(L):
    rng = np.random.default_rng(3)
    n = 5000
    theta = rng.uniform(0.0, np.pi / 2 * 0.999, n)
    cos_i = np.cos(theta).astype(F32)
    eta = rng.uniform(0.1, 4.0, (n, 3)).astype(F32)
    k = rng.uniform(0.0, 6.0, (n, 3)).astype(F32)
    got = fresnel_conductor(L, cos_i, eta, k).astype(np.float64)
    c = cos_i.astype(np.float64)[:, None]
    m = eta.astype(np.float64) + 1j * k.astype(np.float64)  # complex index of the conductor (outside: 1)
    s2 = 1.0 - c * c
    mct = np.sqrt(m * m - s2)  # m cos(theta_t), Snell's law sin(theta_t) = sin(theta_i) / m
    rs = np.abs((c - mct) / (c + mct)) ** 2
    rp = np.abs((m * m * c - mct) / (m * m * c + mct)) ** 2
    want = 0.5 * (rs + rp)
    np.testing.assert_allclose(got, want, rtol=2e-4, atol=2e-6)


def pin_tr_lambda_closed_form(L):
    rng = np.random.default_rng(4)
    for alpha in (0.05, 0.2, 0.5, 0.8):
        theta = rng.uniform(0.0, np.deg2rad(88.0), 3000)
        phi = rng.uniform(0.0, 2.0 * np.pi, 3000)
        wo = _dirs(theta, phi).astype(F32)
        got = tr_eval(L, CORRECTED, alpha, wo, np.tile([0.0, 0.0, 1.0], (3000, 1)))[:, 1].astype(np.float64)
        want = _lambda_closed(float(F32(alpha)), wo[:, 2].astype(np.float64))
        np.testing.assert_allclose(got, want, rtol=2e-4, atol=1e-7)


def pin_tr_visible_normals_normalised(L):
    for alpha in (0.15, 0.5):
        for theta_o in (0.3, 1.0, 1.4):
            wo = _dirs(np.array(theta_o), np.array(0.7)).astype(F32)
            total = _vndf_cells(L, CORRECTED, alpha, wo, 64, 32, 6).sum()
            assert abs(total - 1.0) < 2e-4, (alpha, theta_o, total)


def pin_tr_ref_pdf_ratio(L):
    """Q6: g1 = 1/(1/Lambda) = Lambda, so pdf_ref / pdf_corrected = Lambda (1 + Lambda)."""
    rng = np.random.default_rng(5)
    for alpha in (0.1, 0.4):
        theta = rng.uniform(0.05, 1.5, 4000)
        wo = _dirs(theta, rng.uniform(0, 2 * np.pi, 4000)).astype(F32)
        wh = _dirs(rng.uniform(0, 1.2, 4000), rng.uniform(0, 2 * np.pi, 4000)).astype(F32)
        ref = tr_eval(L, REF, alpha, wo, wh)
        cor = tr_eval(L, CORRECTED, alpha, wo, wh).astype(np.float64)
        lam = ref[:, 1].astype(np.float64)  # the oracle's own Lambda (pinned by its closed form)
        assert np.array_equal(ref[:, [0, 1]], cor[:, [0, 1]].astype(F32))  # D and Lambda are shared
        ok = cor[:, 3] > 1e-30
        np.testing.assert_allclose(ref[ok, 3], cor[ok, 3] * lam[ok] * (1.0 + lam[ok]), rtol=3e-6, atol=0)
        # and the corrected G1 is Smith's 1 / (1 + Lambda)
        np.testing.assert_allclose(cor[:, 2], 1.0 / (1.0 + lam), rtol=3e-7)


def pin_sample11_normal_incidence(L):
    rng = np.random.default_rng(6)
    u1 = rng.uniform(0.0, 0.999, 4000)
    u2 = rng.uniform(0.0, 1.0, 4000)
    for mode in (REF, CORRECTED):
        for c in (1.0, 0.99995):
            s = sample11(L, mode, np.full(4000, c), u1, u2).astype(np.float64)
            r2 = s[:, 0] ** 2 + s[:, 1] ** 2
            u1f = u1.astype(F32).astype(np.float64)
            np.testing.assert_allclose(r2 / (1.0 + r2), u1f, rtol=3e-5, atol=1e-6)
            ang = np.mod(np.arctan2(s[:, 1], s[:, 0]), 2 * np.pi)
            want = np.mod(2 * np.pi * u2.astype(F32).astype(np.float64), 2 * np.pi)
            dist = np.abs(np.angle(np.exp(1j * (ang - want))))
            big = r2 > 1e-8
            assert np.all(dist[big] < 2e-5), dist.max()


def pin_sample11_reference_keeps_the_smaller_root(L):
    rng = np.random.default_rng(7)
    n = 20000
    c = rng.uniform(0.05, 0.999, n)
    u1, u2 = rng.uniform(0, 1, n), rng.uniform(0, 1, n)
    ref = sample11(L, REF, c, u1, u2)
    cor = sample11(L, CORRECTED, c, u1, u2)
    fin = np.isfinite(ref[:, 0]) & np.isfinite(cor[:, 0])
    assert fin.mean() > 0.99
    assert np.all(ref[fin, 0] <= cor[fin, 0])
    same = ref[fin, 0] == cor[fin, 0]
    # both branches occur: where the corrected sampler takes slope_x_1 the two agree exactly
    assert 0.05 < same.mean() < 0.95, same.mean()
    assert np.array_equal(ref[fin][same], cor[fin][same])


def pin_tr_sample_follows_visible_normals(L):
    """Chi-square of the corrected sampler's normals against the oracle's own visible-normal pdf,
    on 16 x 16 cells of (u, phi); the reference semantics (Q7) do not follow it."""
    from scipy.stats import chi2

    rng = np.random.default_rng(8)
    nu, nphi, n = 16, 16, 100000
    for alpha, theta_o in ((0.15, 0.4), (0.5, 1.1), (0.3, 1.45)):
        wo = _dirs(np.array(theta_o), np.array(0.7)).astype(F32)
        cells = _vndf_cells(L, CORRECTED, alpha, wo, nu, nphi, 8)
        expected = cells / cells.sum() * n
        pvals = {}
        for mode in (CORRECTED, REF):
            u1, u2 = rng.uniform(0, 1, n), rng.uniform(0, 1, n)
            wh = tr_sample(L, mode, alpha, wo, u1, u2)
            ok = np.all(np.isfinite(wh), axis=1)
            wh = wh[ok].astype(np.float64)
            iu = np.minimum((_u_of(wh, float(F32(alpha))) * nu).astype(int), nu - 1)
            ip = np.minimum((np.mod(np.arctan2(wh[:, 1], wh[:, 0]), 2 * np.pi) / (2 * np.pi) * nphi).astype(int),
                            nphi - 1)
            counts = np.zeros((nu, nphi))
            np.add.at(counts, (iu, ip), 1.0)
            e = expected * (ok.sum() / n)
            use = e >= 5
            stat = float((((counts - e) ** 2) / np.maximum(e, 1e-12))[use].sum() + counts[~use].sum())
            pvals[mode] = chi2.sf(stat, int(use.sum()) - 1)
        print(f"tr sample chi-square p-values alpha {alpha} theta_o {theta_o}: {pvals}")
        assert pvals[CORRECTED] > 1e-3, (alpha, theta_o, pvals)
        assert pvals[REF] < 1e-6, (alpha, theta_o, pvals)


def pin_tr_sample_flip(L):
    rng = np.random.default_rng(9)
    u1, u2 = rng.uniform(0, 1, 2000), rng.uniform(0, 1, 2000)
    for mode in (REF, CORRECTED):
        wo = _dirs(np.array(0.9), np.array(2.0)).astype(F32)
        a = tr_sample(L, mode, 0.3, wo, u1, u2)
        b = tr_sample(L, mode, 0.3, -wo, u1, u2)
        assert np.array_equal(a, -b)


def _ulp(x):
    return np.spacing(np.abs(x).astype(F32)).astype(np.float64)


def _spawn_inputs(seed, n):
    rng = np.random.default_rng(seed)
    mag = 10.0 ** rng.uniform(-2, 2.5, (n, 3))
    p = (mag * rng.choice([-1.0, 1.0], (n, 3))).astype(F32)
    # error bounds of a few ulps of |p| (gamma(n) |p| sized), some exactly 0
    perr = (np.abs(p.astype(np.float64)) * rng.uniform(0, 6, (n, 3)) * 2.0 ** -23).astype(F32)
    perr[rng.uniform(size=(n, 3)) < 0.1] = 0.0
    nrm = rng.standard_normal((n, 3))
    nrm[: n // 5, rng.integers(0, 3)] = 0.0  # components exactly zero: those coordinates stay put
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    nrm = nrm.astype(F32)
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    cosd = np.einsum("ij,ij->i", d, nrm.astype(np.float64))
    keep = np.abs(cosd) > 0.1
    return p[keep], perr[keep], nrm[keep], d[keep].astype(F32)


def _check_offset(p, perr, n, w, o):
    """o = offset_ray_origin(p, perr, n, w): each coordinate moved along s n by at least
    |(|n|.perr) n_i| and at most one ulp more; unmoved where that offset is 0."""
    p64, e64, n64, o64 = (x.astype(np.float64) for x in (p, perr, n, o))
    s = np.where(np.einsum("ij,ij->i", w.astype(np.float64), n64) < 0.0, -1.0, 1.0)
    dmag = np.einsum("ij,ij->i", np.abs(n64), e64)
    off = (s * dmag)[:, None] * n64
    diff = o64 - p64
    zero = off == 0.0
    assert np.array_equal(o[zero], p[zero]), "a coordinate with zero offset moved"
    nz = ~zero
    assert np.all(np.sign(diff[nz]) == np.sign(off[nz])), "moved the wrong way"
    assert np.all(np.abs(diff[nz]) >= np.abs(off[nz]) * (1.0 - 4e-7)), "moved by less than the offset"
    bound = np.abs(off) * (1.0 + 4e-7) + 1.5 * _ulp(np.abs(p64) + np.abs(off))
    assert np.all(np.abs(diff[nz]) <= bound[nz]), "moved by more than the offset and one ulp"
    side = np.einsum("ij,ij->i", diff, n64) * s
    assert np.all(side[dmag > 0] > 0.0), "origin not on the direction's side of the surface"


def pin_spawn_ray(L):
    p, perr, n, d = _spawn_inputs(10, 6000)
    for mode, tmax in ((REF, T_MAX_SHADOW), (CORRECTED, F32(np.inf))):
        r = spawn_ray(L, mode, p, perr, n, d)
        _check_offset(p, perr, n, d, r[:, 0:3])
        assert np.array_equal(r[:, 4:7], d), "direction changed"
        assert np.all(r[:, 3] == tmax), (mode, r[:5, 3])
        assert np.all(r[:, 7] == F32(0.25))


def pin_spawn_ray_to(L):
    rng = np.random.default_rng(11)
    p0, e0, n0, _ = _spawn_inputs(12, 6000)
    m = p0.shape[0]
    dirn = rng.standard_normal((m, 3))
    dirn /= np.linalg.norm(dirn, axis=1, keepdims=True)
    p1 = (p0.astype(np.float64) + dirn * rng.uniform(0.5, 3.0, (m, 1)) * (1.0 + np.abs(p0))).astype(F32)
    e1 = (np.abs(p1.astype(np.float64)) * rng.uniform(0, 6, (m, 3)) * 2.0 ** -23).astype(F32)
    n1 = rng.standard_normal((m, 3))
    n1 /= np.linalg.norm(n1, axis=1, keepdims=True)
    n1 = n1.astype(F32)
    w0 = (p1.astype(np.float64) - p0.astype(np.float64))
    c0 = np.einsum("ij,ij->i", w0 / np.linalg.norm(w0, axis=1, keepdims=True), n0.astype(np.float64))
    c1 = np.einsum("ij,ij->i", -w0 / np.linalg.norm(w0, axis=1, keepdims=True), n1.astype(np.float64))
    k = (np.abs(c0) > 0.1) & (np.abs(c1) > 0.1)
    p0, e0, n0, p1, e1, n1 = p0[k], e0[k], n0[k], p1[k], e1[k], n1[k]
    r = spawn_ray_to(L, p0, e0, n0, p1, e1, n1)
    o, d, t = r[:, 0:3], r[:, 4:7], r[:, 8:11]
    _check_offset(p0, e0, n0, (p1.astype(np.float64) - p0.astype(np.float64)).astype(F32), o)
    # the target is p1 offset towards the origin (the direction o - p1, rounded to f32 as the reference does)
    _check_offset(p1, e1, n1, (o.astype(np.float64) - p1.astype(np.float64)).astype(F32), t)
    assert np.array_equal(d, (t.astype(np.float64) - o.astype(np.float64)).astype(F32)), "d != target - origin"
    assert np.all(r[:, 3] == T_MAX_SHADOW)
    assert np.all(r[:, 7] == F32(0.25))
    # the segment origin + t_max d ends short of the target surface, on the origin's side of its plane
    so = np.einsum("ij,ij->i", o.astype(np.float64) - p1.astype(np.float64), n1.astype(np.float64))
    end = o.astype(np.float64) + float(T_MAX_SHADOW) * d.astype(np.float64)
    se = np.einsum("ij,ij->i", end - p1.astype(np.float64), n1.astype(np.float64))
    assert np.all(np.sign(se) == np.sign(so)), "the shadow segment crosses the target surface"


PINS = {name: fn for name, fn in list(globals().items()) if name.startswith("pin_")}


def failing_pins(lib_path: str) -> list[str]:
    """Names of the pins a library at lib_path fails."""
    L = bind(C.CDLL(lib_path))
    failed = []
    for name, fn in PINS.items():
        try:
            fn(L)
        except AssertionError:
            failed.append(name)
    return failed


if __name__ == "__main__":
    import json
    import sys

    print(json.dumps(failing_pins(sys.argv[1])))
