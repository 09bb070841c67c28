"""Pins of the oracle that the reference's own golden vectors leave open: known-answer
tests, closed forms, invariants and brute force."""
import math

import numpy as np
import pytest
import scipy.special as sp

import oracle as O


def test_philox_known_answers():
    # Random123 kat_vectors for philox4x32-10
    assert O.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert O.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert O.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_ndtri_against_scipy():
    ps = np.concatenate([np.logspace(-300, -1, 400), np.linspace(0.01, 0.99, 301), 1 - np.logspace(-16, -2, 50)])
    for p in ps:
        ref = sp.ndtri(p)
        assert O.ndtri(p) == pytest.approx(ref, rel=2e-15, abs=1e-300)


def test_trunc_normal_inverts_cdf():
    rng = np.random.default_rng(1)
    for _ in range(2000):
        mu = rng.normal(0, 5)
        sg = math.exp(rng.uniform(-4, 2))
        lo, hi = 0.0, 10.0
        u = rng.uniform()
        x = O.trunc_normal(mu, sg, lo, hi, u)
        assert lo <= x <= hi
        a, b, z = (lo - mu) / sg, (hi - mu) / sg, (x - mu) / sg
        if a > 0:   # upper tail in log space: u = (1 - Q(z)/Q(a)) / (1 - Q(b)/Q(a))
            la = sp.log_ndtr(-a)
            mass, frac = -math.expm1(sp.log_ndtr(-b) - la), -math.expm1(sp.log_ndtr(-z) - la)
        else:
            Fa, Fb, Fz = sp.ndtr(a), sp.ndtr(b), sp.ndtr(z)
            mass, frac = Fb - Fa, (Fz - Fa)
        if mass > 1e-300:
            assert frac / mass == pytest.approx(u, abs=1e-9)


def test_trunc_normal_far_tails_stay_inside_and_ordered():
    xs = [O.trunc_normal(-200.0, 1.0, 0.0, 10.0, u) for u in (0.1, 0.5, 0.9)]
    assert all(0.0 <= x <= 10.0 for x in xs) and xs[0] <= xs[1] <= xs[2]
    # mirror: upper-bound tail
    xs = [O.trunc_normal(300.0, 2.0, 0.0, 10.0, u) for u in (0.1, 0.5, 0.9)]
    assert all(0.0 <= x <= 10.0 for x in xs) and xs[0] <= xs[1] <= xs[2]


def test_resample_closed_form_invariants():
    rng = np.random.default_rng(3)
    for P in (1, 7, 100, 2000):
        w = rng.exponential(size=P) * (rng.uniform(size=P) > 0.4)
        if w.sum() == 0:
            w[0] = 1
        w = O.normalize_quad(w)
        uh = rng.uniform()
        a = O.resample_quad(w, uh, 0.0)
        assert a.sum() == P
        assert np.all(a >= np.floor(P * w) - 1e-9) and np.all(a <= np.floor(P * w) + 1)
        # prefix-sum closed form A_i = floor(P (c_i - u0)) + 1
        u0 = max(uh, 1e-5) / P
        A = np.floor(P * (np.cumsum(w) - u0)) + 1
        np.testing.assert_array_equal(np.cumsum(a)[:-1], A[:-1])


def test_find_next_gamma_bruteforce():
    rng = np.random.default_rng(5)
    L = rng.uniform(10, 5000, 400)
    g, w, ev = O.find_next_gamma(0.0, L)
    lmin = L.min()
    ww = np.exp(-g * (L - lmin))
    assert abs(ww.std() / ww.mean() - 2.0) <= 1e-10
    assert ev == pytest.approx(-math.log(ww.mean()) + g * lmin, rel=1e-13)
    # a slightly larger gamma has a larger CV (monotone), so bisection found the crossing
    w2 = np.exp(-(g * (1 + 1e-6)) * (L - lmin))
    assert w2.std() / w2.mean() > 2.0


def small_problem(ndim_seed=0):
    f = O.Fault([4], [3], -9.0, 11.0, -12.0)
    theta = [-15.0, 19.5, 38.0, -1.5, 0.0, 0.2]
    lap = f.laplacian(theta)
    return f, theta, lap


def test_gradient_matches_finite_difference():
    """smc_slip.f90:829-845 is the reference's own (commented) central-difference check."""
    f, theta, lap = small_problem()
    nd = f.ndim
    rng = np.random.default_rng(7)
    A = rng.normal(size=(nd, nd))
    H = A @ A.T
    b = rng.normal(size=nd)
    a2 = np.full(2 * f.nnode, math.exp(theta[3]))
    s = rng.uniform(0.5, 2, nd)
    gam = 0.37
    g = O.grad(s, f.nnode, f.ndof, lap["li"], lap["lv"], lap["lti"], lap["ltv"], gam, a2, b, H)

    def post(x):
        return gam * O.likelihood(x, H, b, 0.0, 0, 0, 0.0, 0.0) + O.prior(x, a2, lap["li"], lap["lv"], f.nnode)
    h = 1e-6
    for i in range(nd):
        e = np.zeros(nd)
        e[i] = h
        fd = (post(s + e) - post(s - e)) / (2 * h)
        assert g[i] == pytest.approx(fd, rel=1e-6, abs=1e-6)


def test_prior_equals_quadratic_form_of_LL():
    f, theta, lap = small_problem()
    rng = np.random.default_rng(9)
    a2 = np.full(2 * f.nnode, 0.3)
    s = rng.uniform(0, 3, f.ndim)
    p = O.prior(s, a2, lap["li"], lap["lv"], f.nnode)
    assert p == pytest.approx(s @ lap["ll"] @ s / (2 * 0.3), rel=1e-12)
    np.testing.assert_array_equal(lap["ll"], lap["ll"].T)


def test_gram_matches_dense_formula():
    rng = np.random.default_rng(11)
    nobs, nd = 30, 8
    G = rng.normal(size=(nobs, nd))
    d = rng.normal(size=nobs)
    sig = rng.uniform(0.3, 1.0, nobs)
    H, b, dsd = O.gsd_gsg(G, d, sig, 1.0, 1.7)
    W = 1.0 / (sig ** 2 * 1.7)
    np.testing.assert_allclose(H, G.T @ (W[:, None] * G), rtol=1e-13)
    np.testing.assert_allclose(b, G.T @ (W * d), rtol=1e-13)
    assert dsd == pytest.approx(d @ (W * d), rel=1e-13)
    # Gram-form likelihood = residual form
    s = rng.uniform(0, 2, nd)
    L = O.likelihood(s, H, b, dsd, 0, nobs // 3, 0.0, 0.4)
    r = G @ s - d
    assert L == pytest.approx(0.5 * r @ (W * r) + 3 * (nobs // 3) * 0.4 / 2, rel=1e-10)


def test_greens_linearity_and_point_source_limit():
    """G s equals the superposed point sources of calc_responce (linearity in the unit slips)."""
    f = O.Fault([3], [2], -10.0, 10.0, -14.142)
    theta = [-17.3, 17.1, 45.0, -1.0, 0.2, 0.1]
    pts = np.array([[5.0, 5.0, 5.0], [-3.0, -3.0, -3.0]])
    uv = np.eye(3)
    G = f.greens(theta, pts.T.copy(order="F").T if False else np.array([[5.0, 5.0, 5.0], [-3.0, -3.0, -3.0]]),
                 uv, 0, 1)
    assert G.shape == (3, f.ndim)
    # far field decays: displacement from a station 1000 km away is ~1e-6 of a near one
    Gf = f.greens(theta, np.array([[1000.0] * 3, [0.0] * 3]), uv, 0, 1)
    assert np.abs(Gf).max() < 1e-3 * np.abs(G).max()


def test_dc3d0_surface_identity_and_linearity():
    """At z = 0 the real and image UA0 terms cancel; displacement is linear in potency."""
    u1, _ = O.dc3d0(2 / 3, 3.0, -2.0, 0.0, 4.0, 50.0, 1.0, 0.0, 0.0, 0.0)
    u2, _ = O.dc3d0(2 / 3, 3.0, -2.0, 0.0, 4.0, 50.0, 0.0, 1.0, 0.0, 0.0)
    u3, _ = O.dc3d0(2 / 3, 3.0, -2.0, 0.0, 4.0, 50.0, 2.0, -3.0, 0.0, 0.0)
    np.testing.assert_allclose(u3, 2 * u1 - 3 * u2, rtol=1e-13, atol=1e-16)
    # strike-slip on a vertical fault: uz is odd in the along-strike coordinate x
    a, _ = O.dc3d0(2 / 3, 2.5, 1.0, 0.0, 3.0, 90.0, 1.0, 0.0, 0.0, 0.0)
    b_, _ = O.dc3d0(2 / 3, -2.5, 1.0, 0.0, 3.0, 90.0, 1.0, 0.0, 0.0, 0.0)
    assert a[2] == pytest.approx(-b_[2], rel=1e-12)
    # far field ~ 1/R^2
    f1, _ = O.dc3d0(2 / 3, 100.0, 0.0, 0.0, 5.0, 45.0, 0.0, 1.0, 0.0, 0.0)
    f2, _ = O.dc3d0(2 / 3, 200.0, 0.0, 0.0, 10.0, 45.0, 0.0, 1.0, 0.0, 0.0)
    assert np.abs(f2[:3]).max() == pytest.approx(np.abs(f1[:3]).max() / 4, rel=1e-12)


def _tn_fraction(a, b, z):
    """(Phi(z) - Phi(a)) / (Phi(b) - Phi(a)) evaluated stably in every regime."""
    if a >= 0:  # upper tail, in log Q
        la = sp.log_ndtr(-a)
        return -math.expm1(sp.log_ndtr(-z) - la) / -math.expm1(sp.log_ndtr(-b) - la)
    if b <= 0:  # lower tail, in log Phi
        lb = sp.log_ndtr(b)
        ea, ez = math.exp(sp.log_ndtr(a) - lb), math.exp(sp.log_ndtr(z) - lb)
        return (ez - ea) / (1.0 - ea)
    return (sp.ndtr(z) - sp.ndtr(a)) / (sp.ndtr(b) - sp.ndtr(a))


@pytest.mark.parametrize("a,b", [(-12.0, 0.5), (-10.5, 3.0), (-9.5, 9.5), (0.5, 12.0), (9.0, 11.0), (12.0, 12.5),
                                 (-12.5, -12.0), (-11.0, -9.0), (-3.0, 40.0), (20.0, 60.0), (-60.0, -36.0)])
def test_trunc_normal_prec_regimes(a, b):
    """The Gibbs draw (one erfc per bound, skipped below 1e-20 of the mass) inverts the CDF
    in every regime, including the skip thresholds and the Mills-ratio tails."""
    for u in (1e-15, 1e-6, 0.01, 0.3, 0.5, 0.7, 0.99, 1 - 1e-6, 1 - 1e-15):
        s = 2.0
        mu = 1.0
        lo, hi = mu + a / s, mu + b / s
        x = O.trunc_normal_prec(mu, s, lo, hi, u)
        assert lo <= x <= hi
        z = (x - mu) * s
        assert _tn_fraction(a, b, z) == pytest.approx(u, rel=1e-6, abs=1e-12), (a, b, u, z)
    xs = [O.trunc_normal_prec(0.0, 1.0, a, b, u) for u in (0.1, 0.5, 0.9)]
    assert xs[0] <= xs[1] <= xs[2]


def test_ndtri7_matches_scipy():
    """The single-precision inverse normal CDF of the mixed Gibbs draw against scipy's ndtri over the lower
    half: Giles' erfinv polynomial (p > 0.0017, both sides of the branch point) and AS 241 PPND7's tail
    branches (r <= 5 and r > 5): error below 1e-6 of max(1, |z|) (both approximations are accurate to about
    1.3e-7; the rest is single-precision evaluation).  Any one coefficient off by a tenth shows up as an error
    of 2e-6 or more."""
    ps = np.concatenate([np.logspace(-37, -1, 400), np.linspace(0.001, 0.08, 400),
                         np.linspace(0.075, 0.5, 200)]).astype(np.float32)
    for p in ps:
        z = O.ndtri7_lower(float(p))
        ref = sp.ndtri(float(p))
        assert z <= 0.0
        assert abs(z - ref) <= 1e-6 * max(1.0, abs(ref)), (p, z, ref)


def test_erfc_cheb32_matches_scipy():
    """The single-precision erfc of the mixed Gibbs draw (Numerical Recipes' Chebyshev-fitted erfcc)
    against scipy's erfc from 0 to 9 (erfc(9) = 4e-37, still a normal float): relative error below
    4e-7 + 1.5e-7 x^2 (the fit's 1.2e-7, single-precision rounding, and the rounding of the exponent
    argument -x^2, which grows like x^2 2^-24).  A wrong coefficient shows up as an error of 1e-4 or more."""
    xs = np.concatenate([np.linspace(0.0, 2.0, 401), np.linspace(2.0, 9.0, 701)]).astype(np.float32)
    worst = 0.0
    for x in xs:
        v = O.erfc_cheb32(float(x))
        ref = sp.erfc(float(x))
        rel = abs(v - ref) / ref
        worst = max(worst, rel)
        assert rel <= 4e-7 + 1.5e-7 * float(x) ** 2, (x, v, ref)
    assert worst < 1.1e-5


@pytest.mark.parametrize("a,b", [(-7.5, 0.5), (-3.0, 3.0), (-0.5, 20.0), (0.5, 7.9), (2.0, 2.5), (-7.9, -7.0),
                                 (-30.0, 1.0), (6.0, 30.0), (-1e3, 1e3)])
def test_trunc_normal_prec_f32_inverts_cdf(a, b):
    """The single-precision Gibbs draw returns the quantile of the FP64 draw (itself pinned against the
    CDF above) for the same uniform, to single-precision resolution of z, in every regime it accepts
    (both bounds inside 8 sigma on the far side of the mean)."""
    mu, s = 1.0, 2.0
    lo, hi = mu + a / s, mu + b / s
    for u in (2.0 ** -24, 1e-6, 0.01, 0.3, 0.5, 0.7, 0.99, 1 - 1e-6, 1 - 2.0 ** -24):
        uf = float(np.float32(u))
        x = O.trunc_normal_prec_f32(mu, s, lo, hi, uf)
        assert lo <= x <= hi
        z = (x - mu) * s
        z64 = (O.trunc_normal_prec(mu, s, lo, hi, uf) - mu) * s
        assert abs(z - z64) <= 1e-6 * max(1.0, abs(z64)), (a, b, u, z, z64)


def test_trunc_normal_prec_f32_declines_far_tails():
    """Bounds beyond 8 sigma on the far side of the mean are left to the FP64 draw (NaN)."""
    assert math.isnan(O.trunc_normal_prec_f32(0.0, 1.0, 8.5, 20.0, 0.5))
    assert math.isnan(O.trunc_normal_prec_f32(0.0, 1.0, -20.0, -8.5, 0.5))
    assert not math.isnan(O.trunc_normal_prec_f32(0.0, 1.0, 7.5, 20.0, 0.5))
