"""Pins of the oracle's outer-SMC covariance, Cholesky factor and proposal, which the reference computes with
LAPACK/BLAS (dpotrf and dtrmv, smc_fault.f90:522-570, 711-747) and which therefore cannot run here.

* fault_calc_cov_particles: the lower factor L of 0.04 * (weighted covariance) satisfies L L^T = 0.04 Sigma_w,
  with Sigma_w formed independently by numpy, and equals numpy's Cholesky factor; a dimension without spread
  makes the factorisation stop with info = that dimension, as dpotrf reports it.
* the proposal: cur + L z (numpy's triangular product) inside the range; reflected once about the violated
  bound (and clipped) outside it; dimensions whose range is at most 1e-3 wide are left unreflected.
"""
import numpy as np

import oracle as O


def _particles(seed, nd=6, N=300):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(nd, nd))
    X = (A @ rng.normal(size=(nd, N))) + rng.normal(size=(nd, 1)) * 5
    w = rng.uniform(0.1, 1.0, N)
    return np.asfortranarray(X), w / w.sum()


def test_cov_times_004_cholesky_matches_numpy():
    for seed in range(5):
        X, w = _particles(seed)
        m, Lc, info = O.fault_mean_cov(X, w)
        assert info == 0
        mu = X @ w
        np.testing.assert_allclose(m, mu, rtol=1e-13, atol=1e-13)
        D = X - mu[:, None]
        Sw = (D * w) @ D.T
        Lo = np.tril(Lc)
        np.testing.assert_allclose(Lo @ Lo.T, 0.04 * Sw, rtol=0, atol=1e-12 * np.abs(Sw).max())
        np.testing.assert_allclose(Lo, np.linalg.cholesky(0.04 * Sw), rtol=0, atol=1e-12 * np.abs(Lo).max())
        assert np.all(np.diag(Lo) > 0)
        assert np.all(np.triu(Lc, 1) == 0.0)  # the upper triangle is never written


def test_cholesky_reports_the_first_dimension_without_spread():
    X, w = _particles(7)
    X[3, :] = 0.0  # no spread in dimension 4 (1-based): its weighted variance is exactly zero
    _, Lc, info = O.fault_mean_cov(X, w)
    assert info == 4
    assert Lc[3, 3] <= 0.0


def test_proposal_is_cur_plus_lower_factor_times_z_inside_the_range():
    X, w = _particles(3)
    _, Lc, _ = O.fault_mean_cov(X, w)
    Lo = np.tril(Lc)
    rng = np.random.default_rng(0)
    wide = np.array([[-1e6] * 6, [1e6] * 6])
    for _ in range(50):
        cur = rng.normal(size=6)
        z = rng.normal(size=6)
        cand = O.fault_propose(Lc, cur, z, wide)
        np.testing.assert_allclose(cand, cur + Lo @ z, rtol=1e-14, atol=1e-14)


def test_proposal_reflects_into_the_range():
    X, w = _particles(4)
    _, Lc, _ = O.fault_mean_cov(X, w)
    Lo = np.tril(Lc)
    rng = np.random.default_rng(1)
    lo, hi = -0.5 * np.ones(6), 0.5 * np.ones(6)
    hi[5] = lo[5] + 1e-4  # a fixed dimension (width <= 1e-3): not reflected
    rr = np.array([lo, hi])
    wide = np.array([[-1e6] * 6, [1e6] * 6])
    for _ in range(200):
        cur = rng.uniform(lo, np.maximum(hi, lo))
        z = rng.normal(size=6) * 3
        free = O.fault_propose(Lc, cur, z, wide)  # the same product, without reflection
        np.testing.assert_allclose(free, cur + Lo @ z, rtol=1e-13, atol=1e-13)
        cand = O.fault_propose(Lc, cur, z, rr)
        for d in range(5):
            if lo[d] <= free[d] <= hi[d]:
                assert cand[d] == free[d]
            elif free[d] < lo[d]:
                assert cand[d] == min(2 * lo[d] - free[d], hi[d])
            else:
                assert cand[d] == max(2 * hi[d] - free[d], lo[d])
            assert lo[d] <= cand[d] <= hi[d]
        assert cand[5] == free[5]
