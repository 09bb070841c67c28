"""The outer-SMC stage components of the library (include/smcf.h, host code: smcf_fault_*) against the oracle's
restatement of smc_fault.f90 on the same inputs.  They run on the host, so these tests need no GPU; they load
libsmcf.so (which loads the HIP runtime but makes no device call here)."""
import os

import numpy as np
import pytest

import oracle as O
import smc_inversion_fort_amd as S

pytestmark = pytest.mark.skipif(not os.path.exists(S.LIB_PATH), reason="libsmcf.so not built")


def _cloud(seed, nd=6, N=400):
    rng = np.random.default_rng(seed)
    X = np.asfortranarray(rng.normal(size=(nd, N)) * rng.uniform(0.5, 3, (nd, 1)) + rng.normal(size=(nd, 1)))
    L = rng.normal(size=N) * 30 + 200
    return X, L


def _outer_u(seed, stage):
    import ctypes as C
    okey = O.lib().orc_splitmix64(seed ^ 0x5EEDF00D)
    w = (C.c_uint64 * 2)()
    O.lib().orc_draw_u64x2(okey, 17, stage, 0, 0, 0, w)
    return (w[0] >> 11) * 2.0 ** -53, (w[1] >> 11) * 2.0 ** -106


@pytest.mark.parametrize("seed", range(4))
def test_stage_components_match_the_oracle(seed):
    X, L = _cloud(seed)
    N = L.shape[0]
    g, w, ev = S.fault_find_next_gamma(0.0 if seed % 2 == 0 else 0.3, L)
    go, wo, evo = O.find_next_gamma(0.0 if seed % 2 == 0 else 0.3, L, cv=0.2, tol=1e-8, cap=1000)
    assert g == pytest.approx(go, rel=1e-12, abs=1e-15)
    assert ev == pytest.approx(evo, rel=1e-12, abs=1e-12)
    wn = S.fault_normalize_weights(w)
    np.testing.assert_array_equal(wn, O.fault_normalize(w))
    m = S.fault_calc_mean_particles(X, wn)
    cov, info = S.fault_calc_cov_particles(X, wn, m)
    mo, covo, infoo = O.fault_mean_cov(X, wn)
    np.testing.assert_array_equal(m, mo)
    np.testing.assert_array_equal(np.tril(cov), np.tril(covo))
    assert info == infoo == 0
    a = S.fault_resample_particles(wn, 77 + seed, 3)
    uh, ul = _outer_u(77 + seed, 3)
    np.testing.assert_array_equal(a, O.resample_quad(wn, uh, ul))
    assert a.sum() == N
    a2, X2, L2, _, sidx = S.fault_reorder_to_send(a, X, L)
    ao, Xo, Lo = a.copy(), np.array(X, order="F"), L.copy()
    so = np.zeros(N, np.int32)
    O.lib().orc_fault_reorder(O.ip(ao), O.dp(Xo), O.dp(Lo), N, X.shape[0], O.ip(so))
    np.testing.assert_array_equal(sidx, so)
    np.testing.assert_array_equal(a2, ao)
    np.testing.assert_array_equal(X2, Xo)
    np.testing.assert_array_equal(L2, Lo)
    assert np.all(np.diff(a2) <= 0)


def test_gsmc_tempering_and_initial_draws():
    X, L = _cloud(9)
    q = np.random.default_rng(1).normal(size=L.shape[0]) * 5
    g, w, ev = S.fault_find_next_gamma(0.0, L, q)
    go, wo, evo = O.find_next_gamma(0.0, L - q, cv=0.2, tol=1e-8, cap=1000)
    assert g == pytest.approx(go, rel=1e-12)
    rng = np.array([[0.0] * 6, [1.0] * 6])
    U = S.fault_sample_init_particles(6, 500, 5, rng)
    assert U.min() >= 0 and U.max() < 1
    assert abs(U.mean() - 0.5) < 0.03
    A = np.tril(np.random.default_rng(2).normal(size=(6, 6))) + 3 * np.eye(6)
    Z = S.fault_sample_init_particles(6, 4000, 5, None, np.arange(6.0), np.asfortranarray(A))
    np.testing.assert_allclose(Z.mean(axis=1), np.arange(6.0), atol=0.25)
    C = A @ A.T  # sample covariance of 4000 draws: standard error about sqrt(2 C_aa C_bb / 4000) <= 0.3
    np.testing.assert_allclose(np.cov(Z), C, rtol=0, atol=1.2)


def test_mean_std():
    v = np.random.default_rng(3).normal(size=1001)
    m, sd = S.fault_mean_std(v)
    assert m == pytest.approx(v.mean(), rel=1e-13)
    assert sd == pytest.approx(v.std(), rel=1e-13)
