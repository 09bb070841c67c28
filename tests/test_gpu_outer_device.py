"""The outer reweight and resample on the device (smcf_outer_reweight_resample) against the host components that
follow smc_fault.f90:372-637 in the reference's arithmetic order: gamma is equal (its bisection compares the CV with
0.2 at distances far above rounding until it stops), the normalised weights, the evidence increment and the Cholesky
factor agree to rounding, the counts equal those of the host resampler on the same weights, and the host-sort mode
reorders exactly as fault_reorder_to_send; the device radix sort gives descending counts with tied parents in
ascending index.  Then whole outer SMCs with the device stage against the host stage."""
import numpy as np
import pytest

import smc_inversion_fort_amd as S
from smc_inversion_fort_amd import synthetic as syn

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = S.Context(0)
    yield c
    c.close()


def stage_inputs(seed, N, nd=6, gsmc=False):
    rng = np.random.default_rng(seed)
    X = np.asfortranarray(rng.normal(size=(nd, N)) * rng.uniform(0.5, 3.0, (nd, 1)) + rng.normal(size=(nd, 1)))
    L = 50.0 + 20.0 * rng.standard_normal(N) ** 2 + X[0] ** 2
    q = rng.normal(size=N) * 3.0 if gsmc else None
    return X, L, q


@pytest.mark.parametrize("seed,N,gprev,gsmc", [(0, 20000, 0.0, False), (1, 20000, 0.37, True), (2, 1, 0.0, False),
                                               (3, 7, 0.9, False), (4, 1000, 0.999999, False), (5, 4097, 0.2, True)])
def test_device_stage_matches_host_components(ctx, seed, N, gprev, gsmc):
    print("case: device outer stage seed", seed, "N", N, "gprev", gprev, "gsmc", gsmc, flush=True)
    X, L, q = stage_inputs(seed, N, gsmc=gsmc)
    stage, oseed = 3 + seed, 1234 + seed
    g_h, w_h, ev_h = S.fault_find_next_gamma(gprev, L, q)
    d = S.outer_reweight_resample(ctx, X, L, gprev, oseed, stage, q=q, device_sort=False)
    assert d["gamma"] == g_h
    assert abs(d["neglog_evidence"] - ev_h) <= 1e-12 * max(1.0, abs(ev_h))
    wn_h = S.fault_normalize_weights(w_h)
    np.testing.assert_allclose(d["weights"], wn_h, rtol=1e-13, atol=1e-300)
    m_h = S.fault_calc_mean_particles(X, wn_h)
    chol_h, info_h = S.fault_calc_cov_particles(X, wn_h, m_h)
    assert d["info"] == info_h
    if info_h == 0:
        np.testing.assert_allclose(np.tril(d["cov_chol"]), np.tril(chol_h), rtol=1e-10,
                                   atol=1e-12 * np.abs(chol_h).max())
    # counts: the host resampler on the device's weights; then the host reorder of those counts
    a_h = S.fault_resample_particles(d["weights"], oseed, stage)
    assert a_h.sum() == N
    a2, X2, L2, q2, sidx2 = S.fault_reorder_to_send(a_h, X, L, q)
    np.testing.assert_array_equal(d["assigned"], a2)
    np.testing.assert_array_equal(d["sorted_idx"], sidx2)
    np.testing.assert_array_equal(d["particles"], X2)
    np.testing.assert_array_equal(d["likelihood"], L2)
    if gsmc:
        np.testing.assert_array_equal(d["q"], q2)
    # device radix sort: same counts, descending, ties in ascending index, rows carried along
    e = S.outer_reweight_resample(ctx, X, L, gprev, oseed, stage, q=q, device_sort=True)
    perm = e["sorted_idx"] - 1
    assert sorted(perm.tolist()) == list(range(N))
    np.testing.assert_array_equal(e["assigned"], a_h[perm])
    assert np.all(np.diff(e["assigned"]) <= 0)
    ties = np.diff(e["assigned"]) == 0
    assert np.all(np.diff(perm)[ties] > 0)
    np.testing.assert_array_equal(e["particles"], X[:, perm])
    np.testing.assert_array_equal(e["likelihood"], L[perm])
    if gsmc:
        np.testing.assert_array_equal(e["q"], q[perm])


def test_outer_smc_with_device_stage(ctx):
    print("case: outer SMC, host vs device stage", flush=True)
    xmin, xmax, zmin, _ = syn.fault_frame()

    def greens(th, pts, uv, nx, ne):
        obs0 = dict(obs_points=pts, obs_unitvec=uv, obs_sigma=np.ones(pts.shape[1]), dvec=np.zeros(pts.shape[1]),
                    nsar_total=0, ngnss=pts.shape[1] // 3)
        p = S.Problem(ctx, [nx], [ne], xmin, xmax, zmin, obs0, syn.MAX_SLIP, 8)
        g = p.calc_greens_func(th)
        p.close()
        return g
    obs = syn.make_observations("c2", 9, greens, nsta=12, sigma=0.5, nxi=4, neta=3)
    pb = S.Problem(ctx, [4], [3], xmin, xmax, zmin, obs, syn.MAX_SLIP, 96, mover=S.MOVER_GIBBS, gibbs_sweeps=2,
                   max_batch=256)
    rr = syn.prior_range()
    res = {}
    for mode in (0, 1, 2):
        pb.set_outer_stage(mode)
        res[mode] = pb.fault_smc_exec(rr, 128, 77, max_stages=6)
        print("mode", mode, "stages", res[mode]["nstage"], "evidence", res[mode]["evidence"],
              "gamma", np.round(res[mode]["gamma"], 6).tolist(), flush=True)
    pb.set_outer_stage(0)
    h, d1, d2 = res[0], res[1], res[2]
    # the first stage's gamma depends only on the initial wave, identical in every mode
    assert d1["gamma"][0] == h["gamma"][0] and d2["gamma"][0] == h["gamma"][0]
    # host sort: the same parents and draws; the covariance factor differs by rounding only
    assert d1["nstage"] == h["nstage"]
    np.testing.assert_allclose(d1["gamma"], h["gamma"], rtol=1e-6)
    assert abs(d1["evidence"] - h["evidence"]) <= 1e-6 * abs(h["evidence"]) + 1e-6
    # device sort: tied parents in another order (other children keys), the same algorithm
    assert np.all(np.isfinite(d2["particles"])) and np.isfinite(d2["evidence"])
    for d in (d1, d2):
        lo, hi = rr[0], rr[1]
        assert np.all(d["particles"] >= lo[:, None] - 1e-12) and np.all(d["particles"] <= hi[:, None] + 1e-12)
    pb.close()
