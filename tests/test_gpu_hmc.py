"""HMC movers on the device against the oracle, at the sizes of the two HMC configurations.

c1 (8x8 patches, 100 GNSS stations / 300 observations, sigma 5 mm, P = 1000, the reference mover: ESJD-tuned HMC
with ntau = 50, smc_slip.f90:591-1214): over 10 seeds the first three tempering steps agree exactly (1e-8), and
-log Z, the posterior means and the posterior standard deviations agree within Monte Carlo error.  The reference
HMC amplifies rounding differences (a 1e-15 change of H moves gamma by ~1e-3 from about the fifth stage on), so
later stages can only agree statistically.

c3 (32x16 patches, 1000 stations / 3000 observations, sigma 3 mm, P = 2000, HMC with 10 leapfrog steps and the
step size adapted toward 0.65 acceptance): the first three stages against the oracle, then the whole inner SMC on
the device, whose last stages must run near the target acceptance.
"""
import math
import sys

import numpy as np
import pytest

import oracle as O
import smc_inversion_fort_amd as S
from smc_inversion_fort_amd import synthetic as syn

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = S.Context(0)
    yield c
    c.close()


def device_obs(ctx, cfg, seed, **kw):
    xmin, xmax, zmin, _ = syn.fault_frame()

    def greens(th, pts, uv, nx, ne):
        obs0 = dict(obs_points=pts, obs_unitvec=uv, obs_sigma=np.ones(pts.shape[1]), dvec=np.zeros(pts.shape[1]),
                    nsar_total=0, ngnss=pts.shape[1] // 3)
        p = S.Problem(ctx, [nx], [ne], xmin, xmax, zmin, obs0, syn.MAX_SLIP, 8)
        g = p.calc_greens_func(th)
        p.close()
        return g
    return syn.make_observations(cfg, seed, greens, **kw)


def test_c1_full_size_reference_hmc_vs_oracle(ctx):
    say("case: c1 full size, reference HMC, 10 seeds")
    c = syn.CONFIGS["c1"]
    obs = device_obs(ctx, "c1", 0)
    xmin, xmax, zmin, _ = syn.fault_frame()
    P = c["P"]
    pb = S.Problem(ctx, [c["nxi"]], [c["neta"]], xmin, xmax, zmin, obs, syn.MAX_SLIP, P, mover=S.MOVER_HMC)
    po = O.ProblemHolder(O.Fault([c["nxi"]], [c["neta"]], xmin, xmax, zmin), obs, syn.MAX_SLIP, P, mover=0)
    theta = syn.theta_true()
    G = pb.calc_greens_func(theta)
    assert pb.ndim == 112 and pb.nobs == 300
    dn, on, dm, om, ds, os_ = [], [], [], [], [], []
    for k in range(10):
        key = O.inner_key(100 + k, 0, 0, 0)
        d = pb.slip_smc_exec(theta, key)
        o = po.slip_smc(theta, G, key)
        say("seed", k, "stages", d["nstage"], o["nstage"], "-logZ", d["neglog"], o["neglog"])
        np.testing.assert_allclose(d["gamma"][:3], o["gamma"][:3], rtol=0, atol=1e-8)
        assert np.all(d["particles"] >= 0.0) and np.all(d["particles"] <= syn.MAX_SLIP)
        dn.append(d["neglog"]); on.append(o["neglog"])
        dm.append(d["particles"].mean(axis=1)); om.append(o["particles"].mean(axis=1))
        ds.append(d["particles"].std(axis=1)); os_.append(o["particles"].std(axis=1))
    dn, on = np.array(dn), np.array(on)
    se = math.sqrt(dn.var(ddof=1) / 10 + on.var(ddof=1) / 10)
    say("-logZ device %.4f oracle %.4f se %.4f" % (dn.mean(), on.mean(), se))
    assert abs(dn.mean() - on.mean()) < 4 * se + 0.05
    for name, a, b, slack in (("mean", dm, om, 0.02), ("std", ds, os_, 0.01)):
        a, b = np.array(a), np.array(b)
        sem = np.sqrt(a.var(axis=0, ddof=1) / 10 + b.var(axis=0, ddof=1) / 10)
        z = np.abs(a.mean(axis=0) - b.mean(axis=0)) / (sem + 1e-12)
        say("posterior %s: max |diff| %.4f, max z %.2f" % (name, np.abs(a.mean(axis=0) - b.mean(axis=0)).max(),
                                                          z.max()))
        assert np.all(np.abs(a.mean(axis=0) - b.mean(axis=0)) < 4 * sem + slack)
    pb.close()


def test_c3_full_size_hmc_accept(ctx):
    say("case: c3 full size, HMC ntau 10 tuned to 0.65 acceptance")
    c = syn.CONFIGS["c3"]
    obs = device_obs(ctx, "c3", 0)
    xmin, xmax, zmin, _ = syn.fault_frame()
    P = c["P"]
    pb = S.Problem(ctx, [c["nxi"]], [c["neta"]], xmin, xmax, zmin, obs, syn.MAX_SLIP, P,
                   mover=S.MOVER_HMC_ACCEPT, hmc_ntau=10, hmc_accept=0.65)
    assert pb.ndim == 992 and pb.nobs == 3000
    theta = syn.theta_true()
    key = O.inner_key(300, 0, 0, 0)
    G = pb.calc_greens_func(theta)
    po = O.ProblemHolder(O.Fault([c["nxi"]], [c["neta"]], xmin, xmax, zmin), obs, syn.MAX_SLIP, P, mover=2,
                         hmc_ntau=10, hmc_accept=0.65, stage_cap=3)
    o = po.slip_smc(theta, G, key)
    say("oracle first stages: gamma", o["gamma"].tolist(), "acceptance", o["accept"].tolist())
    d = pb.slip_smc_exec(theta, key)
    dt, acc = pb.hmc_accept_state()
    say("device: stages", d["nstage"], "-logZ", d["neglog"], "last-stage acceptance %.3f next dtau %.4g" % (acc, dt))
    np.testing.assert_allclose(d["gamma"][:3], o["gamma"][:3], rtol=0, atol=1e-8)
    assert np.all(d["particles"] >= 0.0) and np.all(d["particles"] <= syn.MAX_SLIP)
    if d["nstage"] <= 200:
        assert d["gamma"][-1] == pytest.approx(1.0, abs=1e-7)
        assert 0.4 <= acc <= 0.9
    pb.close()


def test_hmc_accept_small_vs_oracle(ctx):
    """Mover 2 on a small problem: exact first stages, Monte Carlo agreement over 8 keys."""
    say("case: HMC ntau 10 / 0.65 acceptance, small problem, 8 keys")
    xmin, xmax, zmin, _ = syn.fault_frame()

    def greens(th, pts, uv, nx, ne):
        return O.Fault([nx], [ne], xmin, xmax, zmin).greens(th, pts, uv, 0, pts.shape[1] // 3)
    obs = syn.make_observations("c2", 5, greens, nsta=12, sigma=0.5, nxi=4, neta=3)
    P = 256
    pb = S.Problem(ctx, [4], [3], xmin, xmax, zmin, obs, syn.MAX_SLIP, P, mover=S.MOVER_HMC_ACCEPT, hmc_ntau=10)
    po = O.ProblemHolder(O.Fault([4], [3], xmin, xmax, zmin), obs, syn.MAX_SLIP, P, mover=2, hmc_ntau=10)
    theta = syn.theta_true()
    G = pb.calc_greens_func(theta)
    dn, on, dm, om = [], [], [], []
    for k in range(8):
        key = O.inner_key(4, 2, 1, k)
        d = pb.slip_smc_exec(theta, key)
        o = po.slip_smc(theta, G, key)
        say("key", k, "stages", d["nstage"], o["nstage"], "neglog", d["neglog"], o["neglog"],
            "oracle acceptance (last 3)", np.round(o["accept"][-3:], 3).tolist())
        np.testing.assert_allclose(d["gamma"][:3], o["gamma"][:3], rtol=0, atol=1e-8)
        dn.append(d["neglog"]); on.append(o["neglog"])
        dm.append(d["particles"].mean(axis=1)); om.append(o["particles"].mean(axis=1))
    dn, on = np.array(dn), np.array(on)
    se = math.sqrt(dn.var(ddof=1) / 8 + on.var(ddof=1) / 8)
    say("mean -logZ device", dn.mean(), "oracle", on.mean(), "se", se)
    assert abs(dn.mean() - on.mean()) < 4 * se + 0.05
    dm, om = np.array(dm), np.array(om)
    sem = np.sqrt(dm.var(axis=0, ddof=1) / 8 + om.var(axis=0, ddof=1) / 8)
    assert np.all(np.abs(dm.mean(axis=0) - om.mean(axis=0)) < 4 * sem + 0.02)
    pb.close()


@pytest.mark.parametrize("mover", [S.MOVER_HMC, S.MOVER_HMC_ACCEPT])
def test_hmc_batch_matches_solo(ctx, mover):
    """Instances of one batch finish at different stages; the finished ones must drop out of the stage's wave
    work, and every instance's result must equal its run alone (bit for bit)."""
    say("case: HMC batch of 4 vs solo, mover", mover)
    c = syn.CONFIGS["c1"]
    obs = device_obs(ctx, "c1", 2)
    xmin, xmax, zmin, _ = syn.fault_frame()
    P = 400
    kw = dict(mover=mover, hmc_ntau=10, hmc_accept=0.65) if mover == S.MOVER_HMC_ACCEPT else dict(mover=mover)
    th = syn.sample_thetas(4, 21)
    keys = [O.inner_key(5, 1, 0, k) for k in range(4)]
    pb = S.Problem(ctx, [c["nxi"]], [c["neta"]], xmin, xmax, zmin, obs, syn.MAX_SLIP, P, max_batch=4, **kw)
    nl, ns = pb.fault_calc_likelihood(th, keys)
    say("batch: stages", ns.tolist(), "-log Z", nl.tolist())
    assert len(set(ns.tolist())) > 1, "the geometries should finish at different stages"
    for k in range(4):
        n1, s1 = pb.fault_calc_likelihood(th[:, k:k + 1], keys[k:k + 1])
        assert s1[0] == ns[k] and n1[0] == nl[k], (k, s1[0], ns[k], n1[0], nl[k])
    pb.close()


@pytest.mark.parametrize("mover", [S.MOVER_HMC, S.MOVER_HMC_ACCEPT])
def test_hmc_split_bf16_gradients_vs_fp64(ctx, mover):
    """The split-BF16 form of the HMC movers (leapfrog gradients from split-BF16 products, FP64 Hamiltonian in the
    Metropolis test): at c1 sizes over 8 keys, -log Z and the posterior means agree with the FP64 form within Monte
    Carlo error, and the first tempering step (before any move) is identical."""
    say("case: HMC split-BF16 gradients vs FP64, mover", mover)
    c = syn.CONFIGS["c1"]
    obs = device_obs(ctx, "c1", 3)
    xmin, xmax, zmin, _ = syn.fault_frame()
    P = 1000
    kw = dict(mover=mover, hmc_ntau=10, hmc_accept=0.65) if mover == S.MOVER_HMC_ACCEPT else dict(mover=mover)
    res = {}
    for prec in (S.PRECISION_FP64, S.PRECISION_MIXED_BF16):
        pb = S.Problem(ctx, [c["nxi"]], [c["neta"]], xmin, xmax, zmin, obs, syn.MAX_SLIP, P, precision=prec, **kw)
        theta = syn.theta_true()
        out = [pb.slip_smc_exec(theta, O.inner_key(40 + k, 0, 0, 0)) for k in range(8)]
        res[prec] = out
        pb.close()
    f, m = res[S.PRECISION_FP64], res[S.PRECISION_MIXED_BF16]
    for a, b in zip(f, m):
        assert a["gamma"][0] == b["gamma"][0]
    nf, nm = np.array([r["neglog"] for r in f]), np.array([r["neglog"] for r in m])
    se = math.sqrt(nf.var(ddof=1) / 8 + nm.var(ddof=1) / 8)
    say("-log Z fp64 %.4f split-bf16 %.4f se %.4f; stages" % (nf.mean(), nm.mean(), se),
        [r["nstage"] for r in f], [r["nstage"] for r in m])
    assert abs(nf.mean() - nm.mean()) < 4 * se + 0.05
    mf = np.array([r["particles"].mean(axis=1) for r in f])
    mm = np.array([r["particles"].mean(axis=1) for r in m])
    sem = np.sqrt(mf.var(axis=0, ddof=1) / 8 + mm.var(axis=0, ddof=1) / 8)
    assert np.all(np.abs(mf.mean(axis=0) - mm.mean(axis=0)) < 4 * sem + 0.02 * np.abs(mf.mean(axis=0)).max())

