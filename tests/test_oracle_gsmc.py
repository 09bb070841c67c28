"""Pins of the oracle's gsmc branch of the outer SMC (smc_fault.f90:9-52, 405-485, 708-775):
q against the closed-form Gaussian density, the initial draws against N(qmean, qcov), and the
first tempering step against a brute-force root of CV(exp(-dg (L - q))) = 0.2."""
import math

import numpy as np
import pytest
import scipy.optimize as opt
import scipy.stats as st

import oracle as O
from smc_inversion_fort_amd import synthetic as syn


def tiny_problem(P=32):
    xmin, xmax, zmin, _ = syn.fault_frame()

    def greens(th, pts, uv, nx, ne):
        return O.Fault([nx], [ne], xmin, xmax, zmin).greens(th, pts, uv, 0, pts.shape[1] // 3)

    obs = syn.make_observations("c2", 3, greens, nsta=8, sigma=0.5, nxi=3, neta=2)
    return O.ProblemHolder(O.Fault([3], [2], xmin, xmax, zmin), obs, syn.MAX_SLIP, P, mover=1, gibbs_sweeps=2)


def proposal(nd=6, seed=4, scale=0.3):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(nd, nd)) * scale
    qcov = A @ A.T + np.diag([1.0, 1.0, 4.0, 1.0, 0.3, 0.3])
    return syn.theta_true(), qcov


def test_neglog_q_is_gaussian_density():
    qmean, qcov = proposal()
    inv = np.linalg.inv(qcov)
    ld = np.linalg.slogdet(qcov)[1]
    rng = np.random.default_rng(2)
    for _ in range(50):
        x = qmean + rng.normal(size=6) * 2.0
        ref = -st.multivariate_normal(qmean, qcov).logpdf(x)
        assert O.fault_neglog_q(x, qmean, np.asfortranarray(inv), ld) == pytest.approx(ref, rel=1e-12, abs=1e-12)


def test_initial_particles_follow_qcov():
    """max_stages = 0 returns the initial draws; a non-diagonal qcov tells L z from L^T z."""
    pb = tiny_problem(P=16)
    qmean, qcov = proposal(scale=1.0)
    ch = np.linalg.cholesky(qcov)
    g = (qmean, np.asfortranarray(ch), np.asfortranarray(np.linalg.inv(qcov)), np.linalg.slogdet(qcov)[1])
    N = 1500
    o = pb.fault_smc(syn.prior_range(), N, 11, max_stages=0, gsmc=g)
    X = o["particles"]
    assert o["nstage"] == 0
    d = X - qmean[:, None]
    emp = d @ d.T / N
    sd = np.sqrt(np.diag(qcov))
    corr_err = (emp - qcov) / np.outer(sd, sd)
    assert np.abs(corr_err).max() < 0.15
    wrong = ch.T @ ch
    assert np.abs((qcov - wrong) / np.outer(sd, sd)).max() > 0.6
    assert np.abs((emp - wrong) / np.outer(sd, sd)).max() > 0.4
    assert np.abs(X.mean(axis=1) - qmean).max() < 4 * sd.max() / math.sqrt(N)


def test_first_stage_tempers_L_minus_q():
    pb = tiny_problem()
    qmean, qcov = proposal(seed=9)
    inv = np.linalg.inv(qcov)
    ld = np.linalg.slogdet(qcov)[1]
    g = (qmean, np.asfortranarray(np.linalg.cholesky(qcov)), np.asfortranarray(inv), ld)
    N, seed = 24, 5
    X0 = pb.fault_smc(syn.prior_range(), N, seed, max_stages=0, gsmc=g)["particles"]
    o1 = pb.fault_smc(syn.prior_range(), N, seed, max_stages=1, gsmc=g)
    L = np.array([pb.fault_likelihood(X0[:, i], O.inner_key(seed, 0, 0, i))[0] for i in range(N)])
    q = np.array([-st.multivariate_normal(qmean, qcov).logpdf(X0[:, i]) for i in range(N)])
    R = L - q
    Rm = R.min()

    def cv(gm):
        w = np.exp(-gm * (R - Rm))
        return w.std() / w.mean() - 0.2

    if cv(1.0) <= 0:
        gam = 1.0
    else:
        gam = opt.brentq(cv, 0.0, 1.0, xtol=1e-14)
    assert o1["gamma"][0] == pytest.approx(gam, abs=1e-7)
    ev = -math.log(np.mean(np.exp(-o1["gamma"][0] * (R - Rm)))) + o1["gamma"][0] * Rm
    assert o1["evidence"] == pytest.approx(ev, rel=1e-9, abs=1e-9)
    # tempering L alone would give another step
    Lm = L.min()
    cvL = lambda gm: np.std(np.exp(-gm * (L - Lm))) / np.mean(np.exp(-gm * (L - Lm))) - 0.2
    gL = 1.0 if cvL(1.0) <= 0 else opt.brentq(cvL, 0.0, 1.0, xtol=1e-14)
    assert abs(gL - gam) > 1e-4
