"""Pin of the oracle's outer-SMC composition (fault_smc_exec, smc_fault.f90:785-1075: tempering, weights,
evidence, covariance, resampling, reordering and the Metropolis steps of work_mcmc_sampling, :639-783).

The fault likelihood is replaced by a Gaussian in the 6 geometry parameters, L(theta) = 0.5 sum ((theta_d - m_d)
/ s_d)^2, under the uniform prior on the box of the range.  Then the evidence is a product of normal box
probabilities,

    Z = prod_d sqrt(2 pi) s_d (Phi((b_d - m_d) / s_d) - Phi((a_d - m_d) / s_d)) / (b_d - a_d),

and the posterior is the product of the truncated normals.

The reference moves the children of one parent as one Metropolis chain (child r starts from child r - 1), so
heavier parents take more steps; with the short proposals of 0.04 x the covariance this leaves the population's
marginal off the tempered target and -log Z about 0.1 nats high at any N (at N = 16 000 too).  The pin therefore
has three parts: a vectorised numpy restatement of the textbook form (children moved independently from their
parent) reproduces the closed form, which checks the closed form and the restatement's tempering, evidence and
resampling; the same restatement with chained children reproduces the bias; and the oracle agrees with the
chained restatement within Monte Carlo error, and with the closed form within that bias.
"""
import math

import numpy as np
from scipy import stats

import oracle as O
from smc_inversion_fort_amd import synthetic as syn

LO = np.array([-2.0, -1.0, 0.0, 10.0, -3.0, 1.0])
HI = np.array([2.0, 3.0, 5.0, 14.0, 1.0, 2.0])
# some modes well inside the box, some near or beyond a bound (truncation matters there)
M = np.array([0.3, 2.7, 1.0, 12.0, 1.4, 1.5])
SD = np.array([0.5, 0.4, 1.2, 0.3, 0.6, 2.0])
N, K = 4000, 64


def exact():
    a, b = (LO - M) / SD, (HI - M) / SD
    mass = stats.norm.cdf(b) - stats.norm.cdf(a)
    neglog_z = -float(np.sum(np.log(math.sqrt(2 * math.pi) * SD * mass / (HI - LO))))
    tn = stats.truncnorm(a, b, loc=M, scale=SD)
    return neglog_z, tn.mean(), tn.std()


def like(X):
    return 0.5 * np.sum(((X - M) / SD) ** 2, axis=1)


def restated_smc(rng, chained):
    """The outer SMC written from the algorithm's description with numpy's generator: CV-0.2 bisection on
    gamma, -log of the mean weight (+ dg * min L) as the evidence increment, systematic resampling, proposals
    from the Cholesky factor of 0.04 x the weighted covariance reflected once into the box, one Metropolis step
    per child at the new gamma.  Returns (-log Z, final particles (N x 6))."""
    X = LO + rng.random((N, 6)) * (HI - LO)
    Lv = like(X)
    g, ev = 0.0, 0.0
    while 1.0 - g > 1e-8:
        g0, lmin = g, Lv.min()
        lower, upper, err = g0, 1.0, 1.0
        while err > 1e-8:
            g = (lower + upper) / 2
            w = np.exp(-(g - g0) * (Lv - lmin))
            cv = w.std() / w.mean()
            lower, upper = (lower, g) if cv > 0.2 else (g, upper)
            err = abs(cv - 0.2)
            if abs(g - 1.0) < 1e-8:
                break
        w = np.exp(-(g - g0) * (Lv - lmin))
        ev += -math.log(w.mean()) + (g - g0) * lmin
        w /= w.sum()
        D = X - w @ X
        chol = np.linalg.cholesky(0.04 * (D * w[:, None]).T @ D)
        edges = np.floor(np.concatenate([[0.0], np.cumsum(w)]) * N + rng.random()).astype(int)
        counts = np.diff(np.minimum(edges, N))
        counts[-1] += N - counts.sum()
        parent = np.repeat(np.arange(N), counts)
        rank = np.arange(N) - np.repeat(np.cumsum(counts) - counts, counts)
        Xn, Ln = X[parent].copy(), Lv[parent].copy()
        for r in range(int(rank.max()) + 1 if chained else 1):
            sel = np.nonzero(rank == r)[0] if chained else np.arange(N)
            if r > 0:  # child r continues from child r - 1 of the same parent
                Xn[sel], Ln[sel] = Xn[sel - 1], Ln[sel - 1]
            cand = Xn[sel] + rng.standard_normal((len(sel), 6)) @ chol.T
            cand = np.where(cand < LO, np.minimum(2 * LO - cand, HI), cand)
            cand = np.where(cand > HI, np.maximum(2 * HI - cand, LO), cand)
            lnew = like(cand)
            acc = g * (Ln[sel] - lnew) > np.log(rng.random(len(sel)))
            Xn[sel] = np.where(acc[:, None], cand, Xn[sel])
            Ln[sel] = np.where(acc, lnew, Ln[sel])
        X, Lv = Xn, Ln
    return ev, X


def summary(evs, Xs):
    evs = np.array(evs)
    return (evs.mean(), evs.std(ddof=1) / math.sqrt(len(evs)), np.array([x.mean(axis=0) for x in Xs]),
            np.array([x.std(axis=0) for x in Xs]))


def oracle_runs():
    xmin, xmax, zmin, _ = syn.fault_frame()
    nobs = 3
    obs = dict(obs_points=np.zeros((2, nobs)), obs_unitvec=np.tile([[1.0], [0.0], [0.0]], (1, nobs)),
               obs_sigma=np.ones(nobs), dvec=np.zeros(nobs), nsar_total=0, ngnss=1)
    po = O.ProblemHolder(O.Fault([2], [2], xmin, xmax, zmin), obs, syn.MAX_SLIP, 8, mover=1, test_like=(M, SD))
    evs, Xs = [], []
    for k in range(K):
        r = po.fault_smc(np.vstack([LO, HI]), N, 100 + k, max_stages=60)
        X = r["particles"].T
        assert 1.0 - r["gamma"][-1] <= 1e-8
        assert np.all(X >= LO) and np.all(X <= HI)
        np.testing.assert_allclose(r["likelihood"], like(X), rtol=1e-12)  # each particle carries its own L
        evs.append(r["evidence"])
        Xs.append(X)
    return evs, Xs


def test_outer_smc_composition_against_closed_form_and_restatement():
    nz, tmean, tstd = exact()
    assert nz > 5.0  # the posterior is well inside the prior: the evidence is a real test
    res = {}
    for chained in (False, True):
        rng = np.random.default_rng(7 + chained)
        res[chained] = summary(*zip(*[restated_smc(rng, chained) for _ in range(K)]))
    res["oracle"] = summary(*oracle_runs())
    for name, (m, se, mu, sd) in res.items():
        print("%-7s -log Z %.4f +- %.4f (exact %.4f); posterior mean" % (name, m, se, nz), np.round(mu.mean(0), 3),
              "std", np.round(sd.mean(0), 4))
    # textbook form: unbiased
    m0, se0, _, _ = res[False]
    assert abs(m0 - nz) < 4 * se0 + 0.02
    # chained children: biased high, by a margin well outside the Monte Carlo error
    m1, se1, mu1, sd1 = res[True]
    assert 0.04 < m1 - nz < 0.25
    # the oracle: the chained restatement within Monte Carlo error (evidence, posterior means and standard
    # deviations), the closed form within the bias
    m2, se2, mu2, sd2 = res["oracle"]
    assert abs(m2 - m1) < 4 * math.hypot(se1, se2) + 0.01
    assert abs(m2 - nz) < 0.25
    for a, b in ((mu1, mu2), (sd1, sd2)):
        sem = np.hypot(a.std(axis=0, ddof=1), b.std(axis=0, ddof=1)) / math.sqrt(K)
        assert np.all(np.abs(a.mean(axis=0) - b.mean(axis=0)) < 4 * sem + 0.002 * tstd)
    # posterior means: within Monte Carlo error of the truncated normals (both chained forms)
    for mu in (mu1, mu2):
        sem = mu.std(axis=0, ddof=1) / math.sqrt(K)
        assert np.all(np.abs(mu.mean(axis=0) - tmean) < 5 * sem + 0.02 * tstd)
