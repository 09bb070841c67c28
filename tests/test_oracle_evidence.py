"""End-to-end pin of the oracle's inner SMC (gamma search, evidence accumulation, real(16)
normalisation and resampling, Gibbs mover): on a tiny mesh (ndim = 8) the tempered posterior is
a truncated Gaussian, and -log Z has a closed form as a ratio of Gaussian box probabilities,

  Z = int_box exp(-L(s) - s^T LL s / (2 alpha^2)) ds / int_box exp(-s^T LL s / (2 alpha^2)) ds,
  int_box exp(-s^T A s / 2 + c^T s) ds = (2 pi)^(n/2) |A|^(-1/2) exp(c^T A^-1 c / 2) Pr_{N(A^-1 c, A^-1)}(box),

with L(s) = (s^T H s - 2 b^T s + dsd) / 2 + 3 ngnss log(sigma_gnss^2) / 2 (smc_slip.f90:56-104).
The SMC estimate must match within its Monte Carlo error."""
import math

import numpy as np
import scipy.stats as st

import oracle as O
from smc_inversion_fort_amd import synthetic as syn


def closed_form_neglog_z(F, obs, th, G):
    nd = G.shape[1]
    M = syn.MAX_SLIP
    H, b, dsd = O.gsd_gsg(G, obs["dvec"], obs["obs_sigma"], math.exp(th[4]), math.exp(th[5]))
    prec = F.laplacian(th)["ll"] / math.exp(th[3])

    def log_box_integral(A, c):
        Ai = np.linalg.inv(A)
        m = Ai @ c
        p = st.multivariate_normal.cdf(np.full(nd, M), mean=m, cov=Ai, lower_limit=np.zeros(nd), abseps=1e-13,
                                       releps=1e-7, maxpts=1000000)
        return nd / 2 * math.log(2 * math.pi) - 0.5 * np.linalg.slogdet(A)[1] + 0.5 * c @ m + math.log(p)

    const = 0.5 * dsd + 3 * obs["ngnss"] * th[5] / 2
    return const - log_box_integral(H + prec, b) + log_box_integral(prec, np.zeros(nd))


def test_inner_smc_gibbs_evidence_matches_closed_form():
    xmin, xmax, zmin, _ = syn.fault_frame()
    F = O.Fault([3], [2], xmin, xmax, zmin)

    def greens(th, pts, uv, nx, ne):
        return O.Fault([nx], [ne], xmin, xmax, zmin).greens(th, pts, uv, 0, pts.shape[1] // 3)

    obs = syn.make_observations("c2", 3, greens, nsta=8, sigma=0.5, nxi=3, neta=2)
    th = syn.theta_true().copy()
    G = F.greens(th, obs["obs_points"], obs["obs_unitvec"], 0, obs["ngnss"])
    exact = closed_form_neglog_z(F, obs, th, G)
    po = O.ProblemHolder(F, obs, syn.MAX_SLIP, 2000, mover=1, gibbs_sweeps=5)
    vals = np.array([po.fault_likelihood(th, O.inner_key(seed, 0, 0, 0))[0] for seed in range(8)])
    se = vals.std(ddof=1) / math.sqrt(len(vals))
    assert abs(vals.mean() - exact) < 4 * se + 0.03, (vals.mean(), exact, se)
    # the posterior mean of the final particles agrees with the truncated-Gaussian mean (rejection sampling)
    H, b, _ = O.gsd_gsg(G, obs["dvec"], obs["obs_sigma"], math.exp(th[4]), math.exp(th[5]))
    A = H + F.laplacian(th)["ll"] / math.exp(th[3])
    Ai = np.linalg.inv(A)
    rng = np.random.default_rng(0)
    x = (Ai @ b)[:, None] + np.linalg.cholesky(Ai) @ rng.normal(size=(G.shape[1], 200000))
    x = x[:, np.all((x >= 0) & (x <= syn.MAX_SLIP), axis=0)]
    assert x.shape[1] > 10000
    X = np.concatenate([po.slip_smc(th, G, O.inner_key(s, 0, 0, 0))["particles"] for s in range(2)], axis=1)
    sd = x.std(axis=1)
    assert np.all(np.abs(X.mean(axis=1) - x.mean(axis=1)) < 0.15 * sd)
    assert np.all(np.abs(X.std(axis=1) / sd - 1) < 0.15)
