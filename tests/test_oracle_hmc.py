"""Pins of the oracle's HMC leapfrog (slip_leapfrog, smc_slip.f90:732-917), which the reference cannot
run here (it needs BLAS): on the quadratic tempered potential the trajectory converges to the exact
Hamiltonian flow exp(T [[0, M], [-A, 0]]) at second order in dtau, with T = (ntau + 1/2) dtau
(half kick, ntau drift-kick pairs, closing half drift), and reflections keep it in [0, max_slip]."""
import numpy as np
import scipy.linalg as sl

import oracle as O
from smc_inversion_fort_amd import synthetic as syn


def setup(gamma=0.3):
    xmin, xmax, zmin, _ = syn.fault_frame()

    def greens(th, pts, uv, nx, ne):
        return O.Fault([nx], [ne], xmin, xmax, zmin).greens(th, pts, uv, 0, pts.shape[1] // 3)

    obs = syn.make_observations("c2", 3, greens, nsta=8, sigma=0.5, nxi=3, neta=2)
    po = O.ProblemHolder(O.Fault([3], [2], xmin, xmax, zmin), obs, syn.MAX_SLIP, 16, mover=0)
    th = syn.theta_true()
    G = greens(th, obs["obs_points"], obs["obs_unitvec"], 3, 2)
    nd = G.shape[1]
    g0 = po.grad(th, G, gamma, np.zeros(nd))
    A = np.column_stack([po.grad(th, G, gamma, np.eye(nd)[i]) - g0 for i in range(nd)])
    return po, th, G, nd, A, g0


def test_leapfrog_converges_to_exact_flow_at_second_order():
    gam = 0.3
    po, th, G, nd, A, g0 = setup(gam)
    assert np.abs(A - A.T).max() <= 1e-12 * np.abs(A).max()
    rng = np.random.default_rng(1)
    cur = np.full(nd, 5.0)
    p0 = rng.normal(size=nd)
    cov = rng.uniform(0.5, 1.5, nd)
    T = 0.05
    gen = np.zeros((2 * nd + 1, 2 * nd + 1))
    gen[:nd, nd:2 * nd] = np.diag(cov)
    gen[nd:2 * nd, :nd] = -A
    gen[nd:2 * nd, 2 * nd] = -g0
    z = sl.expm(gen * T) @ np.concatenate([cur, p0, [1.0]])
    errs, dham = [], []
    for ntau in (25, 50, 100, 200):
        c, p, hc, hn = po.leapfrog(th, G, gam, cur, p0, cov, T / (ntau + 0.5), ntau)
        assert c.min() > 0.0 and c.max() < syn.MAX_SLIP  # no reflection on this trajectory
        errs.append(max(np.abs(c - z[:nd]).max(), np.abs(p - z[nd:2 * nd]).max() * 0.1))
        dham.append(abs(hn - hc))
    for k in range(3):
        assert 0.2 < errs[k + 1] / errs[k] < 0.3, errs
        assert 0.2 < dham[k + 1] / dham[k] < 0.3, dham
    assert errs[-1] < 1e-4


def test_leapfrog_reflections_stay_in_box_and_escape_sentinel():
    gam = 1.0
    po, th, G, nd, A, g0 = setup(gam)
    rng = np.random.default_rng(3)
    for _ in range(20):
        cur = rng.uniform(0, syn.MAX_SLIP, nd)
        p0 = rng.normal(size=nd) * 30
        c, p, hc, hn = po.leapfrog(th, G, gam, cur, p0, np.ones(nd), 0.05, 50)
        if hn == 1e20:  # |x| > 1e3: the candidate is the current point
            assert np.array_equal(c, cur)
        else:
            assert c.min() >= 0.0 and c.max() <= syn.MAX_SLIP


def test_accept_mover_settles_near_its_target():
    """Mover 2 (ntau leapfrog steps, step size adapted toward a target mean acceptance): past the first stages
    the stage's mean acceptance probability stays near the target, and the step size follows the adaptation
    rule from stage to stage."""
    xmin, xmax, zmin, _ = syn.fault_frame()

    def greens(th, pts, uv, nx, ne):
        return O.Fault([nx], [ne], xmin, xmax, zmin).greens(th, pts, uv, 0, pts.shape[1] // 3)

    obs = syn.make_observations("c1", 0, greens)  # 8x8 patches, 300 observations: ~27 stages
    th = syn.theta_true()
    G = greens(th, obs["obs_points"], obs["obs_unitvec"], 8, 8)
    for target in (0.65, 0.8):
        po = O.ProblemHolder(O.Fault([8], [8], xmin, xmax, zmin), obs, syn.MAX_SLIP, 400, mover=2, hmc_ntau=10,
                             hmc_accept=target)
        r = po.slip_smc(th, G, O.inner_key(7, 0, 0, 0))
        acc, dt = r["accept"], r["dtau"]
        assert r["nstage"] > 15 and 1.0 - r["gamma"][-1] <= 1e-7
        assert abs(dt[0] - 0.1) < 1e-15
        assert abs(np.median(acc[6:]) - target) < 0.08, acc
        # log dtau(s + 1) - log dtau(s) = 2 (acc(s) - target)
        np.testing.assert_allclose(np.diff(np.log(dt)), 2.0 * (acc[:-1] - target), atol=1e-12)
