"""Component procedures of the inner SMC on the device (include/smcf.h, one geometry) against the oracle:
slip_calc_likelihood, slip_calc_prior, slip_calc_grad (smc_slip.f90:56-182), slip_leapfrog (732-917) on fixed
momenta, slip_gen_init_particles (184-287) in the keyed stream, and the weighted moments (485-570).

The leapfrog is the product's trajectory path (Y = H X on the FP64 matrix cores, then the gradient / kick / drift
kernel), run over a whole 50-step trajectory from the same (cur, p0); the oracle sums H x in another order, so
the two agree to rounding amplified by the trajectory, checked at 1e-9."""
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
def setup():
    xmin, xmax, zmin, _ = syn.fault_frame()
    nxi, neta, P = 4, 3, 48

    def greens(th, pts, uv, nx, ne):
        return O.Fault([nx], [ne], xmin, xmax, zmin).greens(th, pts, uv, 0, pts.shape[1] // 3)

    obs = syn.make_observations("c2", 6, greens, nsta=12, sigma=0.5, nxi=nxi, neta=neta)
    ctx = S.Context(0)
    pb = S.Problem(ctx, [nxi], [neta], xmin, xmax, zmin, obs, syn.MAX_SLIP, P, mover=S.MOVER_HMC)
    f = O.Fault([nxi], [neta], xmin, xmax, zmin)
    theta = syn.theta_true() + np.array([0.4, -0.3, 1.5, 0.2, 0.1, -0.2])
    G = pb.calc_greens_func(theta)  # the device's G on both sides: the checks below are of the later steps
    lap = f.laplacian(theta)
    H, b, dsd = O.gsd_gsg(G, obs["dvec"], obs["obs_sigma"], math.exp(theta[4]), math.exp(theta[5]))
    alpha2 = np.full(2 * f.nnode, math.exp(theta[3]))
    po = O.ProblemHolder(f, obs, syn.MAX_SLIP, P, mover=0)
    d = dict(pb=pb, f=f, theta=theta, G=G, lap=lap, H=H, b=b, dsd=dsd, alpha2=alpha2, po=po, obs=obs, P=P, ctx=ctx)
    yield d
    pb.close()
    ctx.close()


def test_likelihood_prior_grad_vs_oracle(setup):
    say("case: component likelihood / prior / grad vs oracle")
    d = setup
    f, th, nd = d["f"], d["theta"], d["f"].ndim
    X = np.asfortranarray(np.random.default_rng(0).uniform(0, 3.0, size=(nd, 20)))
    Ld = d["pb"].slip_calc_likelihood(th, X)
    Rd = d["pb"].slip_calc_prior(th, X)
    gam = 0.37
    Gd = d["pb"].slip_calc_grad(th, gam, X)
    lap = d["lap"]
    for k in range(X.shape[1]):
        Lo = O.likelihood(X[:, k], d["H"], d["b"], d["dsd"], 0, d["obs"]["ngnss"], th[4], th[5])
        Ro = O.prior(X[:, k], d["alpha2"], lap["li"], lap["lv"], f.nnode)
        go = O.grad(X[:, k], f.nnode, f.ndof, lap["li"], lap["lv"], lap["lti"], lap["ltv"], gam, d["alpha2"], d["b"],
                    d["H"])
        assert Ld[k] == pytest.approx(Lo, rel=1e-10)
        assert Rd[k] == pytest.approx(Ro, rel=1e-12)
        assert np.abs(Gd[:, k] - go).max() <= 1e-10 * np.abs(go).max()


@pytest.mark.parametrize("scale,dtau", [(1.0, 0.01), (25.0, 0.02)])
def test_leapfrog_trajectories_vs_oracle(setup, scale, dtau):
    """scale 25: momenta large enough that trajectories reflect at 0 and max_slip."""
    say("case: component leapfrog vs oracle, momentum scale", scale)
    d = setup
    nd, th, gam = d["f"].ndim, d["theta"], 0.6
    rng = np.random.default_rng(int(scale))
    n = 8
    cur = np.asfortranarray(rng.uniform(0.5, 4.0, size=(nd, n)))
    p0 = np.asfortranarray(rng.normal(size=(nd, n)) * scale)
    cov = rng.uniform(0.3, 1.5, nd)
    cand, p, hc, hn = d["pb"].slip_leapfrog(th, gam, cur, p0, cov, dtau, 50)
    nref = 0
    for k in range(n):
        oc, op, ohc, ohn = d["po"].leapfrog(th, d["G"], gam, cur[:, k], p0[:, k], cov, dtau, 50)
        say("traj", k, "ham_cur", hc[k], ohc, "ham_cand", hn[k], ohn)
        assert hc[k] == pytest.approx(ohc, rel=1e-12)
        if ohn == 1e20:
            assert hn[k] == 1e20
            np.testing.assert_array_equal(cand[:, k], cur[:, k])
            continue
        assert hn[k] == pytest.approx(ohn, rel=1e-9)
        assert np.abs(cand[:, k] - oc).max() <= 1e-9 * syn.MAX_SLIP
        assert np.abs(p[:, k] - op).max() <= 1e-9 * max(1.0, np.abs(op).max())
        assert cand[:, k].min() >= 0.0 and cand[:, k].max() <= syn.MAX_SLIP
        nref += int(np.any(np.sign(p[:, k]) != np.sign(p0[:, k])))
    if scale > 1:
        assert nref > 0  # momentum flips: the reflections were exercised


def test_gen_init_particles_vs_oracle_chain(setup):
    say("case: component prior chain vs oracle")
    d = setup
    f, th, P, nd = d["f"], d["theta"], d["P"], d["f"].ndim
    key = O.inner_key(21, 0, 0, 3)
    X, L, R = d["pb"].slip_gen_init_particles(th, key)
    bm = np.zeros((nd, P), order="F")
    for k in range(P):
        for i in range(nd):
            bm[i, k] = O.draw_normal(key, 1, 0, k, i)
    Xo = O.gibbs_init_chain(P, d["lap"]["ll"], d["alpha2"], syn.MAX_SLIP, bm)
    assert np.abs(X - Xo).max() <= 1e-10
    for k in range(0, P, 7):
        Lo = O.likelihood(X[:, k], d["H"], d["b"], d["dsd"], 0, d["obs"]["ngnss"], th[4], th[5])
        assert L[k] == pytest.approx(Lo, rel=1e-10)
        assert R[k] == pytest.approx(O.prior(X[:, k], d["alpha2"], d["lap"]["li"], d["lap"]["lv"], f.nnode),
                                     rel=1e-12)


def test_moments_vs_oracle(setup):
    say("case: component moments vs oracle")
    rng = np.random.default_rng(5)
    X = np.asfortranarray(rng.uniform(0, 5, size=(30, 700)))
    X[7, :] = 1.25  # a coordinate without spread: variance floored at 1e-3
    w = rng.uniform(0, 1, 700)
    w /= w.sum()
    m, c = setup["ctx"].slip_calc_moments(X, w)
    mo, co = O.mean_cov_quad(X, w)
    np.testing.assert_allclose(m, mo, rtol=1e-15, atol=0)
    np.testing.assert_allclose(c, co, rtol=1e-14, atol=0)
    assert c[7] == 1e-3


def _plane_frame(theta, xmin, xmax, zmin):
    """get_geometry + calc_greens_func's frame of a single plane (init.f90:121-173, gfunc.f90:283-296)."""
    xl, yl, xr, yr, dip = xmin, theta[0], xmax, theta[1], theta[2]
    dr = dip * math.pi / 180.0
    sr = math.atan2(xr - xl, yr - yl)
    xf = (xl + xr) / 2.0 - zmin * math.cos(sr) / (2.0 * math.tan(dr))
    yf = (yl + yr) / 2.0 + zmin * math.sin(sr) / (2.0 * math.tan(dr))
    return xf, yf, zmin / 2.0, sr / math.pi * 180.0, dip


def test_call_dc3d0_vs_oracle(setup):
    say("case: call_dc3d0 vs oracle")
    import ctypes as C
    L = O.lib()
    L.orc_call_dc3d0.argtypes = [C.c_double] * 10 + [C.POINTER(C.c_double)]
    rng = np.random.default_rng(8)
    rows = np.column_stack([rng.uniform(-20, 20, 40), rng.uniform(-20, 20, 40), rng.uniform(-15, -0.5, 40),
                            rng.uniform(-50, 50, 40), rng.uniform(-50, 50, 40), rng.normal(size=40),
                            rng.normal(size=40), rng.uniform(10, 89, 40), rng.uniform(1, 20, 40),
                            rng.uniform(0, 360, 40)])
    u = setup["ctx"].call_dc3d0(rows)
    for k in range(rows.shape[0]):
        out = (C.c_double * 3)()
        L.orc_call_dc3d0(*[float(v) for v in rows[k]], out)
        assert np.abs(u[k] - np.array(out[:])).max() <= 1e-12 * max(1e-30, np.abs(out[:]).max())


def test_calc_responce_reproduces_greens_columns(setup):
    """calc_responce_dist of the unit slip at each DOF (gen_unit_slip: strike -1, dip +1) projected on the
    line-of-sight vectors is that DOF's column of calc_greens_func (gfunc.f90:300-330)."""
    say("case: calc_responce vs calc_greens_func columns")
    d = setup
    pb, th, obs = d["pb"], d["theta"], d["obs"]
    xmin, xmax, zmin, _ = syn.fault_frame()
    mesh = pb.discretize_fault(th)
    xf, yf, zf, strike, dip = _plane_frame(th, xmin, xmax, zmin)
    G = pb.calc_greens_func(th)
    pts = obs["obs_points"][:, ::3]  # one point per GNSS station
    for col in range(0, G.shape[1], 3):
        idof, direction = col // 2, col % 2
        node = mesh["id_dof"][idof]
        slip = np.zeros((2, pb.nnode), order="F")
        slip[direction, node - 1] = -1.0 if direction == 0 else 1.0
        targets = mesh["n2e"][:mesh["n2es"][node - 1], node - 1]
        u = d["ctx"].calc_responce(pts, mesh["cny"], mesh["coor"], slip, xf, yf, zf, strike, dip, targets)
        g = np.zeros(G.shape[0])
        for i in range(pts.shape[1]):
            for c in range(3):
                r = 3 * i + c
                g[r] = np.dot(u[:, i], obs["obs_unitvec"][:, r])
        assert np.abs(g - G[:, col]).max() <= 1e-10 * np.abs(G[:, col]).max(), col


def test_raw_array_forms_vs_oracle(setup):
    say("case: likelihood / prior / grad from raw arrays, Gram from a given G")
    d = setup
    f, th, nd = d["f"], d["theta"], d["f"].ndim
    ctx, lap = d["ctx"], d["lap"]
    X = np.asfortranarray(np.random.default_rng(2).uniform(0, 3.0, size=(nd, 9)))
    Ld = ctx.slip_likelihood_gram(X, d["H"], d["b"], d["dsd"], 0, d["obs"]["ngnss"], th[4], th[5])
    Rd = ctx.slip_prior_ell(X, d["alpha2"], lap["li"], lap["lv"])
    Gd = ctx.slip_grad_ell(X, 0.7, d["H"], d["b"], d["alpha2"], lap["li"], lap["lv"], lap["lti"], lap["ltv"])
    for k in range(X.shape[1]):
        Lo = O.likelihood(X[:, k], d["H"], d["b"], d["dsd"], 0, d["obs"]["ngnss"], th[4], th[5])
        Ro = O.prior(X[:, k], d["alpha2"], lap["li"], lap["lv"], f.nnode)
        go = O.grad(X[:, k], f.nnode, f.ndof, lap["li"], lap["lv"], lap["lti"], lap["ltv"], 0.7, d["alpha2"], d["b"],
                    d["H"])
        assert Ld[k] == pytest.approx(Lo, rel=1e-12)
        assert Rd[k] == pytest.approx(Ro, rel=1e-13)
        assert np.abs(Gd[:, k] - go).max() <= 1e-12 * np.abs(go).max()
    H, b, dsd = d["pb"].slip_calc_gsd_gsg_gmat(d["G"], math.exp(th[4]), math.exp(th[5]))
    assert np.abs(H - d["H"]).max() <= 1e-12 * np.abs(d["H"]).max()
    assert np.abs(b - d["b"]).max() <= 1e-12 * np.abs(d["b"]).max()
    assert dsd == pytest.approx(d["dsd"], rel=1e-13)
