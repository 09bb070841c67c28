"""Mixed-precision Gibbs movers against the FP64 oracle and the FP64 device path.

SMCF_PRECISION_MIXED computes the Gibbs products Q S in FP32 on the matrix cores;
SMCF_PRECISION_MIXED_BF16 forms them from split-BF16 operands (hi.hi + hi.lo + lo.hi, FP32
accumulation).  Both draw the truncated normals in single precision; conditional means, corrections,
likelihood, weights and evidence stay FP64.  It is a different rounding of the same sampler, so it is compared statistically: -log Z and
posterior means over independent keys, within Monte Carlo error plus the stated tolerance (posterior
means within 1% of the largest FP64 mean).  Each test prints the measured error.  Needs an MI355X."""
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


def small_obs(seed=3, nsta=12, nxi=4, neta=3, sigma=0.5):
    def greens(th, pts, uv, nx, ne):
        xmin, xmax, zmin, _ = syn.fault_frame()
        return O.Fault([nx], [ne], xmin, xmax, zmin).greens(th, pts, uv, 0, pts.shape[1] // 3)
    return syn.make_observations("c2", seed, greens, nsta=nsta, sigma=sigma, nxi=nxi, neta=neta)


def device_problem(ctx, obs, P, precision, sweeps=3, max_batch=4):
    xmin, xmax, zmin, _ = syn.fault_frame()
    return S.Problem(ctx, [obs["nxi"]], [obs["neta"]], xmin, xmax, zmin, obs, syn.MAX_SLIP, P,
                     mover=S.MOVER_GIBBS, gibbs_sweeps=sweeps, max_batch=max_batch, precision=precision)


def test_trunc_normal_f32_vs_oracle(ctx):
    """The device's single-precision draw against the oracle's restatement: same declines (NaN),
    values within single-precision resolution of the standardised quantile."""
    say("case: single-precision truncated normal vs oracle")
    rng = np.random.default_rng(11)
    n = 20000
    mu = rng.normal(2.0, 6.0, n) * np.exp(rng.uniform(-3, 1, n))
    sq = np.exp(rng.uniform(-3, 4, n))
    u = ((rng.integers(0, 2 ** 23, n).astype(np.float64) + 0.5) * 2.0 ** -23).astype(np.float32)
    x = ctx.trunc_normal_f32(mu, sq, 0.0, 10.0, u)
    xo = np.array([O.trunc_normal_prec_f32(mu[i], sq[i], 0.0, 10.0, float(u[i])) for i in range(n)])
    nan_d, nan_o = np.isnan(x), np.isnan(xo)
    np.testing.assert_array_equal(nan_d, nan_o)
    ok = ~nan_d
    say("accepted", int(ok.sum()), "of", n)
    assert ok.sum() > n // 2
    assert np.all((x[ok] >= 0) & (x[ok] <= 10))
    dz = np.abs(x[ok] - xo[ok]) * sq[ok]
    zmag = np.maximum(1.0, np.abs((xo[ok] - mu[ok]) * sq[ok]))
    say("max |dz| / max(1,|z|):", float((dz / zmag).max()))
    assert np.all(dz <= 2e-6 * zmag)


MIXED_FORMS = [pytest.param(S.PRECISION_MIXED, id="fp32"), pytest.param(S.PRECISION_MIXED_BF16, id="bf16x3")]


@pytest.mark.parametrize("precision", MIXED_FORMS)
def test_mixed_slip_smc_vs_fp64_oracle(ctx, precision):
    """Inner SMC on a small mesh, 24 keys: mixed device vs FP64 oracle.  The first gamma is identical
    (FP64 likelihood of the same prior particles); -log Z and the posterior means agree within Monte
    Carlo error, posterior means also within 1% of the largest FP64 mean."""
    say("case: mixed inner SMC (precision %d) vs FP64 oracle, 24 keys" % precision)
    obs = small_obs()
    P = 256
    pb = device_problem(ctx, obs, P, precision)
    xmin, xmax, zmin, _ = syn.fault_frame()
    f = O.Fault([obs["nxi"]], [obs["neta"]], xmin, xmax, zmin)
    po = O.ProblemHolder(f, obs, syn.MAX_SLIP, P, mover=1, gibbs_sweeps=3)
    theta = syn.theta_true()
    G = pb.calc_greens_func(theta)
    nl_d, nl_o, m_d, m_o, s_d, s_o = [], [], [], [], [], []
    for k in range(24):
        key = O.inner_key(40 + k, 1, 1, 3)
        d = pb.slip_smc_exec(theta, key)
        o = po.slip_smc(theta, G, key)
        assert d["gamma"][0] == pytest.approx(o["gamma"][0], rel=0, abs=1e-9)
        nl_d.append(d["neglog"]); nl_o.append(o["neglog"])
        m_d.append(d["mean"]); m_o.append(o["mean"])
        s_d.append(d["particles"].std(axis=1)); s_o.append(o["particles"].std(axis=1))
    nl_d, nl_o, m_d, m_o, s_d, s_o = map(np.array, (nl_d, nl_o, m_d, m_o, s_d, s_o))
    dl = nl_d - nl_o
    se = dl.std(ddof=1) / np.sqrt(len(dl))
    say("-log Z: mixed %.5f fp64 %.5f  diff %.5f  se %.5f" % (nl_d.mean(), nl_o.mean(), dl.mean(), se))
    assert abs(dl.mean()) <= 4 * se + 0.02
    dm = m_d.mean(0) - m_o.mean(0)
    sem = (m_d - m_o).std(0, ddof=1) / np.sqrt(len(m_d))
    scale = np.abs(m_o.mean(0)).max()
    say("posterior mean: max |diff| %.3e (%.3f%% of max mean), max |diff|/se %.2f"
        % (np.abs(dm).max(), 100 * np.abs(dm).max() / scale, float((np.abs(dm) / np.maximum(sem, 1e-300)).max())))
    assert np.all(np.abs(dm) <= 4 * sem + 0.01 * scale)
    # posterior standard deviations (particle spread per coordinate), averaged over the keys
    ds = s_d.mean(0) - s_o.mean(0)
    ses = (s_d - s_o).std(0, ddof=1) / np.sqrt(len(s_d))
    sscale = s_o.mean(0).max()
    say("posterior std: max |diff| %.3e (%.3f%% of max std), max |diff|/se %.2f"
        % (np.abs(ds).max(), 100 * np.abs(ds).max() / sscale, float((np.abs(ds) / np.maximum(ses, 1e-300)).max())))
    assert np.all(np.abs(ds) <= 4 * ses + 0.01 * sscale)
    pb.close()


@pytest.mark.parametrize("precision", MIXED_FORMS)
def test_mixed_streamed_slots_match_one_wide_batch(ctx, precision):
    """The mixed paths are deterministic per (geometry, key) whatever the slot count."""
    say("case: mixed (precision %d) streamed slots vs wide batch" % precision)
    obs = small_obs(seed=8)
    th = syn.sample_thetas(8, 31)
    keys = [O.inner_key(5, 0, 0, i) for i in range(8)]
    narrow = device_problem(ctx, obs, 128, precision, max_batch=3)
    wide = device_problem(ctx, obs, 128, precision, max_batch=8)
    a, na = narrow.fault_calc_likelihood(th, keys)
    b, nb = wide.fault_calc_likelihood(th, keys)
    say("stages", na.tolist(), "neglog", a.tolist())
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(na, nb)
    narrow.close()
    wide.close()


def test_mixed_c2_vs_fp64_device():
    """Benchmark size (16x16 patches, 400 stations, P = 2000, 5 sweeps): each mixed form vs the FP64
    device path on 8 geometries.  The stage counts and -log Z agree within Monte Carlo spread; the
    posterior means at the true geometry, averaged over 6 keys, agree within Monte Carlo error plus 1% of
    the largest mean."""
    say("case: mixed forms vs fp64 at 16x16 patches")
    ctx = S.Context(0)
    xmin, xmax, zmin, _ = syn.fault_frame()
    cfg = syn.CONFIGS["c2"]

    def greens(th, pts, uv, nx, ne):
        obs0 = dict(obs_points=pts, obs_unitvec=uv, obs_sigma=np.ones(pts.shape[1]), dvec=np.zeros(pts.shape[1]),
                    nsar_total=0, ngnss=pts.shape[1] // 3)
        p = S.Problem(ctx, [nx], [ne], xmin, xmax, zmin, obs0, syn.MAX_SLIP, 8)
        g = p.calc_greens_func(th)
        p.close()
        return g
    obs = syn.make_observations("c2", 0, greens)
    precs = (S.PRECISION_FP64, S.PRECISION_MIXED, S.PRECISION_MIXED_BF16)
    pbs = {prec: S.Problem(ctx, [cfg["nxi"]], [cfg["neta"]], xmin, xmax, zmin, obs, syn.MAX_SLIP, cfg["P"],
                           mover=S.MOVER_GIBBS, gibbs_sweeps=cfg["sweeps"], max_batch=8, precision=prec)
           for prec in precs}
    th = syn.sample_thetas(8, 3)
    keys = [O.inner_key(2, 0, 0, i) for i in range(8)]
    res = {prec: pbs[prec].fault_calc_likelihood(th, keys) for prec in precs}
    nl64, ns64 = res[S.PRECISION_FP64]
    say("fp64 -logZ", np.round(nl64, 3).tolist(), "stages", ns64.tolist())
    t0 = syn.theta_true()
    K = 6
    mk = {prec: np.array([pbs[prec].slip_smc_exec(t0, O.inner_key(9, 0, 0, k))["mean"] for k in range(K)])
          for prec in precs}
    m64 = mk[S.PRECISION_FP64]
    scale = np.abs(m64.mean(0)).max()
    say("posterior mean, one key: fp64 key0 vs key1 max |diff| = %.2f%% of max mean (Monte Carlo spread)"
        % (100 * np.abs(m64[0] - m64[1]).max() / scale))
    for prec in precs[1:]:
        nlm, nsm = res[prec]
        say("precision %d -logZ" % prec, np.round(nlm, 3).tolist(), "stages", nsm.tolist())
        rel = np.abs(nlm - nl64) / np.abs(nl64)
        say("-log Z relative difference: max %.2e" % rel.max())
        assert rel.max() < 1e-3
        assert np.all(np.abs(nsm - ns64) <= 3)
        # posterior means at the true geometry over 6 keys: the difference of the key averages against its
        # Monte Carlo standard error, plus 1% of the largest mean (the stated tolerance)
        mmx = mk[prec]
        diff = mmx.mean(0) - m64.mean(0)
        se = np.sqrt(m64.var(0, ddof=1) / K + mmx.var(0, ddof=1) / K)
        say("posterior mean, %d-key averages: max |mixed - fp64| = %.3e = %.3f%% of max mean; max |diff|/se = %.2f"
            % (K, np.abs(diff).max(), 100 * np.abs(diff).max() / scale, float((np.abs(diff) / se).max())))
        # 5 standard errors: the largest of 480 coordinates' |N(0, 1)| exceeds 5 with probability 3e-4
        assert np.all(np.abs(diff) <= 5 * se + 0.01 * scale)
    for p in pbs.values():
        p.close()
    ctx.close()
