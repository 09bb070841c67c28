"""InSAR paths on the device (dense per-path inverse covariance in the Gram quantities,
smc_slip.f90:1365-1409) against the oracle.  Needs an MI355X."""
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


def sar_obs(seed=5, npath=2, nper=9, nsta=6):
    xmin, xmax, zmin, _ = syn.fault_frame()

    def greens(th, pts, uv, nx, ne, nsar):
        return O.Fault([nx], [ne], xmin, xmax, zmin).greens(th, pts, uv, nsar, (pts.shape[1] - nsar) // 3)
    return syn.make_observations("c2", seed, greens, nsta=nsta, nxi=4, neta=3, npath=npath, nsar_per_path=nper)


def test_gram_with_sar_paths_vs_oracle(ctx):
    say("case: gram with 2 InSAR paths + GNSS vs oracle")
    obs = sar_obs()
    xmin, xmax, zmin, _ = syn.fault_frame()
    pb = S.Problem(ctx, [4], [3], xmin, xmax, zmin, obs, syn.MAX_SLIP, 64, mover=S.MOVER_GIBBS, max_batch=3)
    th = syn.sample_thetas(3, 12)
    H, b, dsd = pb.slip_calc_gsd_gsg(th)
    for k in range(3):
        G = O.Fault([4], [3], xmin, xmax, zmin).greens(th[:, k], obs["obs_points"], obs["obs_unitvec"],
                                                        obs["nsar_total"], obs["ngnss"])
        Hr, br, dr = O.gsd_gsg(G, obs["dvec"], obs["obs_sigma"], math.exp(th[4, k]), math.exp(th[5, k]),
                               nsar_index=obs["nsar_index"], sar_sinv=obs["sar_sinv"])
        say("theta", k, "max|dH|/max|H| %.2e" % (np.abs(H[:, :, k] - Hr).max() / np.abs(Hr).max()))
        np.testing.assert_allclose(H[:, :, k], Hr, rtol=1e-11, atol=1e-11 * np.abs(Hr).max())
        np.testing.assert_allclose(b[:, k], br, rtol=1e-11, atol=1e-11 * np.abs(br).max())
        assert dsd[k] == pytest.approx(dr, rel=1e-11)
    pb.close()


def test_fault_likelihood_with_sar_vs_oracle(ctx):
    say("case: fault likelihood with InSAR + GNSS vs oracle")
    obs = sar_obs(seed=6)
    xmin, xmax, zmin, _ = syn.fault_frame()
    P = 128
    pb = S.Problem(ctx, [4], [3], xmin, xmax, zmin, obs, syn.MAX_SLIP, P, mover=S.MOVER_GIBBS, gibbs_sweeps=3,
                   max_batch=3)
    po = O.ProblemHolder(O.Fault([4], [3], xmin, xmax, zmin), obs, syn.MAX_SLIP, P, mover=1, gibbs_sweeps=3)
    th = syn.sample_thetas(3, 22)
    keys = [O.inner_key(11, 0, 0, i) for i in range(3)]
    nl, ns = pb.fault_calc_likelihood(th, keys)
    for k in range(3):
        ref, nst = po.fault_likelihood(th[:, k], keys[k])
        say("theta", k, nl[k], ref, ns[k], nst)
        assert ns[k] == nst
        assert nl[k] == pytest.approx(ref, rel=1e-8)
    pb.close()
