"""calc_slip_map on the device (smcf_slip_mean_batch): the plain mean of each geometry's final slip particles
against the oracle's inner SMC, and the map file.  Needs an MI355X."""
import sys

import numpy as np
import pytest

import oracle as O
import smc_inversion_fort_amd as S
from smc_inversion_fort_amd import postprocess as PP
from smc_inversion_fort_amd import synthetic as syn

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_slip_map_vs_oracle(tmp_path):
    say("case: calc_slip_map, 5 geometries in batches of 2, P = 64")
    xmin, xmax, zmin, _ = syn.fault_frame()
    nxi, neta, P = [4], [3], 64

    def greens(th, pts, uv, nx, ne):
        return O.Fault([nx], [ne], xmin, xmax, zmin).greens(th, pts, uv, 0, pts.shape[1] // 3)
    obs = syn.make_observations("c2", 4, greens, nsta=12, nxi=nxi[0], neta=neta[0])
    ctx = S.Context(0)
    pb = S.Problem(ctx, nxi, neta, xmin, xmax, zmin, obs, syn.MAX_SLIP, P, mover=S.MOVER_GIBBS, gibbs_sweeps=3,
                   max_batch=2)
    th = syn.sample_thetas(5, 17)
    keys = [O.inner_key(3, 0, 2, i) for i in range(5)]
    path = tmp_path / "mapslip.dat"
    means, nl = PP.calc_slip_map(pb, th, keys, str(path), nxi, neta)
    f = O.Fault(nxi, neta, xmin, xmax, zmin)
    po = O.ProblemHolder(f, obs, syn.MAX_SLIP, P, mover=1, gibbs_sweeps=3)
    for i in range(5):
        G = f.greens(th[:, i], obs["obs_points"], obs["obs_unitvec"], 0, obs["ngnss"])
        o = po.slip_smc(th[:, i], G, keys[i])
        ref = o["particles"].mean(axis=1)
        say("geometry %d: max |mean - oracle| %.2e, -log Z %.6f vs %.6f" % (i, np.abs(means[:, i] - ref).max(),
                                                                          nl[i], o["neglog"]))
        np.testing.assert_allclose(means[:, i], ref, rtol=0, atol=1e-7)
        assert nl[i] == pytest.approx(o["neglog"], rel=1e-8)
    back = PP.read_slip_map(str(path), 5, f.nnode)
    for i in range(5):
        np.testing.assert_allclose(back[:, i].reshape(-1, 2).T, PP.node_slip(means[:, i], nxi, neta), atol=5e-6)
    pb.close()
    ctx.close()
