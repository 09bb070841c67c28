"""The run-file driver end to end (main.f90's two paths) on a run directory written in the reference's formats,
with two InSAR paths and GNSS stations.  Needs an MI355X."""
import os
import sys

import numpy as np
import pytest

import oracle as O
import smc_inversion_fort_amd as S
from smc_inversion_fort_amd import driver as D
from smc_inversion_fort_amd import synthetic as syn

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def write_run(root, obs, nxi, neta, npart_fault, npart_slip, theta):
    xmin, xmax, zmin, _ = syn.fault_frame()
    os.makedirs(os.path.join(root, "data"), exist_ok=True)
    npath = len(obs["nsar_index"]) - 1
    with open(os.path.join(root, "data", "setting.dat"), "w") as f:
        f.write("number of fault planes\n1\nfix-xbend\nF\nxbend\n\n\n"
                "number of samples for fault\n%d\nnumber of samples for slip\n%d\n\n"
                "observation path\ndata/observation.csv\nnumber of SAR paths\n%d\noutput directory path\noutput/\n\n"
                "max slip(m)\n%.17g\n\nnxi\n%d\nneta\n%d\n\nxmin\n%.17g\nxmax\n%.17g\nzmin\n%.17g\n\nGSMC\nF\n"
                % (npart_fault, npart_slip, npath, syn.MAX_SLIP, nxi, neta, xmin, xmax, zmin))
    with open(os.path.join(root, "data", "observation.csv"), "w") as f:
        f.write("x,y,ex,ey,ez,dlos,sigma,type\n")
        p = 0
        for k in range(obs["dvec"].shape[0]):
            if k < obs["nsar_total"]:
                while k + 1 >= obs["nsar_index"][p + 1]:
                    p += 1
                typ = "InSAR_%d" % p
            else:
                typ = "GNSS"
            vals = list(obs["obs_points"][:, k]) + list(obs["obs_unitvec"][:, k]) + [obs["dvec"][k], obs["obs_sigma"][k]]
            f.write(", ".join("%.17g" % v for v in vals) + ", %s\n" % typ)
    with open(os.path.join(root, "peak_fault.dat"), "w") as f:
        f.write("".join("%.17g\n" % v for v in theta))
    r = syn.prior_range()
    with open(os.path.join(root, "data", "range.dat"), "w") as f:
        for i, name in enumerate(["ymin", "ymax", "dip", "log_alpha2", "log_sigma_sar2", "log_sigma_gnss2"]):
            f.write("%s\n%.17g, %.17g\n" % (name, r[0, i], r[1, i]))


def sar_obs():
    xmin, xmax, zmin, _ = syn.fault_frame()

    def greens(th, pts, uv, nx, ne, nsar):
        return O.Fault([nx], [ne], xmin, xmax, zmin).greens(th, pts, uv, nsar, (pts.shape[1] - nsar) // 3)
    return syn.make_observations("c2", 9, greens, nsta=8, nxi=4, neta=3, npath=2, nsar_per_path=6)


def test_driver_likelihood_and_smc(tmp_path):
    say("case: run-file driver, likelihood and outer SMC with per-stage dumps")
    obs = sar_obs()
    root = str(tmp_path)
    theta = syn.theta_true()
    write_run(root, obs, 4, 3, 8, 64, theta)
    st, ob = D.load_run(root)
    assert ob["nsar_index"].tolist() == obs["nsar_index"].tolist() and ob["ngnss"] == obs["ngnss"]
    np.testing.assert_allclose(ob["sar_sinv"], obs["sar_sinv"], rtol=1e-12, atol=1e-12 * np.abs(obs["sar_sinv"]).max())
    r = D.run_likelihood(root, seed=3, mover=S.MOVER_GIBBS, sweeps=3)
    # the same evaluation through the library directly
    ctx = S.Context(0)
    xmin, xmax, zmin, _ = syn.fault_frame()
    pb = S.Problem(ctx, [4], [3], xmin, xmax, zmin, obs, syn.MAX_SLIP, 64, mover=S.MOVER_GIBBS, gibbs_sweeps=3)
    d = pb.slip_smc_exec(theta, O.inner_key(3, 0, 0, 0))
    say("driver -logZ", r["neglog"], "direct", d["neglog"])
    assert r["neglog"] == pytest.approx(d["neglog"], rel=1e-12)
    pb.close()
    ctx.close()
    lines = open(os.path.join(root, "output", "slip_from_mean_fault.dat")).read().splitlines()
    assert len(lines) == 64 and all(len(ln) == 12 * (2 * 20 + 2) for ln in lines)
    s = D.run_smc(root, seed=4, mover=S.MOVER_GIBBS, sweeps=3, max_stages=2, max_batch=16)
    say("outer stages", s["nstage"], "evidence", s["evidence"])
    for it in range(s["nstage"] + 1):
        rows = open(os.path.join(root, "output", "%d.csv" % it)).read().splitlines()
        assert len(rows) == 8 and all(len(x) == 20 * 7 for x in rows)
    last = np.array([[float(x[20 * k:20 * (k + 1)]) for k in range(7)] for x in
                     open(os.path.join(root, "output", "%d.csv" % s["nstage"])).read().splitlines()])
    np.testing.assert_allclose(last[:, :6], s["particles"].T, atol=6e-6)
    np.testing.assert_allclose(last[:, 6], s["likelihood"], atol=6e-6)
    # post-processing of the last stage's samples: calc_slip_map on the device, then point_cloud
    fault_file = os.path.join(root, "output", "%d.csv" % s["nstage"])
    slip_file = os.path.join(root, "output", "mapslip.dat")
    means, nl = D.run_slip_map(root, fault_file, slip_file, seed=5, mover=S.MOVER_GIBBS, sweeps=3, max_batch=4,
                               offset_sample=0, stride=1)
    rows = open(slip_file).read().splitlines()
    assert len(rows) == 8 and all(len(x) == 12 * 2 * 20 for x in rows)
    assert np.all(np.isfinite(nl)) and np.all(means >= 0) and np.all(means <= syn.MAX_SLIP)
    pts = D.run_point_cloud(root, fault_file, slip_file, os.path.join(root, "output", "points.dat"), offset_sample=0,
                            stride=1)
    rows = open(os.path.join(root, "output", "points.dat1")).read().splitlines()
    assert len(rows) == 8 * 12 * 64 == pts[0].shape[0] and all(len(x) == 6 * 12 for x in rows)
