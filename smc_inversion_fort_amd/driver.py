"""Run files of the reference and its driver, over the device library.

Reads the reference's inputs and writes its outputs in the reference's formats:

  data/setting.dat        read_settings      main.f90:107-157
  observation file        read_observation   init.f90:5-80 (read_observation1 / read_observation2)
  data/range.dat          read_range         main.f90:554-560
  peak_fault.dat          read_values        main.f90:455-459
  data/qmean.dat, qcov.dat, qcov_inv.dat, log_qcov_det.dat (gsmc)   main.f90:309-347
  slip particles (flag_output = 1)   write_slip_particles   smc_slip.f90:1632-1660
  <output_dir><iter>.csv per outer stage   written by the library (smcf_fault_smc_set_output_dir),
                                           smc_fault.f90:923-936, 1054-1067

and runs either of main.f90's two paths:

  python -m smc_inversion_fort_amd.driver likelihood [--dir D]   fault_calc_likelihood at peak_fault.dat with
      the slip particles written to output/slip_from_mean_fault.dat (main.f90:447-483, the active path);
  python -m smc_inversion_fort_amd.driver smc [--dir D]          fault_smc_exec over the prior box of
      data/range.dat, plain or gsmc as setting.dat says (main.f90:513-549);
  python -m smc_inversion_fort_amd.driver slipmap --fault-file F  calc_slip_map (main.f90:553-666);
  python -m smc_inversion_fort_amd.driver points --fault-file F   point_cloud (main.f90:668-830).

Files are read with the semantics of Fortran list-directed input: every READ starts at a new record, its
items may continue over following records (blank records are skipped while values are sought), the rest
of its last record is discarded, values are separated by blanks or commas, and '/' ends a READ early.
"""
import argparse
import math
import os
import re
import sys

import numpy as np


# ------------------------------------------------------------------ list-directed input
class Records:
    """The records of a formatted sequential file, read the way READ(u, *) and READ(u, "(a)") read them."""

    def __init__(self, path):
        with open(path) as f:
            self.lines = f.read().splitlines()
        self.pos = 0
        self.path = path

    def _next_record(self):
        if self.pos >= len(self.lines):
            raise EOFError("%s: end of file" % self.path)
        rec = self.lines[self.pos]
        self.pos += 1
        return rec

    def skip(self):
        """READ(u, *) with no items: the next record is read and discarded."""
        self._next_record()

    def line(self):
        """READ(u, "(a)") into a character variable: the whole next record, trailing blanks removed."""
        return self._next_record().rstrip()

    def values(self, n):
        """READ(u, *) of n items: tokens from as many records as needed (the rest of the last one dropped)."""
        out = []
        while len(out) < n:
            rec = self._next_record()
            for tok in _tokens(rec):
                if tok == "/":
                    return out + [None] * (n - len(out))
                m = re.fullmatch(r"(\d+)\*(.*)", tok)
                if m:
                    out += [m.group(2) or None] * int(m.group(1))
                else:
                    out.append(tok)
                if len(out) >= n:
                    break
        return out[:n]


def _tokens(rec):
    toks, cur, quote = [], "", None
    for ch in rec:
        if quote:
            if ch == quote:
                quote = None
            else:
                cur += ch
        elif ch in "'\"":
            quote = ch
        elif ch in " \t,":
            if cur:
                toks.append(cur)
            cur = ""
        elif ch == "/":
            if cur:
                toks.append(cur)
            toks.append("/")
            return toks
        else:
            cur += ch
    if cur:
        toks.append(cur)
    return toks


def _int(t):
    return int(t)


def _real(t):
    return float(t.replace("d", "e").replace("D", "e"))


def _logical(t):
    s = t.lstrip(".").upper()
    if not s or s[0] not in "TF":
        raise ValueError("not a logical value: %r" % t)
    return s[0] == "T"


# ------------------------------------------------------------------ formatted output
def fortran_f(x, w, d):
    """Fortran Fw.d editing: right-justified in w columns; w asterisks when the value does not fit."""
    s = "%.*f" % (d, x)
    if x != x or len(s) > w:
        return "*" * w
    return s.rjust(w)


# ------------------------------------------------------------------ input files
def read_settings(path):
    """data/setting.dat (main.f90:107-157): a comment record before every value, blank separators."""
    r = Records(path)
    st = {}
    r.skip()
    st["nplane"] = _int(r.values(1)[0])
    n = st["nplane"]
    r.skip()
    st["fix_xbend"] = _logical(r.values(1)[0])
    r.skip()
    xb = r.values(n - 1) if n > 1 else (r.skip() or [])
    st["xbend"] = [_real(t) for t in xb]
    r.skip()
    r.skip()
    st["nparticle_fault"] = _int(r.values(1)[0])
    r.skip()
    st["nparticle_slip"] = _int(r.values(1)[0])
    r.skip()
    r.skip()
    st["observation_path"] = r.line()
    r.skip()
    st["npath"] = _int(r.values(1)[0])
    r.skip()
    st["output_dir"] = r.line()
    r.skip()
    r.skip()
    st["max_slip"] = _real(r.values(1)[0])
    r.skip()
    r.skip()
    st["nxi_ls"] = [_int(t) for t in r.values(n)]
    r.skip()
    st["neta_ls"] = [_int(t) for t in r.values(n)]
    r.skip()
    r.skip()
    st["xmin"] = _real(r.values(1)[0])
    r.skip()
    st["xmax"] = _real(r.values(1)[0])
    r.skip()
    st["zmin"] = _real(r.values(1)[0])
    r.skip()
    r.skip()
    st["gsmc"] = _logical(r.values(1)[0])
    return st


def read_observation(path, npath):
    """Observation file (init.f90:5-80): a header record, then one record per observation row
    "x, y, ex, ey, ez, dlos, sigma, type".  Rows whose type starts with "InSAR" form the paths: a change
    of type starts the next path (the type of the first row seeds the comparison, so InSAR rows are
    expected first, as the reference assumes); rows of type "GNSS" are counted and divided by 3.
    The number of rows is the number of non-blank records after the header (read_observation1)."""
    with open(path) as f:
        nonblank = [ln for ln in f.read().splitlines() if ln.strip()]
    nobs = len(nonblank) - 1
    r = Records(path)
    r.values(1)  # header record: READ(10, *) buf
    pts = np.zeros((2, nobs), order="F")
    uv = np.zeros((3, nobs), order="F")
    sig = np.zeros(nobs)
    d = np.zeros(nobs)
    nsar_ls = [0] * max(npath, 1)
    ipath = 0
    ngnss = 0
    type_prev = None
    for i in range(nobs):
        v = r.values(8)
        pts[0, i], pts[1, i] = _real(v[0]), _real(v[1])
        uv[:, i] = [_real(v[2]), _real(v[3]), _real(v[4])]
        d[i], sig[i] = _real(v[5]), _real(v[6])
        typ = (v[7] or "")[:20]
        if i == 0:
            type_prev = typ
        if typ[:5] == "InSAR":
            if typ == type_prev:
                nsar_ls[ipath] += 1
            else:
                ipath += 1
                if ipath >= npath:
                    raise ValueError("%s: more InSAR paths than npath = %d" % (path, npath))
                nsar_ls[ipath] = 1
            type_prev = typ
        if typ == "GNSS":
            ngnss += 1
    nsar_index = [1]
    for p in range(npath):
        nsar_index.append(nsar_index[-1] + nsar_ls[p])
    return dict(obs_points=pts, obs_unitvec=uv, obs_sigma=sig, dvec=d, nsar_total=nsar_index[-1] - 1,
                ngnss=ngnss // 3, nsar_index=np.array(nsar_index, np.int32), nsar_ls=nsar_ls[:npath])


def read_range(path, ndim_fault):
    """data/range.dat (main.f90:554-560): per parameter a label record, then "lo, hi"; returns 2 x ndim."""
    r = Records(path)
    rng = np.zeros((2, ndim_fault), order="F")
    for i in range(ndim_fault):
        r.line()
        lo, hi = r.values(2)
        rng[:, i] = (_real(lo), _real(hi))
    return rng


def read_values(path, n, one_per_read=False):
    """n reals: READ(u, *) x(1:n) in one statement (one_per_read False) or n statements (peak_fault.dat)."""
    r = Records(path)
    if one_per_read:
        return np.array([_real(r.values(1)[0]) for _ in range(n)])
    return np.array([_real(t) for t in r.values(n)])


# ------------------------------------------------------------------ mesh bookkeeping
def id_dof(nxi_ls, neta_ls):
    """DOF -> node map of discretize_fault (init.f90:226-267), 1-based node ids: every node except the bottom
    row, the first plane's left column and the last plane's right column."""
    out, offset = [], 0
    nplane = len(nxi_ls)
    for ip, (nxi, neta) in enumerate(zip(nxi_ls, neta_ls)):
        for j in range(1, neta + 2):
            for i in range(1, nxi + 2):
                if j == 1 or (ip == 0 and i == 1) or (ip == nplane - 1 and i == nxi + 1):
                    continue
                out.append(i + (nxi + 1) * (j - 1) + offset)
        offset += (nxi + 1) * (neta + 1)
    return np.array(out, np.int32)


def write_slip_particles(path, particles, likelihood, prior, nxi_ls, neta_ls):
    """flag_output = 1 of slip_smc_exec (smc_slip.f90:1632-1660): per particle the two slip components of
    every node (zero off the DOF nodes), its likelihood and prior, each as F12.5."""
    ids = id_dof(nxi_ls, neta_ls)
    nnode = sum((a + 1) * (b + 1) for a, b in zip(nxi_ls, neta_ls))
    slip = np.zeros((2, nnode))
    with open(path, "w") as f:
        for p in range(particles.shape[1]):
            slip[0, ids - 1] = particles[0::2, p]
            slip[1, ids - 1] = particles[1::2, p]
            f.write("".join(fortran_f(slip[k, n], 12, 5) for n in range(nnode) for k in range(2)))
            f.write(fortran_f(likelihood[p], 12, 5) + fortran_f(prior[p], 12, 5) + "\n")


# ------------------------------------------------------------------ driver
def load_run(workdir):
    """Settings, observations and the InSAR covariance of a run directory (main.f90:99-307)."""
    from .sar import calc_sigma_sar_mat
    st = read_settings(os.path.join(workdir, "data", "setting.dat"))
    obs = read_observation(os.path.join(workdir, st["observation_path"]), st["npath"])
    if obs["nsar_total"] > 0:
        obs["sar_sinv"] = calc_sigma_sar_mat(obs["nsar_index"], obs["obs_points"], obs["obs_sigma"])
    n = st["nplane"]
    st["ndim_fault"] = 2 * n + 4 if st["fix_xbend"] else 3 * n + 3
    return st, obs


def make_problem(ctx, st, obs, mover, sweeps, max_batch, precision):
    from . import Problem
    return Problem(ctx, st["nxi_ls"], st["neta_ls"], st["xmin"], st["xmax"], st["zmin"], obs, st["max_slip"],
                   st["nparticle_slip"], mover=mover, gibbs_sweeps=sweeps, max_batch=max_batch,
                   fix_xbend=st["fix_xbend"], xbend=st["xbend"], precision=precision)


def run_likelihood(workdir, seed=0, mover=0, sweeps=5, precision=0, out=sys.stdout):
    """main.f90:447-483: -log Z at peak_fault.dat, slip particles to output/slip_from_mean_fault.dat."""
    from . import Context
    from . import lib
    st, obs = load_run(workdir)
    theta = read_values(os.path.join(workdir, "peak_fault.dat"), st["ndim_fault"], one_per_read=True)
    ctx = Context(0)
    pb = make_problem(ctx, st, obs, mover, sweeps, 1, precision)
    key = int(lib().smcf_inner_key(seed, 0, 0, 0))
    r = pb.slip_smc_exec(theta, key)
    os.makedirs(os.path.join(workdir, "output"), exist_ok=True)
    write_slip_particles(os.path.join(workdir, "output", "slip_from_mean_fault.dat"), r["particles"],
                         r["likelihood"], r["prior"], st["nxi_ls"], st["neta_ls"])
    print("neglog: ", r["neglog"], file=out)
    print("Ndim slip: ", pb.ndim, file=out)
    print("ABIC: ", 2.0 * r["neglog"] + 2.0 * st["ndim_fault"], file=out)
    pb.close()
    ctx.close()
    return r


def run_smc(workdir, seed=0, mover=0, sweeps=5, precision=0, max_stages=100, max_batch=256, out=sys.stdout):
    """main.f90:513-549 (fault_smc_exec): outer SMC over data/range.dat with per-stage dumps in output_dir."""
    from . import Context
    st, obs = load_run(workdir)
    nd = st["ndim_fault"]
    rng = read_range(os.path.join(workdir, "data", "range.dat"), nd)
    gsmc = None
    if st["gsmc"]:
        dd = os.path.join(workdir, "data")
        qmean = read_values(os.path.join(dd, "qmean.dat"), nd)
        qcov = read_values(os.path.join(dd, "qcov.dat"), nd * nd).reshape(nd, nd, order="F")
        qinv = read_values(os.path.join(dd, "qcov_inv.dat"), nd * nd).reshape(nd, nd, order="F")
        logdet = read_values(os.path.join(dd, "log_qcov_det.dat"), 1)[0]
        gsmc = (qmean, np.asfortranarray(np.linalg.cholesky(qcov)), np.asfortranarray(qinv), float(logdet))
    out_dir = os.path.join(workdir, st["output_dir"])
    os.makedirs(out_dir, exist_ok=True)
    ctx = Context(0)
    pb = make_problem(ctx, st, obs, mover, sweeps, max_batch, precision)
    prefix = out_dir if out_dir.endswith("/") else out_dir + "/"
    pb.set_output_dir(prefix)
    r = pb.fault_smc_exec(rng, st["nparticle_fault"], seed, max_stages=max_stages, gsmc=gsmc)
    print("stages: ", r["nstage"], file=out)
    print("evidence: ", r["evidence"], file=out)
    pb.set_output_dir(None)
    pb.close()
    ctx.close()
    return r


def run_slip_map(workdir, fault_path, out_path, seed=0, mover=0, sweeps=5, precision=0, max_batch=256,
                 offset_sample=10, stride=2, out=sys.stdout):
    """calc_slip_map (main.f90:553-666): geometry samples of an outer-SMC stage file (offset_sample records
    skipped, then every stride-th record, nparticle_fault of them), the inner SMC of each on the device, the plain
    mean of its final slip particles on every node written to out_path."""
    from . import Context
    from . import lib
    from . import postprocess as PP
    st, obs = load_run(workdir)
    n = st["nparticle_fault"]
    th = PP.read_samples(fault_path, n, st["ndim_fault"] + 1, offset_sample, stride)[:st["ndim_fault"]]
    keys = [int(lib().smcf_inner_key(seed, 0, 2, i)) for i in range(n)]
    ctx = Context(0)
    pb = make_problem(ctx, st, obs, mover, sweeps, min(max_batch, n), precision)
    means, nl = PP.calc_slip_map(pb, th, keys, out_path, st["nxi_ls"], st["neta_ls"])
    print("slip map of %d geometries written to %s" % (n, out_path), file=out)
    pb.close()
    ctx.close()
    return means, nl


def run_point_cloud(workdir, fault_path, slip_path, out_prefix, offset_sample=10, stride=100, layout="current",
                    out=sys.stdout):
    """point_cloud (main.f90:668-830): geometry samples (every stride-th record after offset_sample) and their
    mean slips (calc_slip_map's records), 8 x 8 points per patch, one file per plane (<out_prefix><plane>)."""
    from . import postprocess as PP
    st = read_settings(os.path.join(workdir, "data", "setting.dat"))
    n_pl = st["nplane"]
    nd = 2 * n_pl + 4 if st["fix_xbend"] else 3 * n_pl + 3
    n = st["nparticle_fault"]
    th = PP.read_samples(fault_path, n, nd + 1, offset_sample, stride)[:nd]
    nnode = sum((a + 1) * (b + 1) for a, b in zip(st["nxi_ls"], st["neta_ls"]))
    slips = PP.read_slip_map(slip_path, n, nnode)
    pts = PP.point_cloud(th, slips, st["nxi_ls"], st["neta_ls"], st["xmin"], st["xmax"], st["zmin"], st["fix_xbend"],
                         st["xbend"], layout=layout)
    PP.write_point_cloud(pts, out_prefix)
    print("point clouds of %d geometries written to %s<plane>" % (n, out_prefix), file=out)
    return pts


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("mode", choices=["likelihood", "smc", "slipmap", "points"])
    ap.add_argument("--fault-file", help="outer-SMC stage file of geometry samples (slipmap, points)")
    ap.add_argument("--slip-file", default="output/more/mapslip.dat", help="slip map written (slipmap) or read "
                    "(points)")
    ap.add_argument("--points-prefix", default="output/shuf/points.dat", help="point files <prefix><plane>")
    ap.add_argument("--dir", default=".", help="run directory holding data/ (the reference's working directory)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--mover", choices=["hmc", "gibbs"], default="hmc")
    ap.add_argument("--sweeps", type=int, default=5)
    ap.add_argument("--precision", choices=["fp64", "mixed", "bf16"], default="fp64",
                    help="Gibbs arithmetic: fp64, mixed (FP32 products) or bf16 (split-BF16 products)")
    ap.add_argument("--max-stages", type=int, default=100)
    ap.add_argument("--slots", type=int, default=256, help="device slots (inner SMC instances in flight)")
    a = ap.parse_args(argv)
    mover = 0 if a.mover == "hmc" else 1
    prec = {"fp64": 0, "mixed": 1, "bf16": 2}[a.precision]
    if a.mode == "likelihood":
        run_likelihood(a.dir, a.seed, mover, a.sweeps, prec)
    elif a.mode == "smc":
        run_smc(a.dir, a.seed, mover, a.sweeps, prec, a.max_stages, a.slots)
    elif a.mode == "slipmap":
        out_path = os.path.join(a.dir, a.slip_file)
        os.makedirs(os.path.dirname(out_path) or ".", exist_ok=True)
        run_slip_map(a.dir, a.fault_file, out_path, a.seed, mover, a.sweeps, prec, a.slots)
    else:
        prefix = os.path.join(a.dir, a.points_prefix)
        os.makedirs(os.path.dirname(prefix) or ".", exist_ok=True)
        run_point_cloud(a.dir, a.fault_file, os.path.join(a.dir, a.slip_file), prefix)


if __name__ == "__main__":
    main()
