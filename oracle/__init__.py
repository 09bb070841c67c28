"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the timed CPU
baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
import this package; the product never does.

Arrays follow the reference: Fortran (column-major) layout, 1-based index values.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

D = C.POINTER(C.c_double)
I = C.POINTER(C.c_int)


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = C.CDLL(path)
        _declare(_LIB)
    return _LIB


class FaultDesc(C.Structure):
    _fields_ = [("nplane", C.c_int), ("nxi_ls", I), ("neta_ls", I), ("xmin", C.c_double),
                ("xmax", C.c_double), ("zmin", C.c_double), ("fix_xbend", C.c_int), ("xbend", D)]


class Problem(C.Structure):
    _fields_ = [("nobs", C.c_int), ("nsar_total", C.c_int), ("ngnss", C.c_int), ("npath", C.c_int),
                ("nsar_index", I), ("obs_points", D), ("obs_unitvec", D), ("obs_sigma", D),
                ("dvec", D), ("sar_sinv", D), ("fault", FaultDesc), ("max_slip", C.c_double),
                ("nparticle_slip", C.c_int), ("mover", C.c_int), ("gibbs_sweeps", C.c_int),
                ("hmc_ntau", C.c_int), ("hmc_accept", C.c_double), ("stage_cap", C.c_int), ("test_like", D)]


class SlipResult(C.Structure):
    _fields_ = [("nstage", C.c_int), ("neglog", C.c_double), ("gamma_trace", D), ("evid_trace", D),
                ("accept_trace", D), ("dtau_trace", D)]


class FaultResult(C.Structure):
    _fields_ = [("nstage", C.c_int), ("evidence", C.c_double), ("gamma_trace", D)]


class Gsmc(C.Structure):
    _fields_ = [("qmean", D), ("qcov_chol", D), ("qcov_inv", D), ("log_det", C.c_double)]


def _declare(L):
    L.orc_philox4x32.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    L.orc_draw_u64x2.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                 C.POINTER(C.c_uint64)]
    L.orc_inner_key.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32]
    L.orc_inner_key.restype = C.c_uint64
    L.orc_splitmix64.argtypes = [C.c_uint64]
    L.orc_splitmix64.restype = C.c_uint64
    L.orc_dc3d0.argtypes = [C.c_double] * 10 + [D, I]
    L.orc_mesh_sizes.argtypes = [C.POINTER(FaultDesc), I, I, I, I]
    L.orc_discretize_fault.argtypes = [C.POINTER(FaultDesc), D, I, D, I, I, I]
    L.orc_gen_laplacian.argtypes = [C.POINTER(FaultDesc), D, C.c_int, I, C.c_int, D, D]
    L.orc_gen_sparse_lmat.argtypes = [D, C.c_int, C.c_int, I, D, I, D]
    L.orc_calc_ll.argtypes = [D, C.c_int, C.c_int, D]
    L.orc_calc_greens_func.argtypes = [C.POINTER(FaultDesc), D, C.c_int, C.c_int, C.c_int, D, D, D]
    L.orc_slip_calc_gsd_gsg.argtypes = [D, D, C.c_int, C.c_int, C.c_int, I, C.c_int, D, C.c_double,
                                        C.c_double, D, D, D, D]
    L.orc_slip_calc_likelihood.argtypes = [D, C.c_int, D, D, C.c_double, C.c_int, C.c_int, C.c_double,
                                           C.c_double]
    L.orc_slip_calc_likelihood.restype = C.c_double
    L.orc_slip_calc_prior.argtypes = [D, D, I, D, C.c_int, D]
    L.orc_slip_calc_prior.restype = C.c_double
    L.orc_slip_calc_grad.argtypes = [D, D, C.c_int, C.c_int, C.c_int, I, D, I, D, C.c_double, D, D, D, D]
    L.orc_slip_gibbs_init_chain.argtypes = [D, C.c_int, C.c_int, D, D, C.c_double, D]
    L.orc_find_next_gamma.argtypes = [C.c_double, D, C.c_int, D, D, C.c_double, C.c_double, C.c_int]
    L.orc_find_next_gamma.restype = C.c_double
    L.orc_slip_normalize_weights.argtypes = [D, C.c_int]
    L.orc_slip_calc_mean_particles.argtypes = [D, D, D, C.c_int, C.c_int]
    L.orc_slip_calc_cov_particles.argtypes = [D, D, D, D, C.c_int, C.c_int]
    L.orc_resample_quad.argtypes = [C.c_int, D, C.c_double, C.c_double, I]
    L.orc_ndtri.argtypes = [C.c_double]
    L.orc_ndtri.restype = C.c_double
    L.orc_trunc_normal.argtypes = [C.c_double] * 5
    L.orc_trunc_normal_prec.argtypes = [C.c_double] * 5
    L.orc_trunc_normal_prec.restype = C.c_double
    L.orc_calc_sigma_sar_mat.argtypes = [C.c_int, I, C.c_int, D, D, D]
    L.orc_calc_sigma_sar_mat.restype = C.c_int
    L.orc_trunc_normal_prec_f32.argtypes = [C.c_double] * 4 + [C.c_float]
    L.orc_trunc_normal_prec_f32.restype = C.c_double
    L.orc_ndtri7_lower.argtypes = [C.c_float]
    L.orc_ndtri7_lower.restype = C.c_float
    L.orc_erfc_cheb32.argtypes = [C.c_float]
    L.orc_erfc_cheb32.restype = C.c_float
    L.orc_trunc_normal.restype = C.c_double
    L.orc_slip_smc_exec.argtypes = [C.POINTER(Problem), D, D, C.c_uint64, D, D, D, D, C.POINTER(SlipResult)]
    L.orc_fault_calc_likelihood.argtypes = [C.POINTER(Problem), D, C.c_uint64, C.POINTER(SlipResult)]
    L.orc_fault_calc_likelihood.restype = C.c_double
    L.orc_fault_calc_likelihood_batch.argtypes = [C.POINTER(Problem), D, C.c_int, C.c_int,
                                                  C.POINTER(C.c_uint64), D]
    L.orc_fault_normalize_weights.argtypes = [D, C.c_int]
    L.orc_fault_calc_mean_particles.argtypes = [D, D, D, C.c_int, C.c_int]
    L.orc_fault_calc_cov_particles.argtypes = [D, D, D, D, C.c_int, C.c_int, I]
    L.orc_qsort.argtypes = [I, I, C.c_int]
    L.orc_fault_propose.argtypes = [C.c_int, D, D, D, D, D]
    L.orc_fault_reorder.argtypes = [I, D, D, C.c_int, C.c_int, I]
    L.orc_fault_smc_exec.argtypes = [C.POINTER(Problem), D, C.c_int, C.c_uint64, D, D, C.c_int,
                                     C.POINTER(FaultResult)]
    L.orc_fault_smc_exec_gsmc.argtypes = [C.POINTER(Problem), D, C.c_int, C.c_uint64, D, D, C.c_int,
                                          C.POINTER(FaultResult), C.POINTER(Gsmc)]
    L.orc_inst_grad.argtypes = [C.POINTER(Problem), D, D, C.c_double, D, D]
    L.orc_hmc_tuning_step.argtypes = [C.POINTER(Problem), D, D, C.c_uint64, C.c_int, C.c_double, D, D, D, D, I, D,
                                      D, D, D]
    L.orc_hmc_tuning_step.restype = C.c_int
    L.orc_draw_normal.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]
    L.orc_draw_normal.restype = C.c_double
    L.orc_draw_clamped_uniform.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]
    L.orc_draw_clamped_uniform.restype = C.c_double
    L.orc_hmc_leapfrog.argtypes = [C.POINTER(Problem), D, D, C.c_double, D, D, D, C.c_double, C.c_int, D, D, D, D]
    L.orc_fault_calc_neglog_q.argtypes = [D, D, D, C.c_double, C.c_int]
    L.orc_fault_calc_neglog_q.restype = C.c_double


def dp(a):
    return a.ctypes.data_as(D)


def ip(a):
    return a.ctypes.data_as(I)


def f64(a, shape=None):
    a = np.asarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape, order="F")
    return np.asfortranarray(a)


def i32(a):
    return np.asfortranarray(np.asarray(a, dtype=np.int32))


# ---------------------------------------------------------------- RNG
def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32(c, k, o)
    return list(o)


def draw_normal(key, site, stage, c2, c1):
    """Box-Muller normal of the keyed stream (sin branch, clamped uniforms)."""
    return lib().orc_draw_normal(key, site, stage, c2, c1)


def draw_clamped_uniform(key, site, stage, c2, c1, word=0):
    return lib().orc_draw_clamped_uniform(key, site, stage, c2, c1, word)


def inner_key(seed, stage, role, slot):
    return lib().orc_inner_key(seed, stage, role, slot)


# ---------------------------------------------------------------- Okada
def dc3d0(alpha, x, y, z, depth, dip, pot1, pot2, pot3, pot4):
    u = np.zeros(12)
    iret = C.c_int(-99)
    lib().orc_dc3d0(alpha, x, y, z, depth, dip, pot1, pot2, pot3, pot4, dp(u), C.byref(iret))
    return u, iret.value


class Fault:
    """Holds the fault description and keeps its arrays alive."""

    def __init__(self, nxi_ls, neta_ls, xmin, xmax, zmin, fix_xbend=False, xbend=()):
        self.nxi = i32(nxi_ls)
        self.neta = i32(neta_ls)
        self.xbend = f64(list(xbend) + [0.0])
        self.desc = FaultDesc(len(self.nxi), ip(self.nxi), ip(self.neta), xmin, xmax, zmin,
                              int(bool(fix_xbend)), dp(self.xbend))
        nn, nd, npt = C.c_int(), C.c_int(), C.c_int()
        self.ndof_index = np.zeros(len(self.nxi) + 1, np.int32)
        lib().orc_mesh_sizes(C.byref(self.desc), C.byref(nn), C.byref(nd), C.byref(npt),
                             ip(self.ndof_index))
        self.nnode, self.ndof, self.npatch = nn.value, nd.value, npt.value
        self.ndim = 2 * self.ndof
        self.ndim_fault = 2 * len(self.nxi) + 4 if fix_xbend else 3 * len(self.nxi) + 3

    def mesh(self, theta):
        theta = f64(theta)
        cny = np.zeros((4, self.npatch), np.int32, order="F")
        coor = np.zeros((2, self.nnode), order="F")
        n2e = np.zeros((4, self.nnode), np.int32, order="F")
        n2es = np.zeros(self.nnode, np.int32)
        id_dof = np.zeros(self.ndof, np.int32)
        lib().orc_discretize_fault(C.byref(self.desc), dp(theta), ip(cny), dp(coor), ip(n2e), ip(n2es),
                                   ip(id_dof))
        return dict(cny=cny, coor=coor, n2e=n2e, n2es=n2es, id_dof=id_dof)

    def laplacian(self, theta):
        theta = f64(theta)
        m = self.mesh(theta)
        nn, nd = self.nnode, self.ndof
        luni = np.zeros((nn, nn), order="F")
        lmat = np.zeros((2 * nn, 2 * nd), order="F")
        lib().orc_gen_laplacian(C.byref(self.desc), dp(theta), nn, ip(m["id_dof"]), nd, dp(luni), dp(lmat))
        li = np.zeros((9, 2 * nn), np.int32, order="F")
        lv = np.zeros((9, 2 * nn), order="F")
        lti = np.zeros((9, 2 * nd), np.int32, order="F")
        ltv = np.zeros((9, 2 * nd), order="F")
        lib().orc_gen_sparse_lmat(dp(lmat), nn, nd, ip(li), dp(lv), ip(lti), dp(ltv))
        ll = np.zeros((2 * nd, 2 * nd), order="F")
        lib().orc_calc_ll(dp(lmat), nn, nd, dp(ll))
        return dict(luni=luni, lmat=lmat, li=li, lv=lv, lti=lti, ltv=ltv, ll=ll, **m)

    def greens(self, theta, obs_points, obs_unitvec, nsar_total, ngnss):
        theta = f64(theta)
        op = f64(obs_points)
        ou = f64(obs_unitvec)
        nobs = op.shape[1]
        g = np.zeros((nobs, self.ndim), order="F")
        lib().orc_calc_greens_func(C.byref(self.desc), dp(theta), nobs, nsar_total, ngnss, dp(op), dp(ou),
                                   dp(g))
        return g


class ProblemHolder:
    """Problem description for the end-to-end oracle (GNSS rows and optional SAR paths)."""

    def __init__(self, fault, obs, max_slip, nparticle_slip, mover, gibbs_sweeps=5, hmc_ntau=10, hmc_accept=0.65,
                 stage_cap=0, test_like=None):
        """test_like = (m, s) replaces the fault likelihood by 0.5 sum ((theta - m) / s)^2 (a test hook for
        the outer SMC's composition; no inner SMC runs)."""
        self.fault = fault
        self.test_like = None if test_like is None else f64(np.concatenate([test_like[0], test_like[1]]))
        self.obs_points = f64(obs["obs_points"])
        self.obs_unitvec = f64(obs["obs_unitvec"])
        self.obs_sigma = f64(obs["obs_sigma"])
        self.dvec = f64(obs["dvec"])
        self.nsar_index = i32(obs.get("nsar_index", [1, 1 + obs.get("nsar_total", 0)]))
        self.sar_sinv = f64(obs.get("sar_sinv", np.zeros(1)))
        nobs = self.dvec.shape[0]
        self.pb = Problem(nobs, int(obs.get("nsar_total", 0)), int(obs["ngnss"]),
                          len(self.nsar_index) - 1, ip(self.nsar_index), dp(self.obs_points),
                          dp(self.obs_unitvec), dp(self.obs_sigma), dp(self.dvec), dp(self.sar_sinv),
                          fault.desc, float(max_slip), int(nparticle_slip), int(mover), int(gibbs_sweeps),
                          int(hmc_ntau), float(hmc_accept), int(stage_cap),
                          None if self.test_like is None else dp(self.test_like))

    def slip_smc(self, theta, gmat, key):
        f = self.fault
        P = self.pb.nparticle_slip
        X = np.zeros((f.ndim, P), order="F")
        L = np.zeros(P)
        PR = np.zeros(P)
        mean = np.zeros(f.ndim)
        gt = np.zeros(202)
        et = np.zeros(202)
        at = np.zeros(202)
        dt = np.zeros(202)
        res = SlipResult(0, 0.0, dp(gt), dp(et), dp(at), dp(dt))
        lib().orc_slip_smc_exec(C.byref(self.pb), dp(f64(theta)), dp(f64(gmat)), key, dp(X), dp(L), dp(PR),
                                dp(mean), C.byref(res))
        n = res.nstage
        return dict(particles=X, likelihood=L, prior=PR, mean=mean, nstage=n, neglog=res.neglog,
                    gamma=gt[:n].copy(), evid=et[:n].copy(), accept=at[:n].copy(), dtau=dt[:n].copy())

    def hmc_tuning_step(self, theta, gmat, key, stage, gamma, X, L, PR, cov, assigned, dtau, log_lo, log_hi):
        """One slip_hmc_tuning call; returns (rc, dtau list, log_lo, log_hi, scores)."""
        P = self.pb.nparticle_slip
        d = f64(dtau).copy()
        lo, hi = C.c_double(log_lo), C.c_double(log_hi)
        sc = np.zeros(P)
        rc = lib().orc_hmc_tuning_step(C.byref(self.pb), dp(f64(theta)), dp(f64(gmat)), key, stage, gamma,
                                       dp(f64(X)), dp(f64(L)), dp(f64(PR)), dp(f64(cov)), ip(i32(assigned)), dp(d),
                                       C.byref(lo), C.byref(hi), dp(sc))
        return rc, d, lo.value, hi.value, sc

    def grad(self, theta, gmat, gamma, s):
        """Tempered gradient of gamma * likelihood + prior on this geometry (slip_calc_grad)."""
        s = f64(s)
        g = np.zeros(s.shape[0])
        lib().orc_inst_grad(C.byref(self.pb), dp(f64(theta)), dp(f64(gmat)), gamma, dp(s), dp(g))
        return g

    def leapfrog(self, theta, gmat, gamma, cur, p0, cov, dtau, ntau):
        """One slip_leapfrog trajectory; returns (cand, p, ham_cur, ham_cand)."""
        cur = f64(cur)
        cand = np.zeros_like(cur)
        p = np.zeros_like(cur)
        hc = C.c_double()
        hn = C.c_double()
        lib().orc_hmc_leapfrog(C.byref(self.pb), dp(f64(theta)), dp(f64(gmat)), gamma, dp(cur), dp(f64(p0)),
                               dp(f64(cov)), dtau, ntau, dp(cand), dp(p), C.byref(hc), C.byref(hn))
        return cand, p, hc.value, hn.value

    def fault_likelihood(self, theta, key):
        gt = np.zeros(202)
        et = np.zeros(202)
        res = SlipResult(0, 0.0, dp(gt), dp(et), None, None)
        v = lib().orc_fault_calc_likelihood(C.byref(self.pb), dp(f64(theta)), key, C.byref(res))
        return v, res.nstage

    def fault_likelihood_batch(self, thetas, keys):
        th = f64(thetas)
        B = th.shape[1]
        k = np.asarray(keys, dtype=np.uint64)
        out = np.zeros(B)
        lib().orc_fault_calc_likelihood_batch(C.byref(self.pb), dp(th), th.shape[0], B,
                                              k.ctypes.data_as(C.POINTER(C.c_uint64)), dp(out))
        return out

    def fault_smc(self, rng, N, seed, max_stages=50, gsmc=None):
        """Outer SMC; gsmc = (qmean, qcov_chol, qcov_inv, log_qcov_det) selects the gsmc branch."""
        nd = self.fault.ndim_fault
        X = np.zeros((nd, N), order="F")
        L = np.zeros(N)
        gt = np.zeros(max_stages + 1)
        res = FaultResult(0, 0.0, dp(gt))
        if gsmc is None:
            lib().orc_fault_smc_exec(C.byref(self.pb), dp(f64(rng)), N, seed, dp(X), dp(L), max_stages,
                                     C.byref(res))
        else:
            keep = [f64(gsmc[0]), f64(gsmc[1]), f64(gsmc[2])]
            g = Gsmc(dp(keep[0]), dp(keep[1]), dp(keep[2]), float(gsmc[3]))
            lib().orc_fault_smc_exec_gsmc(C.byref(self.pb), dp(f64(rng)), N, seed, dp(X), dp(L), max_stages,
                                          C.byref(res), C.byref(g))
        return dict(particles=X, likelihood=L, nstage=res.nstage, evidence=res.evidence,
                    gamma=gt[:res.nstage].copy())


# ---------------------------------------------------------------- inner-SMC components
def likelihood(s, H, b, dsd, nsar, ngnss, lss2, lsg2):
    s = f64(s)
    H = f64(H)
    return lib().orc_slip_calc_likelihood(dp(s), s.shape[0], dp(H), dp(f64(b)), dsd, nsar, ngnss, lss2, lsg2)


def prior(s, alpha2, li, lv, nnode):
    s = f64(s)
    lsv = np.zeros(2 * nnode)
    return lib().orc_slip_calc_prior(dp(s), dp(f64(alpha2)), ip(i32(li)), dp(f64(lv)), nnode, dp(lsv))


def grad(s, nnode, ndof, li, lv, lti, ltv, gamma, alpha2, b, H):
    s = f64(s)
    g = np.zeros(s.shape[0])
    lsv = np.zeros(2 * nnode)
    lib().orc_slip_calc_grad(dp(g), dp(s), s.shape[0], nnode, ndof, ip(i32(li)), dp(f64(lv)), ip(i32(lti)),
                             dp(f64(ltv)), gamma, dp(f64(alpha2)), dp(f64(b)), dp(f64(H)), dp(lsv))
    return g


def calc_sigma_sar_mat(nsar_index, obs_points, obs_sigma):
    """Inverse InSAR covariance per path (init.f90:82-119), column-major blocks concatenated."""
    nsi = i32(nsar_index)
    sizes = np.diff(nsi)
    out = np.zeros(max(1, int((sizes.astype(np.int64) ** 2).sum())))
    pts = f64(obs_points)
    rc = lib().orc_calc_sigma_sar_mat(len(nsi) - 1, ip(nsi), pts.shape[1], dp(pts), dp(f64(obs_sigma)), dp(out))
    if rc != 0:
        raise ValueError("calc_sigma_sar_mat: sigma index past nobs or singular covariance")
    return out


def gsd_gsg(gmat, dvec, obs_sigma, sigma_sar2, sigma_gnss2, nsar_index=(1, 1), sar_sinv=None):
    g = f64(gmat)
    nobs, nd = g.shape
    nsi = i32(nsar_index)
    nsar_total = int(nsi[-1] - nsi[0])
    sinv = f64(sar_sinv if sar_sinv is not None else np.zeros(1))
    b = np.zeros(nd)
    H = np.zeros((nd, nd), order="F")
    dsd = C.c_double()
    lib().orc_slip_calc_gsd_gsg(dp(g), dp(f64(dvec)), nobs, nd, len(nsi) - 1, ip(nsi), nsar_total, dp(sinv),
                                sigma_sar2, sigma_gnss2, dp(f64(obs_sigma)), dp(b), dp(H), C.byref(dsd))
    return H, b, dsd.value


def gibbs_init_chain(P, ll, alpha2, max_slip, bm):
    ll = f64(ll)
    nd = ll.shape[0]
    X = np.zeros((nd, P), order="F")
    lib().orc_slip_gibbs_init_chain(dp(X), P, nd, dp(ll), dp(f64(alpha2)), max_slip, dp(f64(bm)))
    return X


def find_next_gamma(gprev, L, cv=2.0, tol=1e-10, cap=100000):
    L = f64(L)
    w = np.zeros(L.shape[0])
    ev = C.c_double()
    g = lib().orc_find_next_gamma(gprev, dp(L), L.shape[0], dp(w), C.byref(ev), cv, tol, cap)
    return g, w, ev.value


def normalize_quad(w):
    w = f64(w).copy()
    lib().orc_slip_normalize_weights(dp(w), w.shape[0])
    return w


def mean_cov_quad(X, w):
    X = f64(X)
    nd, P = X.shape
    m = np.zeros(nd)
    c = np.zeros(nd)
    lib().orc_slip_calc_mean_particles(dp(X), dp(f64(w)), dp(m), P, nd)
    lib().orc_slip_calc_cov_particles(dp(X), dp(f64(w)), dp(m), dp(c), P, nd)
    return m, c


def resample_quad(w, u_hi, u_lo):
    w = f64(w)
    a = np.zeros(w.shape[0], np.int32)
    lib().orc_resample_quad(w.shape[0], dp(w), u_hi, u_lo, ip(a))
    return a


def ndtri(p):
    return lib().orc_ndtri(p)


def trunc_normal(mu, sigma, lo, hi, u):
    return lib().orc_trunc_normal(mu, sigma, lo, hi, u)


def trunc_normal_prec(mu, s, lo, hi, u):
    """The Gibbs mover's draw: s = sqrt(precision)."""
    return lib().orc_trunc_normal_prec(mu, s, lo, hi, u)


def trunc_normal_prec_f32(mu, s, lo, hi, u):
    """Single-precision draw of the mixed-precision Gibbs mover (NaN where it declines)."""
    return lib().orc_trunc_normal_prec_f32(mu, s, lo, hi, u)


def erfc_cheb32(x):
    return lib().orc_erfc_cheb32(x)


def ndtri7_lower(p):
    """AS 241 PPND7 in single precision for p in (0, 1/2]."""
    return lib().orc_ndtri7_lower(p)


def qsort_desc(arr):
    arr = i32(arr).copy()
    idx = np.arange(1, arr.shape[0] + 1, dtype=np.int32)
    lib().orc_qsort(ip(arr), ip(idx), arr.shape[0])
    return idx


def fault_normalize(w):
    w = f64(w).copy()
    lib().orc_fault_normalize_weights(dp(w), w.shape[0])
    return w


def fault_mean_cov(X, w):
    X = f64(X)
    nd, N = X.shape
    m = np.zeros(nd)
    cov = np.zeros((nd, nd), order="F")
    info = C.c_int()
    lib().orc_fault_calc_mean_particles(dp(X), dp(f64(w)), dp(m), N, nd)
    lib().orc_fault_calc_cov_particles(dp(X), dp(f64(w)), dp(m), dp(cov), N, nd, C.byref(info))
    return m, cov, info.value


def fault_propose(chol, cur, z, rng):
    """work_mcmc_sampling's proposal: cur + chol z, reflected into the range (smc_fault.f90:711-747)."""
    cur = f64(cur)
    zz = f64(z).copy()
    cand = np.zeros_like(cur)
    lib().orc_fault_propose(cur.shape[0], dp(f64(chol)), dp(cur), dp(zz), dp(f64(rng)), dp(cand))
    return cand


def fault_neglog_q(x, qmean, qcov_inv, log_det):
    """fault_calc_neglog_q (smc_fault.f90:9-27)."""
    x = f64(x)
    return lib().orc_fault_calc_neglog_q(dp(x), dp(f64(qmean)), dp(f64(qcov_inv)), float(log_det), x.shape[0])
