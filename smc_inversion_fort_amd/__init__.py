"""MI355X implementation of the nested SMC fault inversion hot path (libsmcf).

Python host mirror of the reference's Fortran procedures over the C ABI of
libsmcf.so (include/smcf.h).  Names follow the reference:

    Problem.calc_greens_func          gfunc.f90:242-348
    Problem.laplacian                 init.f90:295-541 (gen_laplacian, gen_sparse_lmat, calc_ll)
    Problem.slip_calc_gsd_gsg         smc_slip.f90:1347-1443
    Problem.slip_smc_exec             smc_slip.f90:1445-1663
    Problem.fault_calc_likelihood     smc_fault.f90:66-292 (batched over thetas)
    Context.dc3d0                     DC3Dfortran.f90:2-118
    Context.find_next_gamma           smc_slip.f90:314-387 / smc_fault.f90:392-485
    Context.normalize_resample        smc_slip.f90:465-483, 572-589
    qsort                             sort.cpp:3-7

Arrays use the reference layout (Fortran order, 1-based index values).  There is no
CPU fallback: a missing library or device raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsmcf.so")
_LIB = None

MOVER_HMC = 0         # reference HMC: ntau 50, ESJD step-size tuning
MOVER_GIBBS = 1       # truncated-MVN Gibbs sweeps
MOVER_HMC_ACCEPT = 2  # HMC with hmc_ntau steps, step size adapted toward the hmc_accept mean acceptance
PRECISION_FP64 = 0   # every operation in FP64
PRECISION_MIXED = 1  # Gibbs Q.S products in FP32 on the matrix cores, single-precision draws; means/weights FP64
PRECISION_MIXED_BF16 = 2  # as PRECISION_MIXED with the Q.S products from split-BF16 operands (hi.hi + hi.lo + lo.hi)

REQUEST_OVERHEAD_BYTES = 16 << 20  # serving admission: fixed device bytes counted per request beside its problem

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)
_U64 = C.POINTER(C.c_uint64)
_U32 = C.POINTER(C.c_uint32)


class SmcfError(RuntimeError):
    pass


class ProblemDesc(C.Structure):
    _fields_ = [("nplane", C.c_int), ("nxi_ls", _I), ("neta_ls", _I), ("xmin", C.c_double),
                ("xmax", C.c_double), ("zmin", C.c_double), ("fix_xbend", C.c_int), ("xbend", _D),
                ("nobs", C.c_int), ("nsar_total", C.c_int), ("ngnss", C.c_int), ("npath", C.c_int),
                ("nsar_index", _I), ("obs_points", _D), ("obs_unitvec", _D), ("obs_sigma", _D),
                ("dvec", _D), ("sar_sinv", _D), ("max_slip", C.c_double), ("nparticle_slip", C.c_int),
                ("mover", C.c_int), ("gibbs_sweeps", C.c_int), ("precision", C.c_int), ("hmc_ntau", C.c_int),
                ("hmc_accept", C.c_double)]


EXPORTS = [
    "smcf_last_error", "smcf_ctx_create", "smcf_ctx_destroy", "smcf_set_profiling", "smcf_kernel_stats",
    "smcf_problem_create", "smcf_problem_destroy", "smcf_problem_dims", "smcf_dc3d0",
    "smcf_calc_greens_func", "smcf_laplacian", "smcf_slip_calc_gsd_gsg", "smcf_slip_smc_exec",
    "smcf_fault_calc_likelihood_batch", "smcf_find_next_gamma", "smcf_normalize_resample",
    "smcf_trunc_normal", "smcf_trunc_normal_f32", "smcf_philox", "smcf_fault_smc_set_output_dir", "smcf_fault_smc_exec",
    "smcf_fault_smc_exec_gsmc", "smcf_inner_key", "qsort_", "smcf_problem_bytes", "smcf_problem_device_bytes",
    "smcf_server_create", "smcf_server_submit", "smcf_server_wait", "smcf_server_status", "smcf_server_destroy",
    "smcf_slip_mean_batch", "smcf_batch_times", "smcf_batch_admit_times", "smcf_find_next_gamma_noalloc",
    "smcf_slip_calc_likelihood", "smcf_slip_calc_prior", "smcf_slip_calc_grad", "smcf_slip_leapfrog",
    "smcf_slip_gen_init_particles", "smcf_slip_calc_moments", "smcf_slip_hmc_accept_state",
    "smcf_fault_sample_init_particles", "smcf_fault_mean_std", "smcf_fault_find_next_gamma",
    "smcf_fault_normalize_weights", "smcf_fault_calc_mean_particles", "smcf_fault_calc_cov_particles",
    "smcf_fault_resample_particles", "smcf_fault_reorder_to_send", "smcf_work_mcmc_sampling", "smcf_call_dc3d0",
    "smcf_calc_responce", "smcf_slip_likelihood_gram", "smcf_slip_prior_ell", "smcf_slip_grad_ell",
    "smcf_slip_calc_gsd_gsg_gmat", "smcf_discretize_fault", "smcf_server_set_max_running",
    "smcf_outer_reweight_resample", "smcf_fault_smc_set_outer_stage", "smcf_problem_context",
    "smcf_gibbs_block_rows", "smcf_set_gibbs_block_rows", "smcf_group_create", "smcf_group_set_graph",
    "smcf_group_calc_likelihood_batch", "smcf_group_fault_smc_exec", "smcf_group_destroy",
]


def lib():
    """Load libsmcf.so (built in-tree by __graft_entry__.build() / make)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise SmcfError("libsmcf.so is not built: run `make -C smc_inversion_fort_amd`")
        L = C.CDLL(LIB_PATH)
        _declare(L)
        _LIB = L
    return _LIB


def _declare(L):
    L.smcf_last_error.restype = C.c_char_p
    L.smcf_ctx_create.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
    L.smcf_ctx_destroy.argtypes = [C.c_void_p]
    L.smcf_set_profiling.argtypes = [C.c_void_p, C.c_int]
    L.smcf_kernel_stats.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_long), C.POINTER(C.c_double)]
    L.smcf_problem_create.argtypes = [C.c_void_p, C.POINTER(ProblemDesc), C.c_int, C.POINTER(C.c_void_p)]
    L.smcf_problem_destroy.argtypes = [C.c_void_p]
    L.smcf_problem_dims.argtypes = [C.c_void_p, _I, _I, _I, _I, _I]
    L.smcf_dc3d0.argtypes = [C.c_void_p, C.c_int, _D, _D, _I]
    L.smcf_calc_greens_func.argtypes = [C.c_void_p, C.c_int, _D, _D]
    L.smcf_laplacian.argtypes = [C.c_void_p, _D, _I, _D, _I, _D, _D]
    L.smcf_slip_calc_gsd_gsg.argtypes = [C.c_void_p, C.c_int, _D, _D, _D, _D]
    L.smcf_slip_smc_exec.argtypes = [C.c_void_p, _D, C.c_uint64, _D, _D, _D, _D, _I, _D, _D]
    L.smcf_fault_calc_likelihood_batch.argtypes = [C.c_void_p, C.c_int, _D, _U64, _D, _I]
    L.smcf_find_next_gamma.argtypes = [C.c_void_p, C.c_int, C.c_double, _D, C.c_double, C.c_double, C.c_int,
                                       _D, _D, _D]
    L.smcf_find_next_gamma_noalloc.argtypes = [C.c_void_p, C.c_int, C.c_double, _D, C.c_double, C.c_double,
                                               C.c_int, _D, _D, _D]
    L.smcf_normalize_resample.argtypes = [C.c_void_p, C.c_int, _D, C.c_double, C.c_double, _D, _I]
    L.smcf_trunc_normal.argtypes = [C.c_void_p, C.c_int, _D, _D, C.c_double, C.c_double, _D, _D]
    L.smcf_trunc_normal_f32.argtypes = [C.c_void_p, C.c_int, _D, _D, C.c_double, C.c_double, C.POINTER(C.c_float), _D]
    L.smcf_fault_smc_set_output_dir.argtypes = [C.c_void_p, C.c_char_p]
    L.smcf_philox.argtypes = [C.c_void_p, C.c_int, _U32, _U32, _U32]
    L.smcf_fault_smc_exec.argtypes = [C.c_void_p, _D, C.c_int, C.c_uint64, C.c_int, C.c_int, _D, _D, _I, _D, _D]
    L.smcf_fault_smc_exec_gsmc.argtypes = [C.c_void_p, _D, C.c_int, C.c_uint64, C.c_int, C.c_int, _D, _D, _D,
                                           C.c_double, _D, _D, _I, _D, _D]
    L.smcf_inner_key.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32]
    L.smcf_inner_key.restype = C.c_uint64
    L.qsort_.argtypes = [_I, _I, _I]
    L.smcf_problem_bytes.argtypes = [C.POINTER(ProblemDesc), C.c_int, C.POINTER(C.c_size_t)]
    L.smcf_problem_device_bytes.argtypes = [C.c_void_p, C.POINTER(C.c_size_t)]
    L.smcf_server_create.argtypes = [C.c_int, C.c_size_t, C.POINTER(C.c_void_p)]
    L.smcf_server_submit.argtypes = [C.c_void_p, C.POINTER(ProblemDesc), C.c_int, _D, C.c_int, C.c_uint64, C.c_int,
                                     C.POINTER(C.c_int64)]
    L.smcf_server_wait.argtypes = [C.c_void_p, C.c_int64, _D, _D, _I, C.POINTER(C.c_double), _D, _D]
    L.smcf_server_status.argtypes = [C.c_void_p, _I, _I, C.POINTER(C.c_size_t)]
    L.smcf_server_destroy.argtypes = [C.c_void_p]
    L.smcf_server_set_max_running.argtypes = [C.c_void_p, C.c_int]
    L.smcf_slip_mean_batch.argtypes = [C.c_void_p, C.c_int, _D, _U64, _D, _D]
    L.smcf_batch_times.argtypes = [C.c_void_p, C.c_int, _D]
    L.smcf_batch_admit_times.argtypes = [C.c_void_p, C.c_int, _D]
    L.smcf_slip_calc_likelihood.argtypes = [C.c_void_p, _D, C.c_int, _D, _D]
    L.smcf_slip_calc_prior.argtypes = [C.c_void_p, _D, C.c_int, _D, _D]
    L.smcf_slip_calc_grad.argtypes = [C.c_void_p, _D, C.c_double, C.c_int, _D, _D]
    L.smcf_slip_leapfrog.argtypes = [C.c_void_p, _D, C.c_double, C.c_int, _D, _D, _D, C.c_double, C.c_int, _D, _D,
                                     _D, _D, _D, _D]
    L.smcf_slip_calc_gsd_gsg_gmat.argtypes = [C.c_void_p, _D, C.c_double, C.c_double, _D, _D, _D]
    L.smcf_discretize_fault.argtypes = [C.c_void_p, _D, _I, _D, _I, _I, _I, _D, _D]
    L.smcf_slip_gen_init_particles.argtypes = [C.c_void_p, _D, C.c_uint64, _D, _D, _D]
    L.smcf_slip_calc_moments.argtypes = [C.c_void_p, C.c_int, C.c_int, _D, _D, _D, _D]
    L.smcf_slip_hmc_accept_state.argtypes = [C.c_void_p, _D, _D]
    L.smcf_call_dc3d0.argtypes = [C.c_void_p, C.c_int, _D, _D]
    L.smcf_slip_likelihood_gram.argtypes = [C.c_void_p, C.c_int, C.c_int, _D, _D, _D, C.c_double, C.c_int, C.c_int,
                                            C.c_double, C.c_double, _D]
    L.smcf_slip_prior_ell.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, _D, _D, _I, _D, _D]
    L.smcf_slip_grad_ell.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, _D, C.c_double, _D, _D, _D, _I, _D, _I,
                                     _D, _D]
    L.smcf_calc_responce.argtypes = [C.c_void_p, C.c_int, _D, C.c_int, _I, C.c_int, _D, _D, C.c_double, C.c_double,
                                     C.c_double, C.c_double, C.c_double, C.c_int, _I, _D]
    L.smcf_fault_sample_init_particles.argtypes = [C.c_int, C.c_int, C.c_uint64, _D, _D, _D, _D]
    L.smcf_fault_mean_std.argtypes = [C.c_int, _D, _D, _D]
    L.smcf_fault_find_next_gamma.argtypes = [C.c_int, C.c_double, _D, _D, _D, _D, _D]
    L.smcf_fault_normalize_weights.argtypes = [C.c_int, _D]
    L.smcf_fault_calc_mean_particles.argtypes = [C.c_int, C.c_int, _D, _D, _D]
    L.smcf_fault_calc_cov_particles.argtypes = [C.c_int, C.c_int, _D, _D, _D, _D, _I]
    L.smcf_fault_resample_particles.argtypes = [C.c_int, _D, C.c_uint64, C.c_int, _I]
    L.smcf_fault_reorder_to_send.argtypes = [C.c_int, C.c_int, _I, _D, _D, _D, _I]
    L.smcf_outer_reweight_resample.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_uint64, C.c_int,
                                               C.c_int, _D, _D, _D, _D, _D, _D, _D, _I, _I, _I]
    L.smcf_fault_smc_set_outer_stage.argtypes = [C.c_void_p, C.c_int]
    L.smcf_gibbs_block_rows.argtypes = [C.c_int, C.c_int]
    L.smcf_set_gibbs_block_rows.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
    L.smcf_problem_context.argtypes = [C.c_void_p]
    L.smcf_problem_context.restype = C.c_void_p
    _PP = C.POINTER(C.c_void_p)
    L.smcf_group_create.argtypes = [C.c_void_p, C.c_int, _PP, _PP]
    L.smcf_group_set_graph.argtypes = [C.c_void_p, C.c_int]
    L.smcf_group_calc_likelihood_batch.argtypes = [C.c_void_p, _I, _PP, _PP, _PP, _PP]
    L.smcf_group_fault_smc_exec.argtypes = [C.c_void_p, _PP, _I, _U64, C.c_int, _PP, _PP, _I, _D]
    L.smcf_group_destroy.argtypes = [C.c_void_p]
    L.smcf_work_mcmc_sampling.argtypes = [C.c_void_p, C.c_uint64, C.c_int, C.c_double, _D, _D, C.c_int, _I, _D, _D,
                                          _D, _D, C.c_double, C.c_int, _D, _D, _D, _I]


def _check(rc):
    if rc != 0:
        raise SmcfError("libsmcf error %d: %s" % (rc, lib().smcf_last_error().decode()))


def _f64(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def _i32(a):
    return np.asfortranarray(np.asarray(a, dtype=np.int32))


def _dp(a):
    return a.ctypes.data_as(_D)


def _ip(a):
    return a.ctypes.data_as(_I)


def inner_key(seed, outer_stage, role, slot):
    """RNG key of the inner SMC of one outer evaluation."""
    return lib().smcf_inner_key(seed, outer_stage, role, slot)


def qsort(arr, index):
    """sort.cpp:3-7: sort the 1-based `index` in place by descending arr(index)."""
    arr = _i32(arr)
    index = np.ascontiguousarray(index, dtype=np.int32)
    n = C.c_int(index.shape[0])
    lib().qsort_(_ip(arr), _ip(index), C.byref(n))
    return index


# ------------------------------------------------------------------ outer-SMC stage components (host code)
def _opt(a):
    return None if a is None else _dp(a)


def fault_sample_init_particles(ndim, N, seed, range_=None, qmean=None, qcov_chol=None):
    """fault_sample_init_particles (smc_fault.f90:29-64): ndim x N, uniform in range_ or qmean + chol z."""
    X = np.zeros((ndim, N), order="F")
    r = None if range_ is None else _f64(range_)
    qm = None if qmean is None else _f64(qmean)
    qc = None if qcov_chol is None else _f64(qcov_chol)
    _check(lib().smcf_fault_sample_init_particles(ndim, N, C.c_uint64(seed), _opt(r), _opt(qm), _opt(qc), _dp(X)))
    return X


def fault_mean_std(vec):
    """calc_mean_std_vector (smc_fault.f90:372-390)."""
    v = _f64(vec)
    m, sd = C.c_double(), C.c_double()
    _check(lib().smcf_fault_mean_std(v.shape[0], _dp(v), C.byref(m), C.byref(sd)))
    return m.value, sd.value


def fault_find_next_gamma(gamma_prev, likelihood, q=None):
    """fault_find_next_gamma (smc_fault.f90:392-485): (gamma, weights, -log evidence increment)."""
    L = _f64(likelihood)
    qq = None if q is None else _f64(q)
    w = np.zeros(L.shape[0])
    g, ev = C.c_double(), C.c_double()
    _check(lib().smcf_fault_find_next_gamma(L.shape[0], float(gamma_prev), _dp(L), _opt(qq), _dp(w), C.byref(g),
                                            C.byref(ev)))
    return g.value, w, ev.value


def fault_normalize_weights(w):
    w = _f64(w).copy()
    _check(lib().smcf_fault_normalize_weights(w.shape[0], _dp(w)))
    return w


def fault_calc_mean_particles(X, w):
    X = _f64(X)
    m = np.zeros(X.shape[0])
    _check(lib().smcf_fault_calc_mean_particles(X.shape[0], X.shape[1], _dp(X), _dp(_f64(w)), _dp(m)))
    return m


def fault_calc_cov_particles(X, w, mean):
    """Lower Cholesky factor of 0.04 x the weighted covariance and dpotrf's info."""
    X = _f64(X)
    nd = X.shape[0]
    cov = np.zeros((nd, nd), order="F")
    info = C.c_int()
    _check(lib().smcf_fault_calc_cov_particles(nd, X.shape[1], _dp(X), _dp(_f64(w)), _dp(_f64(mean)), _dp(cov),
                                               C.byref(info)))
    return cov, info.value


def fault_resample_particles(w, seed, stage):
    w = _f64(w)
    a = np.zeros(w.shape[0], np.int32)
    _check(lib().smcf_fault_resample_particles(w.shape[0], _dp(w), C.c_uint64(seed), int(stage), _ip(a)))
    return a


def fault_reorder_to_send(assigned, X, L, q=None):
    """Descending reorder through qsort_; returns (assigned, X, L, q, sorted_idx) (copies)."""
    a = np.array(assigned, dtype=np.int32)
    Xc = np.array(_f64(X), order="F")
    Lc = np.array(_f64(L))
    qc = None if q is None else np.array(_f64(q))
    sidx = np.zeros(a.shape[0], np.int32)
    _check(lib().smcf_fault_reorder_to_send(Xc.shape[0], a.shape[0], _ip(a), _dp(Xc), _dp(Lc), _opt(qc), _ip(sidx)))
    return a, Xc, Lc, qc, sidx


def gibbs_block_rows(precision, ndim):
    """Rows per coordinate block the Gibbs sweep kernels use for a problem of this form and ndim (tuning table)."""
    r = lib().smcf_gibbs_block_rows(int(precision), int(ndim))
    if r < 0:
        raise SmcfError("libsmcf error %d: %s" % (r, lib().smcf_last_error().decode()))
    return r


def set_gibbs_block_rows(precision, ndim_lo, ndim_hi, rows):
    """Override the tuning table for ndim_lo..ndim_hi (problems created afterwards)."""
    _check(lib().smcf_set_gibbs_block_rows(int(precision), int(ndim_lo), int(ndim_hi), int(rows)))


def outer_reweight_resample(ctx, X, L, gamma_prev, seed, stage, q=None, device_sort=False):
    """The outer reweight and resample of one stage on the device (smcf_outer_reweight_resample): returns a dict with
    gamma, neglog_evidence, weights (before the reorder), cov_chol, info, and the reordered assigned, X, L, q and
    sorted_idx (1-based)."""
    Xc = np.array(_f64(X), order="F")
    nd, N = Xc.shape
    Lc = np.array(_f64(L))
    qc = None if q is None else np.array(_f64(q))
    g, ev = C.c_double(), C.c_double()
    w = np.zeros(N)
    cov = np.zeros((nd, nd), order="F")
    info = np.zeros(1, np.int32)
    a = np.zeros(N, np.int32)
    sidx = np.zeros(N, np.int32)
    _check(lib().smcf_outer_reweight_resample(ctx.h, nd, N, float(gamma_prev), C.c_uint64(seed), int(stage),
                                              int(bool(device_sort)), _dp(Xc), _dp(Lc), _opt(qc), C.byref(g),
                                              C.byref(ev), _dp(w), _dp(cov), _ip(info), _ip(a), _ip(sidx)))
    return dict(gamma=g.value, neglog_evidence=ev.value, weights=w, cov_chol=cov, info=int(info[0]), assigned=a,
                particles=Xc, likelihood=Lc, q=qc, sorted_idx=sidx)


class Context:
    """One device and one HIP stream."""

    def __init__(self, device=0):
        h = C.c_void_p()
        _check(lib().smcf_ctx_create(device, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().smcf_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_profiling(self, on=True):
        _check(lib().smcf_set_profiling(self.h, int(on)))

    def kernel_stats(self, name):
        n = C.c_long()
        t = C.c_double()
        _check(lib().smcf_kernel_stats(self.h, name.encode(), C.byref(n), C.byref(t)))
        return n.value, t.value

    def dc3d0(self, inputs):
        """inputs: (n, 10) rows of (alpha, x, y, z, depth, dip, pot1..pot4) -> (u (n,12), iret (n,))"""
        a = np.ascontiguousarray(inputs, dtype=np.float64)
        n = a.shape[0]
        u = np.zeros((n, 12))
        iret = np.full(n, -99, dtype=np.int32)
        _check(lib().smcf_dc3d0(self.h, n, _dp(a), _dp(u), _ip(iret)))
        return u, iret

    def find_next_gamma(self, gamma_prev, likelihood, cv=2.0, tol=1e-10, cap=100000):
        L = np.ascontiguousarray(likelihood, dtype=np.float64)
        w = np.zeros_like(L)
        g = C.c_double()
        ev = C.c_double()
        _check(lib().smcf_find_next_gamma(self.h, L.shape[0], gamma_prev, _dp(L), cv, tol, cap, C.byref(g),
                                          _dp(w), C.byref(ev)))
        return g.value, w, ev.value

    def find_next_gamma_noalloc(self, gamma_prev, likelihood, cv=2.0, tol=1e-10, cap=100000):
        """find_next_gamma without device allocations (the library's per-device scratch; P <= 3584)."""
        L = np.ascontiguousarray(likelihood, dtype=np.float64)
        w = np.zeros_like(L)
        g = C.c_double()
        ev = C.c_double()
        _check(lib().smcf_find_next_gamma_noalloc(self.h, L.shape[0], gamma_prev, _dp(L), cv, tol, cap,
                                                  C.byref(g), _dp(w), C.byref(ev)))
        return g.value, w, ev.value

    def slip_calc_moments(self, particles, weights):
        """Weighted mean and variance diagonal (floored at 1e-3) of particles (ndim x P)."""
        X = _f64(particles)
        nd, P = X.shape
        m = np.zeros(nd)
        c = np.zeros(nd)
        _check(lib().smcf_slip_calc_moments(self.h, nd, P, _dp(X), _dp(_f64(weights)), _dp(m), _dp(c)))
        return m, c

    def normalize_resample(self, weights, u_hi, u_lo):
        w = np.ascontiguousarray(weights, dtype=np.float64)
        wn = np.zeros_like(w)
        a = np.zeros(w.shape[0], dtype=np.int32)
        _check(lib().smcf_normalize_resample(self.h, w.shape[0], _dp(w), u_hi, u_lo, _dp(wn), _ip(a)))
        return wn, a

    def trunc_normal(self, mu, sigma, lo, hi, u):
        mu = np.ascontiguousarray(mu, dtype=np.float64)
        sg = np.ascontiguousarray(sigma, dtype=np.float64)
        u = np.ascontiguousarray(u, dtype=np.float64)
        out = np.zeros_like(mu)
        _check(lib().smcf_trunc_normal(self.h, mu.shape[0], _dp(mu), _dp(sg), lo, hi, _dp(u), _dp(out)))
        return out

    def trunc_normal_f32(self, mu, sq, lo, hi, u):
        """Single-precision draw of the mixed-precision Gibbs mover; NaN where it declines."""
        mu = np.ascontiguousarray(mu, dtype=np.float64)
        sq = np.ascontiguousarray(sq, dtype=np.float64)
        u = np.ascontiguousarray(u, dtype=np.float32)
        out = np.zeros_like(mu)
        _check(lib().smcf_trunc_normal_f32(self.h, mu.shape[0], _dp(mu), _dp(sq), lo, hi,
                                           u.ctypes.data_as(C.POINTER(C.c_float)), _dp(out)))
        return out

    def call_dc3d0(self, inputs):
        """call_dc3d0 (gfunc.f90:65-100) for rows (xsource, ysource, zsource, xobs, yobs, uxi, ueta, dip, area,
        strike) of inputs (n x 10); returns uret (n x 3)."""
        a = np.ascontiguousarray(inputs, dtype=np.float64)
        u = np.zeros((a.shape[0], 3))
        _check(lib().smcf_call_dc3d0(self.h, a.shape[0], _dp(a), _dp(u)))
        return u

    def calc_responce(self, pts, cny, coor, slip_dist, xf, yf, zf, strike, dip, target_id):
        """calc_responce (gfunc.f90:102-194) at each column of pts (2 x npts); returns uobs (3 x npts)."""
        p = _f64(pts)
        cn, co, sl, t = _i32(cny), _f64(coor), _f64(slip_dist), _i32(target_id)
        u = np.zeros((3, p.shape[1]), order="F")
        _check(lib().smcf_calc_responce(self.h, p.shape[1], _dp(p), cn.shape[1], _ip(cn), co.shape[1], _dp(co),
                                        _dp(sl), xf, yf, zf, strike, dip, t.shape[0], _ip(t), _dp(u)))
        return u

    def slip_likelihood_gram(self, particles, H, b, dsd, nsar_total, ngnss, lss2, lsg2):
        """slip_calc_likelihood from the Gram quantities, for each column of particles."""
        X = _f64(particles)
        if X.ndim == 1:
            X = X.reshape(-1, 1, order="F")
        out = np.zeros(X.shape[1])
        _check(lib().smcf_slip_likelihood_gram(self.h, X.shape[0], X.shape[1], _dp(X), _dp(_f64(H)), _dp(_f64(b)),
                                               float(dsd), int(nsar_total), int(ngnss), float(lss2), float(lsg2),
                                               _dp(out)))
        return out

    def slip_prior_ell(self, particles, alpha2_full, li, lv):
        X = _f64(particles)
        if X.ndim == 1:
            X = X.reshape(-1, 1, order="F")
        a2 = _f64(alpha2_full)
        out = np.zeros(X.shape[1])
        _check(lib().smcf_slip_prior_ell(self.h, a2.shape[0] // 2, X.shape[0], X.shape[1], _dp(X), _dp(a2),
                                         _ip(_i32(li)), _dp(_f64(lv)), _dp(out)))
        return out

    def slip_grad_ell(self, particles, gamma, H, b, alpha2_full, li, lv, lti, ltv):
        X = _f64(particles)
        if X.ndim == 1:
            X = X.reshape(-1, 1, order="F")
        a2 = _f64(alpha2_full)
        g = np.zeros(X.shape, order="F")
        _check(lib().smcf_slip_grad_ell(self.h, a2.shape[0] // 2, X.shape[0], X.shape[1], _dp(X), float(gamma),
                                        _dp(_f64(H)), _dp(_f64(b)), _dp(a2), _ip(_i32(li)), _dp(_f64(lv)),
                                        _ip(_i32(lti)), _dp(_f64(ltv)), _dp(g)))
        return g

    def philox(self, ctr, key):
        c = np.ascontiguousarray(ctr, dtype=np.uint32).reshape(-1, 4)
        k = np.ascontiguousarray(key, dtype=np.uint32).reshape(-1, 2)
        out = np.zeros_like(c)
        _check(lib().smcf_philox(self.h, c.shape[0], c.ctypes.data_as(_U32), k.ctypes.data_as(_U32),
                                 out.ctypes.data_as(_U32)))
        return out


def problem_desc(nxi_ls, neta_ls, xmin, xmax, zmin, obs, max_slip, nparticle_slip, mover=MOVER_GIBBS,
                 gibbs_sweeps=5, fix_xbend=False, xbend=(), precision=PRECISION_FP64, hmc_ntau=10, hmc_accept=0.65):
    """smcf_problem_desc of an observation set and fault mesh; returns (desc, arrays the desc points into)."""
    keep = dict(
        nxi=_i32(nxi_ls), neta=_i32(neta_ls), xbend=_f64(list(xbend) + [0.0]),
        nsar_index=_i32(obs.get("nsar_index", [1, 1 + int(obs.get("nsar_total", 0))])),
        obs_points=_f64(obs["obs_points"]), obs_unitvec=_f64(obs["obs_unitvec"]),
        obs_sigma=_f64(obs["obs_sigma"]), dvec=_f64(obs["dvec"]),
        sar=_f64(obs["sar_sinv"] if obs.get("sar_sinv") is not None else np.zeros(1)))
    k = keep
    nobs = k["dvec"].shape[0]
    desc = ProblemDesc(len(k["nxi"]), _ip(k["nxi"]), _ip(k["neta"]), xmin, xmax, zmin, int(bool(fix_xbend)),
                       _dp(k["xbend"]), nobs, int(obs.get("nsar_total", 0)), int(obs["ngnss"]),
                       len(k["nsar_index"]) - 1, _ip(k["nsar_index"]), _dp(k["obs_points"]),
                       _dp(k["obs_unitvec"]), _dp(k["obs_sigma"]), _dp(k["dvec"]), _dp(k["sar"]),
                       float(max_slip), int(nparticle_slip), int(mover), int(gibbs_sweeps), int(precision),
                       int(hmc_ntau), float(hmc_accept))
    return desc, keep


def problem_bytes(desc, max_batch):
    """Device bytes a problem of this description allocates (smcf_problem_bytes; no device work)."""
    n = C.c_size_t()
    _check(lib().smcf_problem_bytes(C.byref(desc), int(max_batch), C.byref(n)))
    return n.value


class Server:
    """Serving loop of libsmcf (smcf_server_*): inversion requests (Gibbs mover) admitted by a device-memory
    budget as they arrive and batched continuously: every stage kernel runs over all running requests."""

    def __init__(self, device=0, hbm_budget_bytes=200 << 30):
        h = C.c_void_p()
        _check(lib().smcf_server_create(int(device), C.c_size_t(int(hbm_budget_bytes)), C.byref(h)))
        self.h = h
        self._req = {}

    def submit(self, desc, rng, N, seed, max_stages=50, max_batch=None):
        """Queue one outer SMC (smcf_fault_smc_exec settings); returns the request id."""
        rng = _f64(rng)
        if max_batch is None:
            max_batch = min(int(N), 512)
        rid = C.c_int64()
        _check(lib().smcf_server_submit(self.h, C.byref(desc), int(max_batch), _dp(rng), int(N), C.c_uint64(seed),
                                        int(max_stages), C.byref(rid)))
        self._req[rid.value] = (rng.shape[1], int(N), int(max_stages))
        return rid.value

    def wait(self, rid):
        nf, N, max_stages = self._req.pop(rid)
        X = np.zeros((nf, N), order="F")
        L = np.zeros(N)
        ns = C.c_int()
        ev = C.c_double()
        gt = np.zeros(max_stages + 1)
        times = np.zeros(3)
        _check(lib().smcf_server_wait(self.h, C.c_int64(rid), _dp(X), _dp(L), C.byref(ns), C.byref(ev), _dp(gt),
                                      _dp(times)))
        return dict(particles=X, likelihood=L, nstage=ns.value, evidence=ev.value, gamma=gt[:ns.value].copy(),
                    t_submit=times[0], t_start=times[1], t_done=times[2])

    def set_max_running(self, n):
        """At most n requests (1 .. 256) run at once."""
        _check(lib().smcf_server_set_max_running(self.h, int(n)))

    def status(self):
        q, r, b = C.c_int(), C.c_int(), C.c_size_t()
        _check(lib().smcf_server_status(self.h, C.byref(q), C.byref(r), C.byref(b)))
        return dict(queued=q.value, running=r.value, bytes_in_use=b.value)

    def close(self):
        if self.h:
            lib().smcf_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Problem:
    """Observations + fault mesh + inner-SMC settings resident on the device."""

    def __init__(self, ctx, nxi_ls, neta_ls, xmin, xmax, zmin, obs, max_slip, nparticle_slip,
                 mover=MOVER_GIBBS, gibbs_sweeps=5, max_batch=1, fix_xbend=False, xbend=(),
                 precision=PRECISION_FP64, hmc_ntau=10, hmc_accept=0.65):
        self.ctx = ctx
        self.desc, self._keep = problem_desc(nxi_ls, neta_ls, xmin, xmax, zmin, obs, max_slip, nparticle_slip,
                                             mover, gibbs_sweeps, fix_xbend, xbend, precision, hmc_ntau, hmc_accept)
        h = C.c_void_p()
        _check(lib().smcf_problem_create(ctx.h, C.byref(self.desc), int(max_batch), C.byref(h)))
        self.h = h
        dims = [C.c_int() for _ in range(5)]
        _check(lib().smcf_problem_dims(h, *[C.byref(x) for x in dims]))
        self.nnode, self.ndof, self.ndim, self.ndim_fault, self.nobs = [x.value for x in dims]
        self.P = int(nparticle_slip)
        self.max_batch = int(max_batch)

    def close(self):
        if self.h:
            lib().smcf_problem_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _theta(self, theta):
        th = np.asfortranarray(np.asarray(theta, dtype=np.float64))
        if th.ndim == 1:
            th = th.reshape(-1, 1, order="F")
        if th.shape[0] != self.ndim_fault:
            raise ValueError("theta must have %d rows" % self.ndim_fault)
        return th

    def calc_greens_func(self, theta):
        th = self._theta(theta)
        B = th.shape[1]
        g = np.zeros((self.nobs, self.ndim, B), order="F")
        _check(lib().smcf_calc_greens_func(self.h, B, _dp(th), _dp(g)))
        return g[:, :, 0] if B == 1 else g

    def discretize_fault(self, theta):
        """discretize_fault + the dense luni / lmat of gen_laplacian for one theta."""
        th = self._theta(theta)
        npatch = int(self._keep["nxi"].astype(int) @ self._keep["neta"].astype(int))
        cny = np.zeros((4, npatch), np.int32, order="F")
        coor = np.zeros((2, self.nnode), order="F")
        n2e = np.zeros((4, self.nnode), np.int32, order="F")
        n2es = np.zeros(self.nnode, np.int32)
        idd = np.zeros(self.ndof, np.int32)
        luni = np.zeros((self.nnode, self.nnode), order="F")
        lmat = np.zeros((2 * self.nnode, 2 * self.ndof), order="F")
        _check(lib().smcf_discretize_fault(self.h, _dp(th), _ip(cny), _dp(coor), _ip(n2e), _ip(n2es), _ip(idd),
                                           _dp(luni), _dp(lmat)))
        return dict(cny=cny, coor=coor, n2e=n2e, n2es=n2es, id_dof=idd, luni=luni, lmat=lmat)

    def laplacian(self, theta):
        th = self._theta(theta)
        li = np.zeros((9, 2 * self.nnode), np.int32, order="F")
        lv = np.zeros((9, 2 * self.nnode), order="F")
        lti = np.zeros((9, 2 * self.ndof), np.int32, order="F")
        ltv = np.zeros((9, 2 * self.ndof), order="F")
        ll = np.zeros((self.ndim, self.ndim), order="F")
        _check(lib().smcf_laplacian(self.h, _dp(th), _ip(li), _dp(lv), _ip(lti), _dp(ltv), _dp(ll)))
        return dict(li=li, lv=lv, lti=lti, ltv=ltv, ll=ll)

    def slip_calc_gsd_gsg(self, theta):
        th = self._theta(theta)
        B = th.shape[1]
        H = np.zeros((self.ndim, self.ndim, B), order="F")
        b = np.zeros((self.ndim, B), order="F")
        dsd = np.zeros(B)
        _check(lib().smcf_slip_calc_gsd_gsg(self.h, B, _dp(th), _dp(H), _dp(b), _dp(dsd)))
        if B == 1:
            return H[:, :, 0], b[:, 0], float(dsd[0])
        return H, b, dsd

    def slip_smc_exec(self, theta, key):
        th = self._theta(theta)
        X = np.zeros((self.ndim, self.P), order="F")
        L = np.zeros(self.P)
        PR = np.zeros(self.P)
        mean = np.zeros(self.ndim)
        ns = C.c_int()
        nl = C.c_double()
        gt = np.zeros(202)
        _check(lib().smcf_slip_smc_exec(self.h, _dp(th), int(key), _dp(X), _dp(L), _dp(PR), _dp(mean),
                                        C.byref(ns), C.byref(nl), _dp(gt)))
        return dict(particles=X, likelihood=L, prior=PR, mean=mean, nstage=ns.value, neglog=nl.value,
                    gamma=gt[:ns.value].copy())

    def _cols(self, particles):
        X = _f64(particles)
        if X.ndim == 1:
            X = X.reshape(-1, 1, order="F")
        if X.shape[0] != self.ndim:
            raise ValueError("particles must have %d rows" % self.ndim)
        return X

    def slip_calc_likelihood(self, theta, particles):
        """slip_calc_likelihood for each column of particles (ndim x n) on geometry theta."""
        X = self._cols(particles)
        out = np.zeros(X.shape[1])
        _check(lib().smcf_slip_calc_likelihood(self.h, _dp(self._theta(theta)), X.shape[1], _dp(X), _dp(out)))
        return out

    def slip_calc_prior(self, theta, particles):
        X = self._cols(particles)
        out = np.zeros(X.shape[1])
        _check(lib().smcf_slip_calc_prior(self.h, _dp(self._theta(theta)), X.shape[1], _dp(X), _dp(out)))
        return out

    def slip_calc_grad(self, theta, gamma, particles):
        X = self._cols(particles)
        g = np.zeros(X.shape, order="F")
        _check(lib().smcf_slip_calc_grad(self.h, _dp(self._theta(theta)), float(gamma), X.shape[1], _dp(X), _dp(g)))
        return g

    def slip_leapfrog(self, theta, gamma, cur, p0, cov_diag, dtau, ntau):
        """n independent slip_leapfrog trajectories; returns (cand, p, ham_cur, ham_cand)."""
        X = self._cols(cur)
        Pm = self._cols(p0)
        n = X.shape[1]
        cand = np.zeros(X.shape, order="F")
        pout = np.zeros(X.shape, order="F")
        hc = np.zeros(n)
        hn = np.zeros(n)
        lk = np.zeros(n)
        pr = np.zeros(n)
        _check(lib().smcf_slip_leapfrog(self.h, _dp(self._theta(theta)), float(gamma), n, _dp(X), _dp(Pm),
                                        _dp(_f64(cov_diag)), float(dtau), int(ntau), _dp(cand), _dp(pout), _dp(hc),
                                        _dp(hn), _dp(lk), _dp(pr)))
        self.last_leapfrog_like_prior = (lk, pr)
        return cand, pout, hc, hn

    def slip_calc_gsd_gsg_gmat(self, gmat, sigma_sar2, sigma_gnss2):
        """slip_calc_gsd_gsg from a given Green's matrix: (H, b, dsd)."""
        H = np.zeros((self.ndim, self.ndim), order="F")
        b = np.zeros(self.ndim)
        dsd = C.c_double()
        _check(lib().smcf_slip_calc_gsd_gsg_gmat(self.h, _dp(_f64(gmat)), float(sigma_sar2), float(sigma_gnss2),
                                                 _dp(H), _dp(b), C.byref(dsd)))
        return H, b, dsd.value

    def slip_gen_init_particles(self, theta, key):
        """The prior Gibbs chain of slip_gen_init_particles: (particles ndim x P, likelihood, prior)."""
        X = np.zeros((self.ndim, self.P), order="F")
        L = np.zeros(self.P)
        R = np.zeros(self.P)
        _check(lib().smcf_slip_gen_init_particles(self.h, _dp(self._theta(theta)), C.c_uint64(key), _dp(X), _dp(L),
                                                  _dp(R)))
        return X, L, R

    def hmc_accept_state(self):
        """MOVER_HMC_ACCEPT: (next stage's step size, mean acceptance of the last stage) after slip_smc_exec."""
        dt, ab = C.c_double(), C.c_double()
        _check(lib().smcf_slip_hmc_accept_state(self.h, C.byref(dt), C.byref(ab)))
        return dt.value, ab.value

    def work_mcmc_sampling(self, seed, stage, gamma, range_, cov_chol, assigned, particles, max_batch=None,
                           gsmc=None, q=None):
        """work_mcmc_sampling (smc_fault.f90:639-783) of one outer stage: (particles_new, likelihood_new, q_new,
        accepted).  gsmc = (qmean, qcov_inv, log_qcov_det) with q the parents' -log q."""
        X = _f64(particles)
        nd, N = X.shape
        a = _i32(assigned)
        Xn = np.zeros((nd, N), order="F")
        Ln = np.zeros(N)
        qn = np.zeros(N)
        acc = C.c_int()
        if gsmc is not None:
            qm, qi = _f64(gsmc[0]), _f64(gsmc[1])
            args = (_dp(_f64(q)), _dp(qm), _dp(qi), float(gsmc[2]))
        else:
            args = (None, None, None, 0.0)
        _check(lib().smcf_work_mcmc_sampling(self.h, C.c_uint64(seed), int(stage), float(gamma), _dp(_f64(range_)),
                                             _dp(_f64(cov_chol)), N, _ip(a), _dp(X), *args, int(max_batch or N),
                                             _dp(Xn), _dp(Ln), _dp(qn), C.byref(acc)))
        return Xn, Ln, qn, acc.value

    def set_outer_stage(self, mode):
        """Where fault_smc_exec runs each stage's reweight and resample: 0 host (default), 1 device with the host
        sort, 2 device with the device radix sort."""
        _check(lib().smcf_fault_smc_set_outer_stage(self.h, int(mode)))

    def set_output_dir(self, output_dir):
        """Per-stage dumps <output_dir><iter>.csv of fault_smc_exec (None stops them)."""
        _check(lib().smcf_fault_smc_set_output_dir(self.h, output_dir.encode() if output_dir else None))

    def fault_smc_exec(self, range_, N, seed, max_stages=100, gsmc=None, call_batch=None):
        """Outer SMC over the geometry (smc_fault.f90:785-1075, one rank).

        gsmc = (qmean, qcov_chol, qcov_inv, log_qcov_det) runs the gsmc branch: initial particles from
        N(qmean, qcov), tempering of L - q, the mixed MH ratio (qcov_chol is the lower Cholesky factor,
        as main.f90:332 passes it).  call_batch: geometries per library call (default: a whole wave)."""
        r = _f64(range_)
        cb = int(call_batch or N)
        X = np.zeros((self.ndim_fault, N), order="F")
        L = np.zeros(N)
        ns = C.c_int()
        ev = C.c_double()
        gt = np.zeros(max_stages + 1)
        if gsmc is None:
            _check(lib().smcf_fault_smc_exec(self.h, _dp(r), int(N), int(seed), int(max_stages), cb,
                                             _dp(X), _dp(L), C.byref(ns), C.byref(ev), _dp(gt)))
        else:
            qm, qc, qi = (_f64(a) for a in gsmc[:3])
            nd = self.ndim_fault
            if qm.size != nd or qc.size != nd * nd or qi.size != nd * nd:
                raise SmcfError("gsmc: qmean / qcov shapes do not match ndim_fault")
            _check(lib().smcf_fault_smc_exec_gsmc(self.h, _dp(r), int(N), int(seed), int(max_stages),
                                                  cb, _dp(qm), _dp(qc), _dp(qi), float(gsmc[3]),
                                                  _dp(X), _dp(L), C.byref(ns), C.byref(ev), _dp(gt)))
        return dict(particles=X, likelihood=L, nstage=ns.value, evidence=ev.value, gamma=gt[:ns.value].copy())

    def slip_mean(self, thetas, keys):
        """Plain mean of the final slip particles and -log Z for each column of `thetas` (at most max_batch
        columns per call; calc_slip_map's loop, main.f90:600-640)."""
        th = self._theta(thetas)
        B = th.shape[1]
        k = np.ascontiguousarray(keys, dtype=np.uint64)
        mean = np.zeros((self.ndim, B), order="F")
        nl = np.zeros(B)
        _check(lib().smcf_slip_mean_batch(self.h, B, _dp(np.asfortranarray(th)), k.ctypes.data_as(_U64), _dp(mean),
                                          _dp(nl)))
        return mean, nl

    def batch_times(self, B):
        """Seconds from the start of the last batch call to each of its B geometries' results."""
        t = np.zeros(B)
        _check(lib().smcf_batch_times(self.h, int(B), _dp(t)))
        return t

    def batch_admit_times(self, B):
        """Seconds from the start of the last batch call to each geometry's set-up launch (0: first fill)."""
        t = np.zeros(B)
        _check(lib().smcf_batch_admit_times(self.h, int(B), _dp(t)))
        return t

    def fault_calc_likelihood(self, thetas, keys):
        """-log Z for each column of `thetas` (ndim_fault x B), RNG key per column."""
        th = self._theta(thetas)
        B = th.shape[1]
        k = np.ascontiguousarray(keys, dtype=np.uint64)
        out = np.zeros(B)
        ns = np.zeros(B, dtype=np.int32)
        # the library streams the B geometries through the problem's max_batch device slots
        _check(lib().smcf_fault_calc_likelihood_batch(self.h, B, _dp(np.asfortranarray(th)),
                                                      k.ctypes.data_as(_U64), _dp(out), _ip(ns)))
        return out, ns


def _ptrs(arrays):
    """C array of the data pointers of numpy arrays (None -> NULL)."""
    out = (C.c_void_p * len(arrays))()
    for i, a in enumerate(arrays):
        out[i] = None if a is None else a.ctypes.data
    return out


class Group:
    """Problems (Gibbs mover, one context) whose inner SMCs run in one stage loop: every stage kernel covers the
    instances of all members, whatever their meshes, observation sets and particle counts, and each member's
    results equal those of the same call on the member alone."""

    def __init__(self, ctx, problems, graph=True):
        self.ctx = ctx
        self.problems = list(problems)
        arr = (C.c_void_p * len(self.problems))(*[p.h.value for p in self.problems])
        h = C.c_void_p()
        _check(lib().smcf_group_create(ctx.h, len(self.problems), arr, C.byref(h)))
        self.h = h
        self.set_graph(graph)

    def set_graph(self, on):
        """Capture each call's stage launches into a HIP graph (on) or launch them one by one."""
        _check(lib().smcf_group_set_graph(self.h, int(bool(on))))

    def fault_calc_likelihood(self, thetas, keys):
        """-log Z and stage counts of each member's geometries: thetas[i] (ndim_fault x B_i), keys[i] (B_i)."""
        n = len(self.problems)
        th = [self.problems[i]._theta(thetas[i]) if len(keys[i]) else None for i in range(n)]
        ks = [np.ascontiguousarray(k, dtype=np.uint64) for k in keys]
        B = np.array([len(k) for k in ks], dtype=np.int32)
        out = [np.zeros(b) for b in B]
        ns = [np.zeros(b, dtype=np.int32) for b in B]
        _check(lib().smcf_group_calc_likelihood_batch(self.h, _ip(B), _ptrs(th), _ptrs(ks), _ptrs(out), _ptrs(ns)))
        return out, ns

    def fault_smc_exec(self, ranges, Ns, seeds, max_stages=100):
        """Each member's outer SMC (smcf_fault_smc_exec) at once: a list of dicts as Problem.fault_smc_exec returns
        (without the gamma trace)."""
        n = len(self.problems)
        rg = [_f64(r) for r in ranges]
        Nv = np.array(Ns, dtype=np.int32)
        sd = np.array(seeds, dtype=np.uint64)
        X = [np.zeros((p.ndim_fault, int(N)), order="F") for p, N in zip(self.problems, Nv)]
        L = [np.zeros(int(N)) for N in Nv]
        ns = np.zeros(n, dtype=np.int32)
        ev = np.zeros(n)
        _check(lib().smcf_group_fault_smc_exec(self.h, _ptrs(rg), _ip(Nv), sd.ctypes.data_as(_U64), int(max_stages),
                                               _ptrs(X), _ptrs(L), _ip(ns), _dp(ev)))
        return [dict(particles=X[i], likelihood=L[i], nstage=int(ns[i]), evidence=float(ev[i])) for i in range(n)]

    def close(self):
        if self.h:
            lib().smcf_group_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
