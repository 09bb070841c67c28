"""Benchmark: outer-particle evaluations per second of the nested SMC at 16x16 patches, FP64 and mixed.

The unit of work is one fault_calc_likelihood (smc_fault.f90:66-292) evaluation, as the outer SMC issues
them in its waves (2N per stage, smc_fault.f90:695,734): for one geometry particle the Green's matrix, the
Gram matrices and a full inner slip SMC (P = 2000 particles, truncated-MVN Gibbs, 5 sweeps per stage) tempered
to gamma = 1.  At 16x16 patches and 1200 observations (N = 10 000 outer particles) one outer stage is a wave of
20 000 evaluations streamed through the device's slots (a slot whose inner SMC has finished takes the next
geometry).  The bench times a stretch of such a stream:

  one library call streams  W*Ew + K*E + T  geometries (drawn from the prior) through S slots;
  warm-up step w  = the next Ew completions (Ew >= 2S/W by default, so the warm-up spans two fills of every slot: the
                    first fill's geometries start together, and the slowest of them finish together at the
                    reference's 200-stage limit; a timed window inside that burst overstates the steady rate);
  timed step k    = the next E completions, k = 1..K;
  tail            = T more geometries (default T = S) that keep every slot busy until the last timed
                    completion, as the rest of a 20 000-evaluation wave would.

Step boundaries are stage boundaries at which the host has synchronised with the device and read the finished
geometries' -log Z back, so `ms_per_step * steps` is wall time between two device synchronisations, and `value`
= K*E / that time.  The whole stream's average (ramp-up and drain included) is reported beside it.

Arithmetic forms of the same workload, on the same geometries and keys:
  fp64   every operation in FP64 (the headline `value`);
  mixed  Gibbs products Q S in FP32 on the matrix cores, single-precision truncated-normal draws;
         conditional means, likelihood, weights and evidence in FP64;
  bf16   as mixed, with Q S formed from split-BF16 operands (hi.hi + hi.lo + lo.hi, FP32 accumulation).
The mixed forms are reported under "mixed" (the faster one) and "mixed_forms", each with its speed-up over
FP64 and its -log Z difference from the FP64 form.

Prints one JSON line with `roofline` (the Gibbs-sweep kernel, timed with HIP events on the library's stream),
`cpu_baseline` (the C oracle on this host's cores, a time-bounded sample of whole evaluations of the same
workload) and `serving` (Poisson arrivals of inversion requests admitted by a device-memory budget).  A
wall-clock guard skips legs that would not fit --budget-s and still prints the line.
"""
import argparse
import json
import os
import sys
import time


ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "outer-SMC particles/s at 16x16 patches, mixed vs FP64; p50/p99 solve latency"
FP64_PEAK_TFLOPS = 78.6       # MI355X FP64 vector = FP64 matrix
FP32_MFMA_PEAK_TFLOPS = 157.3  # MI355X dense FP32 matrix
BF16_MFMA_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 matrix
T0 = time.perf_counter()


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--evals", type=int, default=96,
                    help="geometry evaluations completed per timed step (20 x 96 = 1920 timed completions, about "
                         "50 s of the FP64 form; the count of completions in such a window fluctuates by about "
                         "1/sqrt(1920) = 2.3 %%, against 2.8 %% at 64 per step)")
    ap.add_argument("--slots", type=int, default=None,
                    help="device slots (inner SMCs in flight; default 1024, and 256 for c3, whose 1024 slots would "
                         "need 278 GB)")
    ap.add_argument("--warmup-fills", type=float, default=2.0,
                    help="slot fills the warm-up spans (Ew >= fills * slots / warmup); 1 halves the stream of the "
                         "32x16-patch HMC workload, whose library call otherwise runs for minutes")
    ap.add_argument("--tail", type=int, default=-1,
                    help="geometries streamed after the last timed one to keep the slots busy (-1: --slots)")
    ap.add_argument("--cpu-seconds", type=float, default=30.0,
                    help="time budget of the CPU baseline: whole oracle evaluations are run until it is spent")
    ap.add_argument("--cpu-min-evals", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--precision", choices=["all", "fp64", "mixed", "bf16"], nargs="+", default=["fp64", "bf16"],
                    help="arithmetic forms to run (the first one run gives `value`)")
    ap.add_argument("--serve-seconds", type=float, default=20.0,
                    help="length of the serving run's arrival stream; 0 skips it")
    ap.add_argument("--serve-rate", type=float, default=2.0, help="mean request arrivals per second")
    ap.add_argument("--serve-sizes", choices=["bench", "c5"], default="bench",
                    help="request meshes and station counts: bench (4x4..12x8 patches, 50..400 stations) or c5 "
                         "(4x4..32x16 patches, 50..1000 stations)")
    ap.add_argument("--budget-s", type=float, default=480.0,
                    help="wall-clock budget of the whole run: a leg whose estimate does not fit is skipped")
    ap.add_argument("--batch-n", type=int, default=64, help="inversions of the batching leg (0 skips it)")
    ap.add_argument("--batch-slots", type=int, default=16, help="device slots per inversion in the batching leg")
    ap.add_argument("--batch-evals", type=int, default=32, help="geometries per inversion in the batching leg")
    ap.add_argument("--no-batch-outer", action="store_true",
                    help="c4: leave out the inversions' outer SMCs on the group's threads (the group call, its "
                         "kernel-by-kernel repeat and the solo sample still run)")
    ap.add_argument("--workload", choices=["c2", "c3", "c4", "c5"], default="c2",
                    help="c2: the headline (16x16 patches, Gibbs x5, FP64 and mixed forms); c3: 32x16 patches, "
                         "3000 observations, HMC with 10 leapfrog steps tuned to 0.65 acceptance, FP64; c4: the "
                         "batching leg alone (64 inversions of different sizes as one group) with a solo comparison; "
                         "c5: the serving leg alone (--serve-rate, --serve-seconds)")
    ap.add_argument("--autotune", action="store_true",
                    help="time the Gibbs sweep kernels' block-row variants per mesh size and form, print the table "
                         "(the library's tuning table is this output) and exit")
    ap.add_argument("--copy-probe", action="store_true",
                    help="fill and copy 1 GiB on the device first (known-size kernels that calibrate rocprofv3's "
                         "FETCH_SIZE / WRITE_SIZE against the library's kernels in the same run)")
    a = ap.parse_args()
    if a.slots is None:
        a.slots = 256 if a.workload == "c3" else 1024
    return a


def elapsed():
    return time.perf_counter() - T0


def traffic_from_profiles(form, instances_per_launch):
    """HBM bytes per launch of the form's Gibbs kernel: the per-instance bytes of the rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary committed for it (total bytes over the profiled run's instance-launches) times this run's
    mean instances per launch; null when there is no summary."""
    path = os.path.join(ROOT, "profiles", "gibbs_c2_traffic_%s.txt" % form)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("hbm_bytes_per_instance_launch:"):
            return float("%.4g" % (float(line.split(":", 1)[1]) * instances_per_launch))
    return None


FORM_NAMES = {0: "fp64", 1: "mixed", 2: "bf16"}


def keys_for(stream, n):
    """RNG keys of n geometries of one stream (the 64-bit products wrap modulo 2^64 by design)."""
    with np.errstate(over="ignore"):
        base = np.uint64(0x9E3779B97F4A7C15) * np.uint64(stream + 1)
        return (np.arange(n, dtype=np.uint64) * np.uint64(0xBF58476D1CE4E5B9)) ^ base


def stream_plan(a):
    W, K, E = a.warmup, a.steps, a.evals
    Ew = max(E, int(np.ceil(a.warmup_fills * a.slots / W))) if W > 0 else 0
    tail = a.slots if a.tail < 0 else a.tail
    return W, K, E, Ew, tail, W * Ew + K * E + tail


C3_NTAU, C3_ACCEPT = 10, 0.65


def run_form(a, S, syn, obs, precision):
    """One arithmetic form: one streamed batch call, cut into warm-up and timed steps by completion order."""
    cfg = syn.CONFIGS[a.workload]
    xmin, xmax, zmin, _ = syn.fault_frame()
    name = FORM_NAMES[precision]
    W, K, E, Ew, tail, total = stream_plan(a)
    ctx = S.Context(0)
    if a.workload == "c2":
        pb = S.Problem(ctx, [cfg["nxi"]], [cfg["neta"]], xmin, xmax, zmin, obs, syn.MAX_SLIP, cfg["P"],
                       mover=S.MOVER_GIBBS, gibbs_sweeps=cfg["sweeps"], max_batch=a.slots, precision=precision)
    else:
        pb = S.Problem(ctx, [cfg["nxi"]], [cfg["neta"]], xmin, xmax, zmin, obs, syn.MAX_SLIP, cfg["P"],
                       mover=S.MOVER_HMC_ACCEPT, hmc_ntau=C3_NTAU, hmc_accept=C3_ACCEPT, max_batch=a.slots,
                       precision=precision)
    ndim, P, sweeps = pb.ndim, cfg["P"], cfg["sweeps"]
    print("bench[%s]: %s %dx%d patches ndim=%d nobs=%d P=%d; streaming %d geometries through %d slots "
          "(warm-up %d x %d, timed %d x %d, tail %d)" % (name, a.workload, cfg["nxi"], cfg["neta"], ndim, pb.nobs, P,
                                                          total, a.slots, W, Ew, K, E, tail), flush=True)
    th = syn.sample_thetas(total, a.seed)
    keys = keys_for(a.seed, total)
    ctx.set_profiling(True)
    t0 = time.perf_counter()
    nl, ns = pb.fault_calc_likelihood(th, keys)
    wall = time.perf_counter() - t0
    t_done = pb.batch_times(total)
    t_adm = pb.batch_admit_times(total)
    n_gibbs, ms_gibbs = ctx.kernel_stats("gibbs" if a.workload == "c2" else "hmc")
    order = np.argsort(t_done, kind="stable")
    c = t_done[order]
    n_warm = W * Ew
    t_begin = c[n_warm - 1] if n_warm > 0 else 0.0
    ends = np.array([c[n_warm + (k + 1) * E - 1] for k in range(K)])
    step_s = np.diff(np.concatenate([[t_begin], ends]))
    for k in range(K):
        print("step %d: %.3f s (geometries %d..%d completed)" % (k, step_s[k], n_warm + k * E, n_warm + (k + 1) * E - 1),
              flush=True)
    timed_s = float(ends[-1] - t_begin)
    timed = order[n_warm:n_warm + K * E]
    lat = t_done[timed] - t_adm[timed]
    # moves per geometry: one per stage, except the 201st stage of a geometry that reaches the reference's
    # 200-stage limit (it returns the sentinel without moving)
    moves = float((ns - (nl >= 1e10)).sum())
    # algorithmic flops of the Gibbs kernel over the whole call: per particle-sweep 2*ndim^2 (the products Q S)
    flops = moves * P * sweeps * 2.0 * ndim * ndim
    if a.workload == "c3":
        # the HMC stage (every kernel of one stage's trajectories): per child ntau + 2 products Y = H x
        # (initial gradient, ntau kicks, final likelihood) of 2*ndim^2 flop each
        flops = moves * P * (C3_NTAU + 2) * 2.0 * ndim * ndim
        peak, kernel = (FP64_PEAK_TFLOPS, "HMC stage (k_hmc_leap: Q_g X on the %s matrix cores with the kick and "
                                          "drift in its epilogue, k_hs_work: H X in FP64, + FP64 VALU wave kernels; "
                                          "timed per stage; FP64-equivalent flops against the FP64 peak)"
                                          % ("FP64" if name == "fp64" else "split-BF16 (wide waves)"))
    elif name == "fp64":
        peak, kernel = FP64_PEAK_TFLOPS, "k_gibbs (FP64 matrix cores + FP64 VALU draws)"
    elif name == "mixed":
        peak, kernel = (FP32_MFMA_PEAK_TFLOPS,
                        "k_gibbs_mixed (FP32 matrix cores + FP32 VALU draws; peak = dense FP32 matrix)")
    else:
        # three BF16 products per FP32-equivalent product
        flops *= 3.0
        peak, kernel = (BF16_MFMA_PEAK_TFLOPS,
                        "k_gibbs_bf (split-BF16 matrix cores, 3 products per term, + FP32 VALU draws; "
                        "achieved counts the 3 BF16 products; peak = dense BF16 matrix)")
    achieved = flops / (ms_gibbs * 1e-3) / 1e12 if ms_gibbs > 0 else 0.0
    inst_per_launch = moves / max(n_gibbs, 1)
    traffic = traffic_from_profiles(name, inst_per_launch) if a.workload == "c2" else None
    avg_ms = ms_gibbs / max(n_gibbs, 1)
    roof = {"bound": "mfma", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
            "frac": round(achieved / peak, 4), "traffic": traffic, "kernel": kernel,
            "launches": n_gibbs, "avg_launch_ms": round(avg_ms, 4),
            "mean_instances_per_launch": round(inst_per_launch, 1),
            # the measured HBM bytes per launch (rocprofv3, profiles/) over this run's average launch time
            "hbm_tb_per_s": round(traffic / (avg_ms * 1e-3) / 1e12, 3) if traffic and avg_ms > 0 else None}
    out = dict(value=K * E / timed_s, ms_per_step=1e3 * timed_s / K, step_s=step_s, stages=ns, neglog=nl,
               stages_timed=ns[timed], eval_lat=(round(float(np.percentile(lat, 50)), 3),
                                                 round(float(np.percentile(lat, 99)), 3)),
               # instance_stages_per_s: the inner SMCs' stages (each a resampling, the moves and the likelihood of
               # P particles) over the stream's wall time, a rate that does not depend on how many stages the
               # geometries happen to need
               stream={"geometries": total, "wall_s": round(wall, 2), "evals_per_s": round(total / wall, 3),
                       "instance_stages_per_s": round(moves / wall, 1)},
               roofline=roof, ndim=ndim, nobs=pb.nobs, P=P, sweeps=sweeps)
    pb.close()
    ctx.close()
    print("bench[%s]: %.2f outer particles/s over %d timed evaluations (stream %.1f s, %.2f/s overall)"
          % (name, out["value"], K * E, wall, total / wall), flush=True)
    return out


def cpu_info():
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = None
    return model, os.cpu_count(), aff


def run_cpu(a, syn, obs, head):
    """The C oracle (the same FP64 workload: per-DOF Green's loop, dense LtL, Gibbs x5, OpenMP over the inner
    particles) timed on whole evaluations until --cpu-seconds is spent."""
    model, ncpu, aff = cpu_info()
    # this job's share of the host: OMP_NUM_THREADS when set (16 per GPU on the GPU hosts), else the affinity set
    threads = int(os.environ.get("OMP_NUM_THREADS") or min(aff or ncpu or 1, 16))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    import oracle as O  # loads libgomp, which reads OMP_NUM_THREADS now
    cfg = syn.CONFIGS["c2"]
    xmin, xmax, zmin, _ = syn.fault_frame()
    f = O.Fault([cfg["nxi"]], [cfg["neta"]], xmin, xmax, zmin)
    po = O.ProblemHolder(f, obs, syn.MAX_SLIP, head["P"], mover=1, gibbs_sweeps=head["sweeps"])
    th = syn.sample_thetas(max(a.cpu_min_evals, 64), a.seed)
    keys = keys_for(a.seed, th.shape[1])
    t0 = time.perf_counter()
    n = 0
    while n < th.shape[1] and (n < a.cpu_min_evals or time.perf_counter() - t0 < a.cpu_seconds):
        po.fault_likelihood(th[:, n], int(keys[n]))
        n += 1
        print("cpu: %d evaluation(s) in %.1f s" % (n, time.perf_counter() - t0), flush=True)
    ct = time.perf_counter() - t0
    return {"value": round(n / ct, 5), "unit": "outer particles/s", "cores": threads, "kind": "port",
            "sample": "the first %d geometries of the FP64 stream (c2, P=2000, Gibbs x5), whole evaluations, "
                      "%.1f s on %d OpenMP threads" % (n, ct, threads),
            "cpu_model": model, "nproc": ncpu, "affinity_cpus": aff}


def run_serving(a, S, syn):
    """Serving form, scaled to a bench run: inversion requests arrive as a Poisson stream for
    --serve-seconds at --serve-rate per second; each carries its own mesh (log-uniform 4x4 .. 12x8 patches),
    station set (log-uniform 50 .. 400 GNSS stations), noise (2 .. 10 mm) and outer particle count
    (log-uniform 64 .. 512), and runs an outer SMC of 3 stages (inner P = 512, Gibbs x3, split-BF16 form).
    Requests are admitted by a 200 GB device-memory budget and overlap on the GPU.  Latency is submission to
    posterior-ready, on the server's clock.  Observation sets are generated before the clock starts."""
    rng = np.random.default_rng(np.random.SeedSequence([a.seed, 0x5E7E]))
    arrivals = []
    t = 0.0
    while True:
        t += rng.exponential(1.0 / a.serve_rate)
        if t > a.serve_seconds:
            break
        arrivals.append(t)
    xmin, xmax, zmin, _ = syn.fault_frame()
    gctx = S.Context(0)
    reqs = []
    # request mesh and station ranges: the bench's default (4x4 .. 12x8 patches, 50 .. 400 stations) or, with
    # --serve-sizes c5, the serving configuration's own (4x4 .. 32x16 patches, 50 .. 1000 stations)
    px, pe, ps = ((12, 8, 400) if a.serve_sizes == "bench" else (32, 16, 1000))
    for k, _t in enumerate(arrivals):
        nxi = int(round(np.exp(rng.uniform(np.log(4), np.log(px)))))
        neta = int(round(np.exp(rng.uniform(np.log(4), np.log(pe)))))
        nsta = int(round(np.exp(rng.uniform(np.log(50), np.log(ps)))))
        sigma = rng.uniform(0.2, 1.0)
        N = int(round(np.exp(rng.uniform(np.log(64), np.log(512)))))

        def greens(th, pts, uv, nx, ne):
            obs0 = dict(obs_points=pts, obs_unitvec=uv, obs_sigma=np.ones(pts.shape[1]),
                        dvec=np.zeros(pts.shape[1]), nsar_total=0, ngnss=pts.shape[1] // 3)
            p = S.Problem(gctx, [nx], [ne], xmin, xmax, zmin, obs0, syn.MAX_SLIP, 8)
            g = p.calc_greens_func(th)
            p.close()
            return g
        obs = syn.make_observations("c2", 1000 + k, greens, nsta=nsta, sigma=sigma, nxi=nxi, neta=neta)
        desc = S.problem_desc([nxi], [neta], xmin, xmax, zmin, obs, syn.MAX_SLIP, 512, mover=S.MOVER_GIBBS,
                              gibbs_sweeps=3, precision=S.PRECISION_MIXED_BF16)
        reqs.append(dict(desc=desc, N=N, nsta=nsta, obs=obs, nxi=nxi, neta=neta))
    gctx.close()
    srv = S.Server(0, hbm_budget_bytes=200 << 30)
    rr = syn.prior_range()
    print("serving: %d requests over %.0f s" % (len(reqs), a.serve_seconds), flush=True)
    t0 = time.perf_counter()
    ids = []
    for k, r in enumerate(reqs):
        dt = arrivals[k] - (time.perf_counter() - t0)
        if dt > 0:
            time.sleep(dt)
        ids.append(srv.submit(r["desc"][0], rr, r["N"], 7000 + k, max_stages=3, max_batch=r["N"]))
    res = [srv.wait(i) for i in ids]
    srv.close()
    lat = np.array([x["t_done"] - x["t_submit"] for x in res])
    span = max(x["t_done"] for x in res) - min(x["t_submit"] for x in res)
    # outer-particle evaluations: N initial + 2N per completed stage (parents re-evaluated, children)
    evals = sum(r["N"] * (1 + 2 * max(x["nstage"] - 1, 0)) for r, x in zip(reqs, res))
    print("serving: p50 %.3f s  p99 %.3f s  over %d requests" % (np.percentile(lat, 50), np.percentile(lat, 99),
                                                                   len(lat)), flush=True)
    # posterior error of the served (split-BF16) results against the same requests in the FP64 form, for the two
    # requests with the fewest outer particles: the same seed drives both outer SMCs, so the differences are those
    # of the arithmetic form carried through the outer stages, where a likelihood that differs in its fifth digit
    # can change which parents the resampling picks (the negative log evidence, in nats; the posterior mean of each
    # geometry parameter in units of its FP64 posterior standard deviation)
    errs = []
    ectx = S.Context(0)
    for k in sorted(range(len(reqs)), key=lambda j: reqs[j]["N"])[:2]:
        r = reqs[k]
        p64 = S.Problem(ectx, [r["nxi"]], [r["neta"]], xmin, xmax, zmin, r["obs"], syn.MAX_SLIP, 512,
                        mover=S.MOVER_GIBBS, gibbs_sweeps=3, max_batch=r["N"], precision=S.PRECISION_FP64)
        f = p64.fault_smc_exec(rr, r["N"], 7000 + k, max_stages=3)
        nds = p64.ndim
        p64.close()
        x64, xbf = f["particles"], res[k]["particles"]
        sd = np.maximum(x64.std(axis=1), 1e-12)
        errs.append({"request": k, "N": r["N"], "ndim_slip": nds,
                     "neglog_evidence_fp64": float("%.6g" % f["evidence"]),
                     "neglog_evidence_diff": float("%.3g" % abs(res[k]["evidence"] - f["evidence"])),
                     "mean_diff_in_sd_max": float("%.3g" % float(np.max(np.abs(xbf.mean(axis=1) - x64.mean(axis=1))
                                                                          / sd)))})
    ectx.close()
    print("serving: vs the FP64 form", errs, flush=True)
    return {"requests": len(res), "seconds": a.serve_seconds, "arrival_rate_per_s": a.serve_rate,
            "error_vs_fp64": errs,
            "p50_latency_s": round(float(np.percentile(lat, 50)), 3),
            "p99_latency_s": round(float(np.percentile(lat, 99)), 3),
            "max_latency_s": round(float(lat.max()), 3),
            "requests_per_s": round(len(res) / span, 3), "outer_evals_per_s": round(evals / span, 2),
            "form": "bf16", "budget_gb": 200,
            "workload": "Poisson arrivals; per request 4x4..%dx%d patches, 50..%d stations, 2..10 mm noise, "
                        "N 64..512 outer particles, 3 outer stages, inner P=512 Gibbs x3" % (px, pe, ps)}


def device_greens(S, syn, ctx, frame):
    """G (nobs x ndim) at a geometry from the device's calc_greens_func, for generating data."""
    xmin, xmax, zmin, _ = frame

    def greens(th, pts, uv, nx, ne):
        obs0 = dict(obs_points=pts, obs_unitvec=uv, obs_sigma=np.ones(pts.shape[1]), dvec=np.zeros(pts.shape[1]),
                    nsar_total=0, ngnss=pts.shape[1] // 3)
        p = S.Problem(ctx, [nx], [ne], xmin, xmax, zmin, obs0, syn.MAX_SLIP, 8)
        g = p.calc_greens_func(th)
        p.close()
        return g
    return greens


def run_batch(a, S, syn, solo_sample=0, extended=False):
    """Batching form: a.batch_n independent inversions drawn from the seed (patch grids 4x4 .. 24x24, 50 .. 800
    stations, faults of any strike with dip 20 .. 80 deg and length 10 .. 60 km, noise 2 .. 10 mm, inner P
    500 .. 2000), split-BF16 form, Gibbs x5, as one group: every member streams a.batch_evals geometries (thetas from
    its own prior) through its a.batch_slots slots, and every stage of all members is one HIP graph launch (each
    Gibbs kernel class one launch over the instances of all members).  Value = evaluations completed per second of
    the group call (set-up of every geometry included).  solo_sample > 0 also times that many members' calls one
    after another (each alone on the device), for comparison."""
    insts = syn.batch_instances(a.batch_n, a.seed)
    ctx = S.Context(0)
    probs, ths, keys = [], [], []
    for it in insts:
        frame = it["frame"]
        obs = syn.make_observations("c2", it["seed"], device_greens(S, syn, ctx, frame), nsta=it["nsta"],
                                    sigma=it["sigma"], nxi=it["nxi"], neta=it["neta"], frame=frame)
        xmin, xmax, zmin, _ = frame
        probs.append(S.Problem(ctx, [it["nxi"]], [it["neta"]], xmin, xmax, zmin, obs, syn.MAX_SLIP, it["P"],
                               mover=S.MOVER_GIBBS, gibbs_sweeps=5, max_batch=a.batch_slots,
                               precision=S.PRECISION_MIXED_BF16))
        ths.append(syn.sample_thetas(a.batch_evals, it["seed"], frame=frame))
        keys.append(keys_for(it["seed"], a.batch_evals))
    nd = [p.ndim for p in probs]
    dev_bytes = sum(S.problem_bytes(p.desc, a.batch_slots) for p in probs)
    print("batch: %d inversions, ndim %d..%d, P %d..%d, %d slots and %d geometries each, %.1f GB"
          % (len(probs), min(nd), max(nd), min(p.P for p in probs), max(p.P for p in probs), a.batch_slots,
             a.batch_evals, dev_bytes / 1e9), flush=True)
    g = S.Group(ctx, probs, graph=True)
    t0 = time.perf_counter()
    nl, ns = g.fault_calc_likelihood(ths, keys)
    wall = time.perf_counter() - t0
    evals = len(probs) * a.batch_evals
    stages = np.concatenate(ns)
    out = {"inversions": len(probs), "evals": evals, "wall_s": round(wall, 2), "evals_per_s": round(evals / wall, 3),
           "slots_per_inversion": a.batch_slots, "form": "bf16", "graph": True,
           "ndim_range": [min(nd), max(nd)], "nobs_range": [min(p.nobs for p in probs), max(p.nobs for p in probs)],
           "inner_particles_range": [min(p.P for p in probs), max(p.P for p in probs)],
           "mean_inner_stages": round(float(stages.mean()), 2), "device_gb": round(dev_bytes / 1e9, 2),
           "workload": "%d inversions drawn from the seed: patch grids 4x4..24x24, 50..800 GNSS stations, strike "
                       "0..360, dip 20..80, length 10..60 km, 2..10 mm noise, inner P 500..2000, Gibbs x5, "
                       "split-BF16; %d geometries per inversion from its prior" % (len(probs), a.batch_evals)}
    # latency per instance: each inversion's last result, and each geometry's set-up launch to result, from the
    # start of the group call
    done = np.array([float(p.batch_times(a.batch_evals).max()) for p in probs])
    lat = np.concatenate([p.batch_times(a.batch_evals) - p.batch_admit_times(a.batch_evals) for p in probs])
    out["inversion_done_s"] = {"p50": round(float(np.percentile(done, 50)), 3),
                               "p99": round(float(np.percentile(done, 99)), 3), "max": round(float(done.max()), 3)}
    out["eval_latency_s"] = {"p50": round(float(np.percentile(lat, 50)), 3),
                             "p99": round(float(np.percentile(lat, 99)), 3)}
    print("batch: %d evaluations in %.2f s (%.2f/s), mean %.1f inner stages; inversions done p50 %.2f s, p99 %.2f s"
          % (evals, wall, evals / wall, stages.mean(), out["inversion_done_s"]["p50"], out["inversion_done_s"]["p99"]),
          flush=True)
    if extended:
        # the same call with the stage launched kernel by kernel (identical results)
        g.set_graph(False)
        t0 = time.perf_counter()
        nl2, _ = g.fault_calc_likelihood(ths, keys)
        w2 = time.perf_counter() - t0
        g.set_graph(True)
        same = all(np.array_equal(x, y) for x, y in zip(nl, nl2))
        out["no_graph"] = {"wall_s": round(w2, 2), "evals_per_s": round(evals / w2, 3), "same_results": bool(same)}
        print("batch: without the stage graph %.2f s (%.2f/s), results %s" % (w2, evals / w2,
                                                                             "identical" if same else "DIFFER"),
              flush=True)
    if extended and not a.no_batch_outer:
        # every inversion's outer SMC (fault_smc_exec on the group's threads), N scaled down, one stage
        Ns = [max(16, it["N"] // 250) for it in insts]
        ranges = [syn.prior_range(frame=it["frame"]) for it in insts]
        t0 = time.perf_counter()
        res = g.fault_smc_exec(ranges, Ns, [a.seed * 1000 + k for k in range(len(insts))], max_stages=1)
        wo = time.perf_counter() - t0
        oevals = sum(N * (1 + 2 * r["nstage"]) for N, r in zip(Ns, res))
        out["outer_smc"] = {"inversions": len(insts), "N_range": [min(Ns), max(Ns)], "max_stages": 1,
                            "wall_s": round(wo, 2), "evals": oevals, "evals_per_s": round(oevals / wo, 3),
                            "note": "N_i = max(16, N_i / 250) of the configuration's 1000..8000; evaluations counted as "
                                    "N initial + 2N per stage"}
        print("batch: 64 outer SMCs (N %d..%d, one stage) in %.2f s, %.2f evaluations/s" % (min(Ns), max(Ns), wo,
                                                                                           oevals / wo), flush=True)
    g.close()
    if solo_sample > 0:
        t0 = time.perf_counter()
        for k in range(solo_sample):
            probs[k].fault_calc_likelihood(ths[k], keys[k])
        ws = time.perf_counter() - t0
        out["solo_sample"] = {"inversions": solo_sample, "evals": solo_sample * a.batch_evals, "wall_s": round(ws, 2),
                              "evals_per_s": round(solo_sample * a.batch_evals / ws, 3)}
        print("batch: the first %d inversions one after another: %.2f evaluations/s"
              % (solo_sample, solo_sample * a.batch_evals / ws), flush=True)
    for p in probs:
        p.close()
    ctx.close()
    return out


TUNE_MESHES = [(4, 4), (8, 8), (12, 12), (16, 16), (24, 16), (24, 24)]


def run_autotune(a, S, syn):
    """For each mesh and form, the Gibbs kernel time per instance-stage of every block-row variant on the same
    inner SMCs (32 geometries through 32 slots, P = 2000, Gibbs x5), printed as one JSON line per case and a final
    table of the fastest variant per ndim."""
    xmin, xmax, zmin, _ = syn.fault_frame()
    ctx = S.Context(0)

    def greens(th, pts, uv, nx, ne):
        obs0 = dict(obs_points=pts, obs_unitvec=uv, obs_sigma=np.ones(pts.shape[1]), dvec=np.zeros(pts.shape[1]),
                    nsar_total=0, ngnss=pts.shape[1] // 3)
        p = S.Problem(ctx, [nx], [ne], xmin, xmax, zmin, obs0, syn.MAX_SLIP, 8)
        g = p.calc_greens_func(th)
        p.close()
        return g
    table = []
    for nxi, neta in TUNE_MESHES:
        obs = syn.make_observations("c2", a.seed, greens, nxi=nxi, neta=neta)
        th = syn.sample_thetas(32, a.seed)
        keys = keys_for(a.seed, 32)
        for form, rows_list in (("fp64", (16, 32)), ("bf16", (16, 32, 48))):
            prec = S.PRECISION_FP64 if form == "fp64" else S.PRECISION_MIXED_BF16
            best = None
            for rows in rows_list:
                S.set_gibbs_block_rows(prec, 1, 1 << 30, rows)
                pb = S.Problem(ctx, [nxi], [neta], xmin, xmax, zmin, obs, syn.MAX_SLIP, 2000, mover=S.MOVER_GIBBS,
                               gibbs_sweeps=5, max_batch=32, precision=prec)
                ctx.set_profiling(True)
                nl, ns = pb.fault_calc_likelihood(th, keys)
                n, ms = ctx.kernel_stats("gibbs")
                moves = float((ns - (nl >= 1e10)).sum())
                us = 1e3 * ms / moves
                print(json.dumps({"autotune": form, "nxi": nxi, "neta": neta, "ndim": pb.ndim, "rows": rows,
                                  "us_per_instance_stage": round(us, 2), "launches": n}), flush=True)
                if best is None or us < best[1]:
                    best = (rows, us)
                ndim = pb.ndim
                pb.close()
            table.append({"form": form, "ndim": ndim, "rows": best[0], "us_per_instance_stage": round(best[1], 2)})
    ctx.close()
    print(json.dumps({"autotune_table": table}), flush=True)


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py runs on one GPU only (this build has no multi-GPU path); got --gpus %d" % a.gpus,
              file=sys.stderr)
        sys.exit(2)
    if a.steps < 1 or a.evals < 1 or a.slots < 1:
        print("--steps, --evals and --slots must be >= 1", file=sys.stderr)
        sys.exit(2)
    import smc_inversion_fort_amd as S
    from smc_inversion_fort_amd import synthetic as syn

    if a.autotune:
        run_autotune(a, S, syn)
        return
    if a.workload == "c5":
        sv = run_serving(a, S, syn)
        print(json.dumps({"metric": "p50/p99 latency of inversion requests arriving at %.1f/s" % a.serve_rate,
                          "value": sv["p50_latency_s"], "unit": "s", "n_gpus": 1, "higher_is_better": False,
                          "dtype": "mixed (split-bf16 matrix + fp32 draws; fp64 means/likelihood)",
                          "data": "synthetic", "serving": sv}), flush=True)
        return
    if a.workload == "c4":
        b = run_batch(a, S, syn, solo_sample=4, extended=True)
        print(json.dumps({"metric": "outer-particle evaluations/s of %d independent inversions as one group"
                                    % a.batch_n, "value": b["evals_per_s"], "unit": "evaluations/s", "n_gpus": 1,
                          "higher_is_better": True, "dtype": "mixed (split-bf16 matrix + fp32 draws; fp64 "
                          "means/likelihood)", "data": "synthetic", "batching": b}), flush=True)
        return
    if a.copy_probe:
        import torch
        x = torch.empty(1 << 27, dtype=torch.float64, device="cuda")
        x.fill_(1.0)
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        del x, y
        print("copy probe: filled and copied 1 GiB", flush=True)

    xmin, xmax, zmin, _ = syn.fault_frame()
    gctx = S.Context(0)

    # data: d = G(theta_true) s_true + noise, G from the device
    def greens(th, pts, uv, nx, ne):
        obs0 = dict(obs_points=pts, obs_unitvec=uv, obs_sigma=np.ones(pts.shape[1]), dvec=np.zeros(pts.shape[1]),
                    nsar_total=0, ngnss=pts.shape[1] // 3)
        p = S.Problem(gctx, [nx], [ne], xmin, xmax, zmin, obs0, syn.MAX_SLIP, 8)
        g = p.calc_greens_func(th)
        p.close()
        return g
    obs = syn.make_observations(a.workload, a.seed, greens)
    gctx.close()
    if a.workload == "c3":
        # the HMC mover's forms: FP64 and split-BF16 leapfrog gradients; the CPU baseline and the serving leg belong
        # to the headline
        a.precision = [f for f in a.precision if f in ("fp64", "bf16")] or ["fp64"]
        a.no_cpu, a.serve_seconds = True, 0.0

    forms = []
    for f in a.precision:
        for g in (["fp64", "mixed", "bf16"] if f == "all" else [f]):
            if g not in forms:
                forms.append(g)
    prec = {"fp64": S.PRECISION_FP64, "mixed": S.PRECISION_MIXED, "bf16": S.PRECISION_MIXED_BF16}
    skipped = []
    res = {}
    for i, f in enumerate(forms):
        # the first form is the headline and always runs; a later one is skipped when the one before it, scaled
        # by the same stream, would overrun the budget
        if i > 0 and elapsed() + res[forms[0]]["stream"]["wall_s"] * 1.1 > a.budget_s:
            skipped.append(f)
            continue
        res[f] = run_form(a, S, syn, obs, prec[f])
    forms = [f for f in forms if f in res]
    head = res[forms[0]]
    cpu = None
    if not a.no_cpu:
        if elapsed() + a.cpu_seconds + 30 <= a.budget_s:
            cpu = run_cpu(a, syn, obs, head)
        else:
            skipped.append("cpu_baseline")
    batching = None
    if a.batch_n > 0 and a.workload == "c2":
        if elapsed() + 90 <= a.budget_s:
            batching = run_batch(a, S, syn)
        else:
            skipped.append("batching")
    serving = None
    if a.serve_seconds > 0:
        if elapsed() + 3 * a.serve_seconds + 20 <= a.budget_s:
            serving = run_serving(a, S, syn)
        else:
            skipped.append("serving")

    W, K, E, Ew, tail, total = stream_plan(a)
    metric = METRIC if a.workload == "c2" else ("outer-SMC particles/s at 32x16 patches, HMC (10 leapfrog steps, "
                                                "0.65 acceptance), FP64")
    out = {"metric": metric, "value": round(head["value"], 4), "unit": "outer particles/s", "n_gpus": 1,
           "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(head["ms_per_step"], 2),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": {"fp64": "f64", "mixed": "mixed (fp32 matrix + fp32 draws; fp64 means/likelihood)",
                     "bf16": "mixed (split-bf16 matrix + fp32 draws; fp64 means/likelihood)"}[forms[0]],
           "data": "synthetic (seeded GNSS stations, Gaussian slip bump, %d mm noise; thetas from the prior)"
                   % (5 if a.workload == "c2" else 3),
           "config": {"workload": ("c2: 16x16 patches, 400 GNSS stations (1200 obs), inner P=2000, truncated-MVN "
                                   "Gibbs x5 per stage" if a.workload == "c2" else
                                   "c3: 32x16 patches, 1000 GNSS stations (3000 obs, 3 mm), inner P=2000, HMC with "
                                   "%d leapfrog steps, step size adapted toward %.2f acceptance" % (C3_NTAU, C3_ACCEPT))
                                  + "; one outer-SMC wave stream of %d geometries through %d device slots, step = %d "
                                    "completed evaluations" % (total, a.slots, E),
                      "form": forms[0], "evals_per_step": E, "warmup_evals_per_step": Ew, "tail_evals": tail,
                      "slots": a.slots, "ndim_slip": head["ndim"], "nobs": head["nobs"],
                      "inner_particles": head["P"],
                      "mean_inner_stages": round(float(head["stages"].mean()), 2),
                      "mean_inner_stages_timed": round(float(head["stages_timed"].mean()), 2),
                      "p50_step_s": round(float(np.percentile(head["step_s"], 50)), 3),
                      "p99_step_s": round(float(np.percentile(head["step_s"], 99)), 3),
                      # per geometry: from its set-up launch to the stage boundary that returned its -log Z
                      "p50_eval_latency_s": head["eval_lat"][0], "p99_eval_latency_s": head["eval_lat"][1],
                      "stream": head["stream"]},
           "roofline": head["roofline"], "cpu_baseline": cpu, "serving": serving, "batching": batching,
           "skipped_legs": skipped, "wall_s": round(elapsed(), 1)}
    if "fp64" in res and len(res) > 1:
        f64 = res["fp64"]
        mixed = {}
        for name in forms:
            if name == "fp64":
                continue
            m = res[name]
            # geometries whose inner SMC ran past 200 stages return the reference's sentinel -log Z = 1e10
            # (smc_slip.f90:1557-1560); near that limit one form can cross it and the other not
            sent_m, sent_f = m["neglog"] >= 1e10, f64["neglog"] >= 1e10
            both = ~sent_m & ~sent_f
            rel = np.abs(m["neglog"][both] - f64["neglog"][both]) / np.abs(f64["neglog"][both])
            if rel.size == 0:  # every geometry at the sentinel in one form or the other
                rel = np.array([np.nan])
            mixed[name] = {"value": round(m["value"], 4), "unit": "outer particles/s",
                           "speedup_vs_fp64": round(m["value"] / f64["value"], 3),
                           "ms_per_step": round(m["ms_per_step"], 2),
                           "mean_inner_stages": round(float(m["stages"].mean()), 2),
                           "p50_eval_latency_s": m["eval_lat"][0], "p99_eval_latency_s": m["eval_lat"][1],
                           "stream": m["stream"],
                           "neglogz_rel_diff_vs_fp64": {"median": float("%.3g" % np.median(rel)),
                                                        "p99": float("%.3g" % np.percentile(rel, 99)),
                                                        "max": float("%.3g" % rel.max()),
                                                        "sentinel_in_one_form_only": int((sent_m != sent_f).sum())},
                           "roofline": m["roofline"]}
        best = max(mixed, key=lambda k: mixed[k]["value"])
        out["mixed"] = dict(form=best, **mixed[best])
        out["mixed_forms"] = {k: {kk: v[kk] for kk in ("value", "speedup_vs_fp64", "ms_per_step",
                                                        "neglogz_rel_diff_vs_fp64")} for k, v in mixed.items()}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
