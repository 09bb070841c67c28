"""Benchmark: outer-particle evaluations per second of the nested SMC at 16x16 patches, FP64 and mixed.

A step is one wave of fault_calc_likelihood (smc_fault.f90:66-292) evaluations, as the outer
SMC issues them: E geometry particles drawn from the prior, streamed through S device slots
(a slot whose inner SMC has finished takes the next geometry).  Per geometry: the Green's
matrix, the Gram matrices and a full inner slip SMC (P = 2000 particles, truncated-MVN Gibbs,
5 sweeps per stage) tempered to gamma = 1.  Observations stay resident on the device; each
step uploads only the thetas and reads back one -log Z per geometry.

Arithmetic forms of the same workload, on the same geometries and keys:
  fp64   every operation in FP64 (the headline `value`);
  mixed  Gibbs products Q S in FP32 on the matrix cores, single-precision truncated-normal draws;
         conditional means, likelihood, weights and evidence in FP64;
  bf16   as mixed, with Q S formed from split-BF16 operands (hi.hi + hi.lo + lo.hi, FP32 accumulation).
The mixed forms are reported under "mixed" (the faster one) and "mixed_forms", each with its speed-up
over FP64 and its -log Z difference from the FP64 form.

Prints one JSON line with `roofline` (Gibbs-sweep kernel, timed with HIP events on the
library's stream) and `cpu_baseline` (the C oracle on this host's cores, a bounded sample of
the same workload).
"""
import argparse
import json
import os
import sys
import time

# The serving run overlaps many requests, each with its own two HIP streams; the HIP runtime maps streams onto
# GPU_MAX_HW_QUEUES hardware queues per process (4 by default), so with more requests in flight than queues their
# kernels serialize.  16 queues let eight requests run side by side (read once, at HIP initialisation, which
# happens below this line).  The single-problem forms use two streams and are unaffected.
os.environ["GPU_MAX_HW_QUEUES"] = "16"

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "outer-SMC particles/s at 16x16 patches, mixed vs FP64; p50/p99 solve latency"
FP64_PEAK_TFLOPS = 78.6       # MI355X FP64 vector = FP64 matrix
FP32_MFMA_PEAK_TFLOPS = 157.3  # MI355X dense FP32 matrix
BF16_MFMA_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 matrix


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--evals", type=int, default=1024, help="geometry evaluations per step (one wave)")
    ap.add_argument("--warmup-evals", type=int, default=128,
                    help="geometry evaluations per warm-up step (loads the kernels and touches every code path; 0: "
                         "as --evals)")
    ap.add_argument("--slots", type=int, default=1024, help="device slots (instances in flight)")
    ap.add_argument("--cpu-evals", type=int, default=3, help="oracle evaluations timed for the CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--precision", choices=["all", "fp64", "mixed", "bf16"], nargs="+", default=["fp64", "bf16"],
                    help="arithmetic forms to run (the first one run gives `value`)")
    ap.add_argument("--serve-seconds", type=float, default=20.0,
                    help="length of the serving run (Poisson arrivals of small inversions); 0 skips it")
    ap.add_argument("--serve-rate", type=float, default=2.0, help="mean request arrivals per second")
    ap.add_argument("--copy-probe", action="store_true",
                    help="fill and copy 1 GiB on the device first (known-size kernels that calibrate rocprofv3's "
                         "FETCH_SIZE / WRITE_SIZE against the library's kernels in the same run)")
    return ap.parse_args()


def traffic_from_profiles(form):
    """HBM bytes per launch of the form's Gibbs kernel, from the rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary committed for it (null when there is none)."""
    path = os.path.join(ROOT, "profiles", "gibbs_c2_traffic_%s.txt" % form)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("hbm_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


FORM_NAMES = {0: "fp64", 1: "mixed", 2: "bf16"}


def run_form(a, S, syn, obs, precision, keys_for):
    """Warm-up and timed steps of one arithmetic form; returns its figures."""
    cfg = syn.CONFIGS["c2"]
    xmin, xmax, zmin, _ = syn.fault_frame()
    ctx = S.Context(0)
    B = a.evals
    pb = S.Problem(ctx, [cfg["nxi"]], [cfg["neta"]], xmin, xmax, zmin, obs, syn.MAX_SLIP, cfg["P"],
                   mover=S.MOVER_GIBBS, gibbs_sweeps=cfg["sweeps"], max_batch=a.slots, precision=precision)
    ndim, P, sweeps = pb.ndim, cfg["P"], cfg["sweeps"]
    name = FORM_NAMES[precision]
    print("bench[%s]: c2 16x16 patches ndim=%d nobs=%d P=%d evals/step=%d slots=%d"
          % (name, ndim, pb.nobs, P, B, a.slots), flush=True)
    BW = a.warmup_evals or B
    for w in range(a.warmup):
        t0 = time.time()
        nl, ns = pb.fault_calc_likelihood(syn.sample_thetas(BW, 1000 + w), keys_for(1000 + w, BW))
        print("warmup step %d: %.2f s, stages mean %.1f max %d" % (w, time.time() - t0, ns.mean(), ns.max()),
              flush=True)
    ctx.set_profiling(True)
    stages, lat, neglog, evl = [], [], [], []
    t_start = time.perf_counter()
    for k in range(a.steps):
        t0 = time.perf_counter()
        nl, ns = pb.fault_calc_likelihood(syn.sample_thetas(B, k), keys_for(k))
        lat.append(time.perf_counter() - t0)
        evl.append(pb.batch_times(B))
        stages.append(ns)
        neglog.append(nl)
        print("step %d: %.3f s, stages mean %.1f, -logZ median %.3f" % (k, lat[-1], ns.mean(), np.median(nl)),
              flush=True)
    elapsed = time.perf_counter() - t_start
    n_gibbs, ms_gibbs = ctx.kernel_stats("gibbs")
    stages = np.concatenate(stages)
    # algorithmic flops of the Gibbs kernel: per particle-sweep 2*ndim^2 (the products Q S)
    flops = float(stages.sum()) * P * sweeps * 2.0 * ndim * ndim
    name = FORM_NAMES[precision]
    if name == "fp64":
        peak, kernel = FP64_PEAK_TFLOPS, "k_gibbs (FP64 matrix cores + FP64 VALU draws)"
    elif name == "mixed":
        peak, kernel = (FP32_MFMA_PEAK_TFLOPS,
                        "k_gibbs_mixed (FP32 matrix cores + FP32 VALU draws; peak = dense FP32 matrix)")
    else:
        # three BF16 products per FP32-equivalent product: the comparable peak is a third of dense BF16
        flops *= 3.0
        peak, kernel = (BF16_MFMA_PEAK_TFLOPS,
                        "k_gibbs_bf (split-BF16 matrix cores, 3 products per term, + FP32 VALU draws; "
                        "achieved counts the 3 BF16 products; peak = dense BF16 matrix)")
    achieved = flops / (ms_gibbs * 1e-3) / 1e12 if ms_gibbs > 0 else 0.0
    traffic = traffic_from_profiles(name)
    avg_ms = ms_gibbs / max(n_gibbs, 1)
    roof = {"bound": "mfma", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
            "frac": round(achieved / peak, 4), "traffic": traffic, "kernel": kernel,
            "launches": n_gibbs, "avg_launch_ms": round(avg_ms, 4),
            # the measured HBM bytes per launch (rocprofv3, profiles/) over this run's average launch time
            "hbm_tb_per_s": round(traffic / (avg_ms * 1e-3) / 1e12, 3) if traffic and avg_ms > 0 else None}
    evl = np.concatenate(evl)
    out = dict(value=B * a.steps / elapsed, ms_per_step=1e3 * elapsed / a.steps, stages=stages, lat=lat,
               eval_lat=(round(float(np.percentile(evl, 50)), 3), round(float(np.percentile(evl, 99)), 3)),
               neglog=np.concatenate(neglog), roofline=roof, ndim=ndim, nobs=pb.nobs, P=P, sweeps=sweeps)
    pb.close()
    ctx.close()
    return out


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
    for k, _t in enumerate(arrivals):
        nxi = int(round(np.exp(rng.uniform(np.log(4), np.log(12)))))
        neta = int(round(np.exp(rng.uniform(np.log(4), np.log(8)))))
        nsta = int(round(np.exp(rng.uniform(np.log(50), np.log(400)))))
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
        reqs.append(dict(desc=desc, N=N, nsta=nsta))
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
    return {"requests": len(res), "seconds": a.serve_seconds, "arrival_rate_per_s": a.serve_rate,
            "p50_latency_s": round(float(np.percentile(lat, 50)), 3),
            "p99_latency_s": round(float(np.percentile(lat, 99)), 3),
            "max_latency_s": round(float(lat.max()), 3),
            "requests_per_s": round(len(res) / span, 3), "outer_evals_per_s": round(evals / span, 2),
            "form": "bf16", "budget_gb": 200,
            "workload": "Poisson arrivals; per request 4x4..12x8 patches, 50..400 stations, 2..10 mm noise, "
                        "N 64..512 outer particles, 3 outer stages, inner P=512 Gibbs x3"}


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py runs on one GPU only (this build has no multi-GPU path); got --gpus %d" % a.gpus,
              file=sys.stderr)
        sys.exit(2)
    import smc_inversion_fort_amd as S
    from smc_inversion_fort_amd import synthetic as syn

    if a.copy_probe:
        import torch
        x = torch.empty(1 << 27, dtype=torch.float64, device="cuda")
        x.fill_(1.0)
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        del x, y
        print("copy probe: filled and copied 1 GiB", flush=True)

    cfg = syn.CONFIGS["c2"]
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
    obs = syn.make_observations("c2", a.seed, greens)
    gctx.close()
    B = a.evals

    def keys_for(step, n=B):
        # the 64-bit products wrap modulo 2^64 by design (seed mixing); numpy would warn on the scalar wrap
        with np.errstate(over="ignore"):
            base = np.uint64(0x9E3779B97F4A7C15) * np.uint64(step + 1)
            return (np.arange(n, dtype=np.uint64) * np.uint64(0xBF58476D1CE4E5B9)) ^ base

    forms = []
    for f in a.precision:
        for g in (["fp64", "mixed", "bf16"] if f == "all" else [f]):
            if g not in forms:
                forms.append(g)
    prec = {"fp64": S.PRECISION_FP64, "mixed": S.PRECISION_MIXED, "bf16": S.PRECISION_MIXED_BF16}
    res = {f: run_form(a, S, syn, obs, prec[f], keys_for) for f in forms}
    head = res[forms[0]]
    cpu = None
    if not a.no_cpu:
        import oracle as O
        threads = min(os.cpu_count() or 1, 16)
        os.environ.setdefault("OMP_NUM_THREADS", str(threads))
        f = O.Fault([cfg["nxi"]], [cfg["neta"]], xmin, xmax, zmin)
        po = O.ProblemHolder(f, obs, syn.MAX_SLIP, head["P"], mover=1, gibbs_sweeps=head["sweeps"])
        th = syn.sample_thetas(a.cpu_evals, 0)
        t0 = time.perf_counter()
        for k in range(a.cpu_evals):
            po.fault_likelihood(th[:, k], int(keys_for(0)[k]))
        ct = time.perf_counter() - t0
        cpu = {"value": round(a.cpu_evals / ct, 5), "unit": "outer particles/s", "cores": threads, "kind": "port",
               "sample": "%d c2 outer-particle evaluation(s) (same workload, P=2000, Gibbs x5, FP64) in %.1f s"
                         % (a.cpu_evals, ct)}

    serving = run_serving(a, S, syn) if a.serve_seconds > 0 else None

    def latency(r):
        return (round(float(np.percentile(r["lat"], 50)), 3), round(float(np.percentile(r["lat"], 99)), 3))
    p50, p99 = latency(head)
    out = {"metric": METRIC, "value": round(head["value"], 4), "unit": "outer particles/s", "n_gpus": 1,
           "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(head["ms_per_step"], 2),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": {"fp64": "f64", "mixed": "mixed (fp32 matrix + fp32 draws; fp64 means/likelihood)",
                     "bf16": "mixed (split-bf16 matrix + fp32 draws; fp64 means/likelihood)"}[forms[0]],
           "data": "synthetic (seeded GNSS stations, Gaussian slip bump, 5 mm noise)",
           "config": {"workload": "c2: 16x16 patches, 400 GNSS stations (1200 obs), inner P=2000, "
                                  "truncated-MVN Gibbs x5 per stage; waves of %d geometry particles "
                                  "streamed through %d device slots" % (B, a.slots),
                      "form": forms[0], "evals_per_step": B, "slots": a.slots, "ndim_slip": head["ndim"],
                      "nobs": head["nobs"], "inner_particles": head["P"],
                      "mean_inner_stages": round(float(head["stages"].mean()), 2),
                      "p50_wave_latency_s": p50, "p99_wave_latency_s": p99,
                      # per geometry: from the wave's submission to the stage boundary that returned its -log Z
                      "p50_eval_latency_s": head["eval_lat"][0], "p99_eval_latency_s": head["eval_lat"][1]},
           "roofline": head["roofline"], "cpu_baseline": cpu, "serving": serving}
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
            m50, m99 = latency(m)
            mixed[name] = {"value": round(m["value"], 4), "unit": "outer particles/s",
                           "speedup_vs_fp64": round(m["value"] / f64["value"], 3),
                           "ms_per_step": round(m["ms_per_step"], 2),
                           "mean_inner_stages": round(float(m["stages"].mean()), 2),
                           "p50_wave_latency_s": m50, "p99_wave_latency_s": m99,
                           "p50_eval_latency_s": m["eval_lat"][0], "p99_eval_latency_s": m["eval_lat"][1],
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
