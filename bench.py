"""Benchmark: outer-particle evaluations per second of the nested SMC at 16x16 patches.

A step is one wave of fault_calc_likelihood (smc_fault.f90:66-292) evaluations, as the outer
SMC issues them: E geometry particles drawn from the prior, streamed through S device slots
(a slot whose inner SMC has finished takes the next geometry).  Per geometry: the Green's
matrix, the Gram matrices and a full inner slip SMC (P = 2000 particles, truncated-MVN Gibbs,
5 sweeps per stage) tempered to gamma = 1.  Observations stay resident on the device; each
step uploads only the thetas and reads back one -log Z per geometry.

Prints one JSON line with `roofline` (Gibbs-sweep kernel, FP64, timed with HIP events on
the library's stream) and `cpu_baseline` (the C oracle on this host's cores, a bounded
sample of the same workload).
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
FP64_PEAK_TFLOPS = 78.6


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--evals", type=int, default=1024, help="geometry evaluations per step (one wave)")
    ap.add_argument("--slots", type=int, default=256, help="device slots (instances in flight)")
    ap.add_argument("--cpu-evals", type=int, default=1, help="oracle evaluations timed for the CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--copy-probe", action="store_true",
                    help="fill and copy 1 GiB on the device first (known-size kernels that calibrate rocprofv3's "
                         "FETCH_SIZE / WRITE_SIZE against the library's kernels in the same run)")
    return ap.parse_args()


def traffic_from_profiles():
    path = os.path.join(ROOT, "profiles", "gibbs_c2_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("hbm_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


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
    ctx = S.Context(0)
    # data: d = G(theta_true) s_true + noise, G from the device
    gen = {}

    def greens(th, pts, uv, nx, ne):
        obs0 = dict(obs_points=pts, obs_unitvec=uv, obs_sigma=np.ones(pts.shape[1]), dvec=np.zeros(pts.shape[1]),
                    nsar_total=0, ngnss=pts.shape[1] // 3)
        p = S.Problem(ctx, [nx], [ne], xmin, xmax, zmin, obs0, syn.MAX_SLIP, 8)
        g = p.calc_greens_func(th)
        p.close()
        return g
    obs = syn.make_observations("c2", a.seed, greens)
    B = a.evals
    pb = S.Problem(ctx, [cfg["nxi"]], [cfg["neta"]], xmin, xmax, zmin, obs, syn.MAX_SLIP, cfg["P"],
                   mover=S.MOVER_GIBBS, gibbs_sweeps=cfg["sweeps"], max_batch=a.slots)
    ndim, P, sweeps = pb.ndim, cfg["P"], cfg["sweeps"]

    def keys_for(step):
        base = np.uint64(0x9E3779B97F4A7C15) * np.uint64(step + 1)
        return (np.arange(B, dtype=np.uint64) * np.uint64(0xBF58476D1CE4E5B9)) ^ base

    print("bench: c2 16x16 patches ndim=%d nobs=%d P=%d evals/step=%d slots=%d" % (ndim, pb.nobs, P, B, a.slots),
          flush=True)
    for w in range(a.warmup):
        t0 = time.time()
        nl, ns = pb.fault_calc_likelihood(syn.sample_thetas(B, 1000 + w), keys_for(1000 + w))
        print("warmup step %d: %.2f s, stages mean %.1f max %d" % (w, time.time() - t0, ns.mean(), ns.max()),
              flush=True)
    ctx.set_profiling(True)
    stages = []
    lat = []
    t_start = time.perf_counter()
    for k in range(a.steps):
        t0 = time.perf_counter()
        nl, ns = pb.fault_calc_likelihood(syn.sample_thetas(B, k), keys_for(k))
        lat.append(time.perf_counter() - t0)
        stages.append(ns)
        print("step %d: %.3f s, stages mean %.1f, -logZ median %.3f" % (k, lat[-1], ns.mean(), np.median(nl)),
              flush=True)
    elapsed = time.perf_counter() - t_start
    n_gibbs, ms_gibbs = ctx.kernel_stats("gibbs")
    value = B * a.steps / elapsed
    stages = np.concatenate(stages)
    # algorithmic flops of the Gibbs kernel: per particle-sweep 2*ndim^2 (one Q-column pass)
    flops = float(stages.sum()) * P * sweeps * 2.0 * ndim * ndim
    achieved = flops / (ms_gibbs * 1e-3) / 1e12 if ms_gibbs > 0 else 0.0
    roof = {"bound": "mfma", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / FP64_PEAK_TFLOPS, 4), "traffic": traffic_from_profiles(),
            "kernel": "k_gibbs (FP64 VALU; MI355X FP64 vector = matrix peak)", "launches": n_gibbs,
            "avg_launch_ms": round(ms_gibbs / max(n_gibbs, 1), 4)}
    cpu = None
    if not a.no_cpu:
        import oracle as O
        threads = min(os.cpu_count() or 1, 16)
        os.environ.setdefault("OMP_NUM_THREADS", str(threads))
        f = O.Fault([cfg["nxi"]], [cfg["neta"]], xmin, xmax, zmin)
        po = O.ProblemHolder(f, obs, syn.MAX_SLIP, P, mover=1, gibbs_sweeps=sweeps)
        th = syn.sample_thetas(a.cpu_evals, 0)
        t0 = time.perf_counter()
        for k in range(a.cpu_evals):
            po.fault_likelihood(th[:, k], int(keys_for(0)[k]))
        ct = time.perf_counter() - t0
        cpu = {"value": round(a.cpu_evals / ct, 5), "unit": "outer particles/s", "cores": threads, "kind": "port",
               "sample": "%d c2 outer-particle evaluation(s) (same workload, P=2000, Gibbs x5) in %.1f s"
                         % (a.cpu_evals, ct)}
    out = {"metric": METRIC, "value": round(value, 4), "unit": "outer particles/s", "n_gpus": 1,
           "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(1e3 * elapsed / a.steps, 2),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (seeded GNSS stations, Gaussian slip bump, 5 mm noise)",
           "config": {"workload": "c2: 16x16 patches, 400 GNSS stations (1200 obs), inner P=2000, "
                                  "truncated-MVN Gibbs x5 per stage, FP64; waves of %d geometry particles "
                                  "streamed through %d device slots" % (B, a.slots),
                      "evals_per_step": B, "slots": a.slots, "ndim_slip": ndim, "nobs": pb.nobs,
                      "inner_particles": P,
                      "mean_inner_stages": round(float(stages.mean()), 2),
                      "p50_wave_latency_s": round(float(np.percentile(lat, 50)), 3),
                      "p99_wave_latency_s": round(float(np.percentile(lat, 99)), 3)},
           "roofline": roof, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)
    pb.close()
    ctx.close()


if __name__ == "__main__":
    main()
