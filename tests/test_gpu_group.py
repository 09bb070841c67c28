"""Groups (smcf_group_*): independent inversions with different meshes, observation sets, faults, particle counts
and arithmetic forms in one stage loop; every member's results equal those of the same call on the member alone,
with the stage captured in a HIP graph or launched kernel by kernel.  Needs an MI355X."""
import sys
import time

import numpy as np
import pytest

import smc_inversion_fort_amd as S
from smc_inversion_fort_amd import synthetic as syn

pytestmark = pytest.mark.gpu

FORMS = [S.PRECISION_FP64, S.PRECISION_MIXED_BF16, S.PRECISION_MIXED]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def device_greens(ctx, frame):
    xmin, xmax, zmin, _ = frame

    def greens(th, pts, uv, nx, ne):
        obs0 = dict(obs_points=pts, obs_unitvec=uv, obs_sigma=np.ones(pts.shape[1]), dvec=np.zeros(pts.shape[1]),
                    nsar_total=0, ngnss=pts.shape[1] // 3)
        p = S.Problem(ctx, [nx], [ne], xmin, xmax, zmin, obs0, syn.MAX_SLIP, 8)
        g = p.calc_greens_func(th)
        p.close()
        return g
    return greens


def build(ctx, inst, precision, max_batch, sweeps=3):
    frame = inst["frame"]
    obs = syn.make_observations("c2", inst["seed"], device_greens(ctx, frame), nsta=inst["nsta"], sigma=inst["sigma"],
                                nxi=inst["nxi"], neta=inst["neta"], frame=frame)
    xmin, xmax, zmin, _ = frame
    return S.Problem(ctx, [inst["nxi"]], [inst["neta"]], xmin, xmax, zmin, obs, syn.MAX_SLIP, inst["P"],
                     mover=S.MOVER_GIBBS, gibbs_sweeps=sweeps, max_batch=max_batch, precision=precision)


def keys_of(k, n):
    return [S.inner_key(5000 + k, 0, 0, j) for j in range(n)]


def test_group_matches_solo_every_form():
    """Twelve inversions (4x4 .. 10x10 patches, 30 .. 120 stations, P 96 .. 400, the three forms) with four
    geometries each through three slots: the group's -log Z and stage counts equal each member's solo call bit for
    bit, with and without the stage graph, and a member with no geometries leaves the others unchanged."""
    insts = syn.batch_instances(12, seed=3, patches=(4, 10), nsta=(30, 120), P=(96, 400))
    ctx = S.Context(0)
    probs = [build(ctx, it, FORMS[k % 3], 3) for k, it in enumerate(insts)]
    ths = [syn.sample_thetas(4, 40 + k, frame=it["frame"]) for k, it in enumerate(insts)]
    keys = [keys_of(k, 4) for k in range(len(insts))]
    say("case: group of 12, ndim", [p.ndim for p in probs], "P", [p.P for p in probs])
    solo = [p.fault_calc_likelihood(th, kk) for p, th, kk in zip(probs, ths, keys)]
    g = S.Group(ctx, probs, graph=True)
    t0 = time.perf_counter()
    nl, ns = g.fault_calc_likelihood(ths, keys)
    say("group (graph) %.2f s; stages" % (time.perf_counter() - t0), [list(x) for x in ns])
    for k in range(len(probs)):
        np.testing.assert_array_equal(nl[k], solo[k][0], err_msg="member %d" % k)
        np.testing.assert_array_equal(ns[k], solo[k][1], err_msg="member %d" % k)
    g.set_graph(False)
    nl2, ns2 = g.fault_calc_likelihood(ths, keys)
    for k in range(len(probs)):
        np.testing.assert_array_equal(nl2[k], solo[k][0], err_msg="member %d (no graph)" % k)
        np.testing.assert_array_equal(ns2[k], solo[k][1], err_msg="member %d (no graph)" % k)
    # member 0 without geometries
    g.set_graph(True)
    nl3, _ = g.fault_calc_likelihood([ths[0][:, :0]] + ths[1:], [[]] + keys[1:])
    assert len(nl3[0]) == 0
    for k in range(1, len(probs)):
        np.testing.assert_array_equal(nl3[k], solo[k][0], err_msg="member %d (member 0 empty)" % k)
    g.close()
    for p in probs:
        p.close()
    ctx.close()


def test_group_of_64_batching_instances():
    """64 inversions drawn from the batching configuration's distribution (faults of any strike, dip 20 .. 80,
    length 10 .. 60 km; patch grids and station counts scaled to 4x4 .. 12x12 and 50 .. 200, P 128 .. 512),
    split-BF16 form, two geometries each: every member equals its solo run."""
    insts = syn.batch_instances(64, seed=11, patches=(4, 12), nsta=(50, 200), P=(128, 512))
    ctx = S.Context(0)
    probs = [build(ctx, it, S.PRECISION_MIXED_BF16, 2) for it in insts]
    ths = [syn.sample_thetas(2, 70 + k, frame=it["frame"]) for k, it in enumerate(insts)]
    keys = [keys_of(100 + k, 2) for k in range(len(insts))]
    say("case: group of 64, ndim %d .. %d, P %d .. %d" % (min(p.ndim for p in probs), max(p.ndim for p in probs),
                                                          min(p.P for p in probs), max(p.P for p in probs)))
    t0 = time.perf_counter()
    solo = [p.fault_calc_likelihood(th, kk) for p, th, kk in zip(probs, ths, keys)]
    t_solo = time.perf_counter() - t0
    g = S.Group(ctx, probs)
    t0 = time.perf_counter()
    nl, ns = g.fault_calc_likelihood(ths, keys)
    t_group = time.perf_counter() - t0
    say("64 members one after another %.2f s, as one group %.2f s" % (t_solo, t_group))
    for k in range(len(probs)):
        np.testing.assert_array_equal(nl[k], solo[k][0], err_msg="member %d" % k)
        np.testing.assert_array_equal(ns[k], solo[k][1], err_msg="member %d" % k)
    g.close()
    for p in probs:
        p.close()
    ctx.close()


def test_group_outer_smc_matches_solo():
    """Four inversions' outer SMCs (N = 24, two stages) on the group's threads, their waves served by one stage
    loop: particles, likelihoods, stage counts and evidence equal each member's solo smcf_fault_smc_exec."""
    insts = syn.batch_instances(4, seed=21, patches=(4, 8), nsta=(30, 80), P=(96, 256))
    ctx = S.Context(0)
    probs = [build(ctx, it, FORMS[k % 2], 16) for k, it in enumerate(insts)]
    ranges = [syn.prior_range(frame=it["frame"]) for it in insts]
    N, stages = 24, 2
    seeds = [900 + k for k in range(len(insts))]
    say("case: group outer SMC of 4, ndim", [p.ndim for p in probs])
    solo = [p.fault_smc_exec(r, N, sd, max_stages=stages) for p, r, sd in zip(probs, ranges, seeds)]
    g = S.Group(ctx, probs)
    t0 = time.perf_counter()
    res = g.fault_smc_exec(ranges, [N] * len(probs), seeds, max_stages=stages)
    say("group outer SMC %.2f s; evidence" % (time.perf_counter() - t0), [round(r["evidence"], 6) for r in res])
    for k in range(len(probs)):
        assert res[k]["nstage"] == solo[k]["nstage"]
        assert res[k]["evidence"] == solo[k]["evidence"]
        np.testing.assert_array_equal(res[k]["particles"], solo[k]["particles"], err_msg="member %d" % k)
        np.testing.assert_array_equal(res[k]["likelihood"], solo[k]["likelihood"], err_msg="member %d" % k)
    g.close()
    for p in probs:
        p.close()
    ctx.close()


def test_group_members_vs_oracle():
    """Three FP64 members with different meshes, station sets and faults in one group call: every geometry's -log Z
    and stage count equal the CPU oracle's inner SMC of that member (1e-8 relative)."""
    import oracle as O
    insts = syn.batch_instances(3, seed=31, patches=(4, 7), nsta=(20, 50), P=(96, 160))
    ctx = S.Context(0)
    probs, orcs, ths, keys = [], [], [], []
    for k, it in enumerate(insts):
        xmin, xmax, zmin, _ = it["frame"]
        f = O.Fault([it["nxi"]], [it["neta"]], xmin, xmax, zmin)

        def greens(th, pts, uv, nx, ne, f=f):
            return f.greens(th, pts, uv, 0, pts.shape[1] // 3)
        obs = syn.make_observations("c2", it["seed"], greens, nsta=it["nsta"], sigma=it["sigma"], nxi=it["nxi"],
                                    neta=it["neta"], frame=it["frame"])
        probs.append(S.Problem(ctx, [it["nxi"]], [it["neta"]], xmin, xmax, zmin, obs, syn.MAX_SLIP, it["P"],
                               mover=S.MOVER_GIBBS, gibbs_sweeps=3, max_batch=2))
        orcs.append(O.ProblemHolder(f, obs, syn.MAX_SLIP, it["P"], mover=1, gibbs_sweeps=3))
        ths.append(syn.sample_thetas(2, 60 + k, frame=it["frame"]))
        keys.append([O.inner_key(61 + k, 0, 0, j) for j in range(2)])
    say("case: group of 3 vs oracle, ndim", [p.ndim for p in probs])
    g = S.Group(ctx, probs)
    nl, ns = g.fault_calc_likelihood(ths, keys)
    for k in range(3):
        for j in range(2):
            ref, nst = orcs[k].fault_likelihood(ths[k][:, j], keys[k][j])
            say("member", k, "geometry", j, nl[k][j], ref, ns[k][j], nst)
            assert ns[k][j] == nst
            assert nl[k][j] == pytest.approx(ref, rel=1e-8)
    g.close()
    for p in probs:
        p.close()
    ctx.close()
