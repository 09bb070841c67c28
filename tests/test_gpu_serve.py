"""Serving loop (smcf_server_*): inversion requests admitted by a device-memory budget, run concurrently,
each result identical alone and in the mix.  Needs an MI355X."""
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


# (seed, patches along strike, down dip, stations): three inversions of different sizes
REQUESTS = [(3, 4, 3, 12), (5, 5, 4, 20), (7, 6, 3, 16)]
N, P, STAGES = 48, 96, 3


def request_desc(seed, nxi, neta, nsta):
    xmin, xmax, zmin, _ = syn.fault_frame()

    def greens(th, pts, uv, nx, ne):
        return O.Fault([nx], [ne], xmin, xmax, zmin).greens(th, pts, uv, 0, pts.shape[1] // 3)
    obs = syn.make_observations("c2", seed, greens, nsta=nsta, nxi=nxi, neta=neta)
    return S.problem_desc([nxi], [neta], xmin, xmax, zmin, obs, syn.MAX_SLIP, P, mover=S.MOVER_GIBBS,
                          gibbs_sweeps=3)


def same(a, b):
    assert a["nstage"] == b["nstage"]
    np.testing.assert_array_equal(a["particles"], b["particles"])
    np.testing.assert_array_equal(a["likelihood"], b["likelihood"])
    np.testing.assert_array_equal(a["gamma"], b["gamma"])
    assert a["evidence"] == b["evidence"]


def test_requests_same_alone_and_in_the_mix():
    """Each request alone (one at a time), then all submitted together: bit-identical results; the
    concurrent run admits all three at once and their execution intervals overlap."""
    rng = syn.prior_range()
    descs = [request_desc(*r) for r in REQUESTS]
    sizes = [S.problem_bytes(d[0], N) for d in descs]
    say("case: serving, request device bytes", sizes)
    srv = S.Server(0, hbm_budget_bytes=sum(sizes) + 3 * S.REQUEST_OVERHEAD_BYTES + (1 << 20))
    srv.set_max_running(4)
    alone = []
    for k, (d, _) in enumerate(descs):
        rid = srv.submit(d, rng, N, 100 + k, max_stages=STAGES, max_batch=N)
        alone.append(srv.wait(rid))
    say("alone: stages", [a["nstage"] for a in alone], "evidence", [round(a["evidence"], 6) for a in alone])
    ids = [srv.submit(d, rng, N, 100 + k, max_stages=STAGES, max_batch=N) for k, (d, _) in enumerate(descs)]
    mix = [srv.wait(i) for i in ids]
    srv.close()
    for a, m in zip(alone, mix):
        same(a, m)
    starts = [m["t_start"] for m in mix]
    dones = [m["t_done"] for m in mix]
    say("mix: admitted at", np.round(starts, 3).tolist(), "done at", np.round(dones, 3).tolist())
    # all three admitted before the first finished: they ran concurrently
    assert max(starts) < min(dones)


def test_admission_waits_for_memory():
    """With a budget that holds one request at a time, the second is admitted only after the first has
    finished and released its memory, and its result is unchanged."""
    rng = syn.prior_range()
    d0, _k0 = request_desc(*REQUESTS[0])
    d1, _k1 = request_desc(*REQUESTS[1])
    b0, b1 = S.problem_bytes(d0, N), S.problem_bytes(d1, N)
    say("case: serving admission, bytes", b0, b1)
    srv = S.Server(0, hbm_budget_bytes=max(b0, b1) + S.REQUEST_OVERHEAD_BYTES + (b0 + b1) // 4)
    i0 = srv.submit(d0, rng, N, 11, max_stages=STAGES, max_batch=N)
    i1 = srv.submit(d1, rng, N, 12, max_stages=STAGES, max_batch=N)
    r0, r1 = srv.wait(i0), srv.wait(i1)
    say("first done at %.3f, second admitted at %.3f" % (r0["t_done"], r1["t_start"]))
    assert r1["t_start"] >= r0["t_done"]
    ref = S.Server(0, hbm_budget_bytes=b1 + S.REQUEST_OVERHEAD_BYTES + (1 << 20))
    r1b = ref.wait(ref.submit(d1, rng, N, 12, max_stages=STAGES, max_batch=N))
    ref.close()
    srv.close()
    same(r1, r1b)


def test_running_cap_queues_the_rest():
    """With room for both in memory but one running place, the second request waits for the first."""
    rng = syn.prior_range()
    d0, _k0 = request_desc(*REQUESTS[0])
    d1, _k1 = request_desc(*REQUESTS[2])
    say("case: serving running cap")
    srv = S.Server(0, hbm_budget_bytes=8 << 30)
    srv.set_max_running(1)
    i0 = srv.submit(d0, rng, N, 21, max_stages=STAGES, max_batch=N)
    i1 = srv.submit(d1, rng, N, 22, max_stages=STAGES, max_batch=N)
    r0, r1 = srv.wait(i0), srv.wait(i1)
    srv.close()
    say("first done at %.3f, second admitted at %.3f" % (r0["t_done"], r1["t_start"]))
    assert r1["t_start"] >= r0["t_done"]


def device_obs(ctx, seed, nxi, neta, nsta, sigma):
    xmin, xmax, zmin, _ = syn.fault_frame()

    def greens(th, pts, uv, nx, ne):
        obs0 = dict(obs_points=pts, obs_unitvec=uv, obs_sigma=np.ones(pts.shape[1]), dvec=np.zeros(pts.shape[1]),
                    nsar_total=0, ngnss=pts.shape[1] // 3)
        p = S.Problem(ctx, [nx], [ne], xmin, xmax, zmin, obs0, syn.MAX_SLIP, 8)
        g = p.calc_greens_func(th)
        p.close()
        return g
    return syn.make_observations("c2", seed, greens, nsta=nsta, sigma=sigma, nxi=nxi, neta=neta)


def test_request_at_serving_maxima_same_alone_and_in_the_mix():
    """A request at the largest serving sizes (32x16 patches, 1000 GNSS stations, split-BF16 Gibbs) gives the
    same result alone and while two smaller requests run beside it."""
    say("case: serving, a 32x16-patch 1000-station request alone and in the mix")
    xmin, xmax, zmin, _ = syn.fault_frame()
    ctx = S.Context(0)
    big = device_obs(ctx, 31, 32, 16, 1000, 0.5)
    small = [device_obs(ctx, 32 + k, 6, 4, 60, 0.4) for k in range(2)]
    ctx.close()
    Pb, Nb = 128, 8
    db = S.problem_desc([32], [16], xmin, xmax, zmin, big, syn.MAX_SLIP, Pb, mover=S.MOVER_GIBBS, gibbs_sweeps=3,
                        precision=S.PRECISION_MIXED_BF16)
    ds = [S.problem_desc([6], [4], xmin, xmax, zmin, o, syn.MAX_SLIP, 96, mover=S.MOVER_GIBBS, gibbs_sweeps=3,
                         precision=S.PRECISION_MIXED_BF16) for o in small]
    rng = syn.prior_range()
    srv = S.Server(0, hbm_budget_bytes=200 << 30)
    srv.set_max_running(3)
    alone = srv.wait(srv.submit(db[0], rng, Nb, 41, max_stages=2, max_batch=Nb))
    say("alone: stages", alone["nstage"], "evidence", alone["evidence"], "time %.2f s"
        % (alone["t_done"] - alone["t_start"]))
    ib = srv.submit(db[0], rng, Nb, 41, max_stages=2, max_batch=Nb)
    iss = [srv.submit(d[0], rng, 24, 50 + k, max_stages=2, max_batch=24) for k, d in enumerate(ds)]
    mixed = srv.wait(ib)
    others = [srv.wait(i) for i in iss]
    srv.close()
    say("in the mix: admitted at %.3f, done at %.3f; small requests done at %s"
        % (mixed["t_start"], mixed["t_done"], [round(o["t_done"], 3) for o in others]))
    same(alone, mixed)
    assert min(o["t_start"] for o in others) < mixed["t_done"]  # they overlapped with the big one
