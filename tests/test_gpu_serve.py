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
    srv = S.Server(0, hbm_budget_bytes=sum(sizes) + (1 << 20))
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
    srv = S.Server(0, hbm_budget_bytes=max(b0, b1) + (b0 + b1) // 4)
    i0 = srv.submit(d0, rng, N, 11, max_stages=STAGES, max_batch=N)
    i1 = srv.submit(d1, rng, N, 12, max_stages=STAGES, max_batch=N)
    r0, r1 = srv.wait(i0), srv.wait(i1)
    say("first done at %.3f, second admitted at %.3f" % (r0["t_done"], r1["t_start"]))
    assert r1["t_start"] >= r0["t_done"]
    ref = S.Server(0, hbm_budget_bytes=b1 + (1 << 20))
    r1b = ref.wait(ref.submit(d1, rng, N, 12, max_stages=STAGES, max_batch=N))
    ref.close()
    srv.close()
    same(r1, r1b)
