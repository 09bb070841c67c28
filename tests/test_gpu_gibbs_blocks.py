"""The FP64 inner SMC (slip_smc_exec, smc_slip.f90:1445-1663) with the truncated-MVN Gibbs mover in place of the
reference's HMC moves (smc_slip.f90:591-730), against the oracle on meshes whose slip vector spans several
coordinate blocks of the sweep kernel.

k_gibbs processes coordinates in blocks (a matrix-core product over every coordinate for the block, then per
16-coordinate tile the left-looking corrections and draws), so a slip vector of one partial block, as on the
4x3-patch mesh of the other parity checks, leaves the hand-over between blocks and tiles unexercised.  Here:
8x8 patches (ndim 112), 12x11 (ndim 242: the last tile holds 2 coordinates) and the 16x16 mesh of the benchmark
(ndim 480, 5 sweeps per stage), there also with P = 600: a workgroup sweeps 256 particles of an instance, so
three workgroups, the last one partly filled.  Same stages, gammas to 1e-9, -log Z to 1e-8 relative, particles
to 1e-7, as the small-mesh check."""
import sys

import numpy as np
import pytest

import oracle as O
import smc_inversion_fort_amd as S
from smc_inversion_fort_amd import synthetic as syn

pytestmark = pytest.mark.gpu

CASES = [((8, 8), 128, 3, 12), ((12, 11), 128, 3, 20), ((16, 16), 128, 5, 30), ((16, 16), 600, 5, 30)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = S.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("mesh,P,sweeps,nsta", CASES, ids=["%dx%d-P%d" % (c[0] + (c[1],)) for c in CASES])
def test_gibbs_blocks_vs_oracle(ctx, mesh, P, sweeps, nsta):
    nxi, neta = mesh
    say("case: FP64 Gibbs inner SMC %dx%d patches, P=%d, %d sweeps, %d stations vs oracle" % (nxi, neta, P, sweeps,
                                                                                           nsta))
    xmin, xmax, zmin, _ = syn.fault_frame()

    def greens(th, pts, uv, nx, ne):
        return O.Fault([nx], [ne], xmin, xmax, zmin).greens(th, pts, uv, 0, pts.shape[1] // 3)

    obs = syn.make_observations("c2", 3, greens, nsta=nsta, sigma=0.5, nxi=nxi, neta=neta)
    pb = S.Problem(ctx, [nxi], [neta], xmin, xmax, zmin, obs, syn.MAX_SLIP, P, mover=S.MOVER_GIBBS,
                   gibbs_sweeps=sweeps, max_batch=1)
    try:
        f = O.Fault([nxi], [neta], xmin, xmax, zmin)
        po = O.ProblemHolder(f, obs, syn.MAX_SLIP, P, mover=1, gibbs_sweeps=sweeps)
        theta = syn.theta_true()
        key = O.inner_key(7, 1, 1, 3)
        d = pb.slip_smc_exec(theta, key)
        G = pb.calc_greens_func(theta)
        o = po.slip_smc(theta, G, key)
        say("  ndim %d, stages %d / %d, -log Z %.12g / %.12g, max |particle diff| %.3g, max |gamma diff| %.3g"
            % (f.ndim, d["nstage"], o["nstage"], d["neglog"], o["neglog"],
               float(np.abs(d["particles"] - o["particles"]).max()),
               float(np.abs(np.asarray(d["gamma"]) - np.asarray(o["gamma"])).max())))
        assert d["nstage"] == o["nstage"]
        np.testing.assert_allclose(d["gamma"], o["gamma"], rtol=0, atol=1e-9)
        assert d["neglog"] == pytest.approx(o["neglog"], rel=1e-8)
        np.testing.assert_allclose(d["particles"], o["particles"], rtol=0, atol=1e-7)
        np.testing.assert_allclose(d["likelihood"], o["likelihood"], rtol=1e-8)
        np.testing.assert_allclose(d["mean"], o["mean"], rtol=1e-8, atol=1e-10)
    finally:
        pb.close()


def test_batch_path_16x16_vs_oracle(ctx):
    """The batched entry point (fault_calc_likelihood: the stage engine with its item lists, packed prior chains and
    per-stage graph) at the benchmark's mesh and sweep count: each geometry's stage count and -log Z against the
    oracle's fault likelihood (its own Green's matrix and Laplacian), 4 geometries through 4 slots."""
    say("case: batched fault likelihood 16x16 patches, P=128, 5 sweeps, 4 geometries vs oracle")
    xmin, xmax, zmin, _ = syn.fault_frame()

    def greens(th, pts, uv, nx, ne):
        return O.Fault([nx], [ne], xmin, xmax, zmin).greens(th, pts, uv, 0, pts.shape[1] // 3)

    obs = syn.make_observations("c2", 4, greens, nsta=30, sigma=0.5, nxi=16, neta=16)
    P = 128
    pb = S.Problem(ctx, [16], [16], xmin, xmax, zmin, obs, syn.MAX_SLIP, P, mover=S.MOVER_GIBBS, gibbs_sweeps=5,
                   max_batch=4)
    try:
        po = O.ProblemHolder(O.Fault([16], [16], xmin, xmax, zmin), obs, syn.MAX_SLIP, P, mover=1, gibbs_sweeps=5)
        th = syn.sample_thetas(4, 77)
        keys = [O.inner_key(31, 0, 0, i) for i in range(4)]
        nl, ns = pb.fault_calc_likelihood(th, keys)
        for k in range(4):
            ref, nst = po.fault_likelihood(th[:, k], keys[k])
            say("  geometry %d: stages %d / %d, -log Z %.12g / %.12g" % (k, ns[k], nst, nl[k], ref))
            assert ns[k] == nst
            assert nl[k] == pytest.approx(ref, rel=1e-8)
    finally:
        pb.close()
