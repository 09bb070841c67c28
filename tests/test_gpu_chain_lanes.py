"""The prior Gibbs chain of slip_gen_init_particles (smc_slip.f90:184-287) at every lane count the library picks.

k_init_chain pipelines one particle chain over L lanes per slip component (L = 2 .. 32, chosen from the mesh's DOF
count and coupling band) and packs 32 / L chains of one geometry into a wavefront.  On meshes of 8x8, 16x16,
4x40 and 3x80 patches (4, 8, 16 and 32 lanes per component; 4x3 patches, 2 lanes, are covered with the other
components) this checks, against the oracle's sequential chain on the same keyed normals:

* the stand-alone entry point (one chain per wavefront) to 1e-10, as the component check of the chain: the
  two sum in different orders and the chain carries their rounding from particle to particle (up to 2e-12
  measured), while a lane reading another lane's ring moves slips by a good part of their size;
* a batch of 32 / L + 1 geometries admitted together (a full wavefront of packed chains and one more) against
  each geometry run alone, bit for bit, so the packed chains of lanes beyond the first read and write their own
  rings."""
import sys

import numpy as np
import pytest

import oracle as O
import smc_inversion_fort_amd as S
from smc_inversion_fort_amd import synthetic as syn

pytestmark = pytest.mark.gpu

# (nxi, neta) -> lanes per component the library picks for them
MESHES = [((8, 8), 4), ((16, 16), 8), ((4, 40), 16), ((3, 80), 32)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = S.Context(0)
    yield c
    c.close()


def build(ctx, nxi, neta, P, max_batch):
    xmin, xmax, zmin, _ = syn.fault_frame()

    def greens(th, pts, uv, nx, ne):
        return O.Fault([nx], [ne], xmin, xmax, zmin).greens(th, pts, uv, 0, pts.shape[1] // 3)

    obs = syn.make_observations("c2", 11, greens, nsta=12, sigma=0.5, nxi=nxi, neta=neta)
    pb = S.Problem(ctx, [nxi], [neta], xmin, xmax, zmin, obs, syn.MAX_SLIP, P, mover=S.MOVER_GIBBS,
                   gibbs_sweeps=2, max_batch=max_batch)
    return pb, O.Fault([nxi], [neta], xmin, xmax, zmin)


@pytest.mark.parametrize("mesh,lanes", MESHES, ids=["%dx%d" % m for m, _ in MESHES])
def test_chain_vs_oracle(ctx, mesh, lanes):
    nxi, neta = mesh
    P = 2 * lanes + 7  # every lane runs at least two particles, the ring wraps
    say("case: prior chain %dx%d patches (%d lanes), P=%d, stand-alone entry point vs oracle" % (nxi, neta, lanes, P))
    pb, f = build(ctx, nxi, neta, P, 1)
    try:
        th = syn.theta_true() + np.array([0.3, -0.2, 1.1, 0.2, 0.1, -0.1])
        key = O.inner_key(17, 0, 0, lanes)
        X, _, _ = pb.slip_gen_init_particles(th, key)
        nd = f.ndim
        bm = np.zeros((nd, P), order="F")
        for k in range(P):
            for i in range(nd):
                bm[i, k] = O.draw_normal(key, 1, 0, k, i)
        lap = f.laplacian(th)
        alpha2 = np.full(2 * f.nnode, np.exp(th[3]))
        Xo = O.gibbs_init_chain(P, lap["ll"], alpha2, syn.MAX_SLIP, bm)
        err = float(np.abs(X - Xo).max())
        say("  ndim %d, max |device - oracle| = %.3g, largest slip %.3g" % (nd, err, float(Xo.max())))
        assert err <= 1e-10
        assert np.all(X >= 0) and np.all(X <= syn.MAX_SLIP)
    finally:
        pb.close()


@pytest.mark.parametrize("mesh,lanes", MESHES, ids=["%dx%d" % m for m, _ in MESHES])
def test_packed_chains_match_alone(ctx, mesh, lanes):
    nxi, neta = mesh
    n = 32 // lanes + 1
    say("case: packed prior chains %dx%d patches (%d lanes), %d geometries together vs alone" % (nxi, neta, lanes, n))
    pb, _ = build(ctx, nxi, neta, 32, n)
    try:
        th = syn.sample_thetas(n, 40 + lanes)
        keys = [O.inner_key(23, 0, 0, i) for i in range(n)]
        nl, ns = pb.fault_calc_likelihood(th, keys)
        say("  together: -log Z", nl.tolist(), "stages", ns.tolist())
        for k in range(n):
            a, sa = pb.fault_calc_likelihood(th[:, k:k + 1], keys[k:k + 1])
            assert a[0] == nl[k] and sa[0] == ns[k], "geometry %d: alone %r (%d stages)" % (k, a[0], sa[0])
    finally:
        pb.close()
