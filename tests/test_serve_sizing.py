"""Device-memory sizing of a problem description (smcf_problem_bytes), the admission measure of the serving
loop.  No device work: runs without a GPU."""
import numpy as np

import oracle as O
import smc_inversion_fort_amd as S
from smc_inversion_fort_amd import synthetic as syn


def _desc(precision=S.PRECISION_FP64, nxi=4, neta=3, nsta=12, P=96):
    xmin, xmax, zmin, _ = syn.fault_frame()

    def greens(th, pts, uv, nx, ne):
        return O.Fault([nx], [ne], xmin, xmax, zmin).greens(th, pts, uv, 0, pts.shape[1] // 3)
    obs = syn.make_observations("c2", 3, greens, nsta=nsta, nxi=nxi, neta=neta)
    return S.problem_desc([nxi], [neta], xmin, xmax, zmin, obs, syn.MAX_SLIP, P, mover=S.MOVER_GIBBS,
                          gibbs_sweeps=3, precision=precision)


def test_problem_bytes_grows_with_slots_and_sizes():
    d, _k = _desc()
    b = [S.problem_bytes(d, n) for n in (1, 2, 4, 8)]
    assert all(x > 0 for x in b)
    assert b[0] < b[1] < b[2] < b[3]
    # slots are the same size: the increments are linear in the slot count (up to 256-byte rounding)
    per = [(b[i + 1] - b[i]) / n for i, n in ((0, 1), (1, 2), (2, 4))]
    assert max(per) - min(per) <= 256 * 64
    # the split-BF16 form keeps extra copies of the particles and of H
    dbf, _k2 = _desc(S.PRECISION_MIXED_BF16)
    assert S.problem_bytes(dbf, 4) > b[2]
    # more particles, more bytes
    dp, _k3 = _desc(P=192)
    assert S.problem_bytes(dp, 4) > b[2]


def test_problem_bytes_rejects_invalid_descriptions():
    """The sizing pass applies problem_create's checks (no device needed): the FP32 form needs the Gibbs mover (HMC
    movers take FP64 or split-BF16), unknown precisions and oversized particle counts are refused with the C ABI's
    error codes."""
    import pytest
    for kw, code in ((dict(precision=S.PRECISION_MIXED), 4), (dict(precision=7), 1)):
        xmin, xmax, zmin, _ = syn.fault_frame()

        def greens(th, pts, uv, nx, ne):
            return O.Fault([nx], [ne], xmin, xmax, zmin).greens(th, pts, uv, 0, pts.shape[1] // 3)
        obs = syn.make_observations("c2", 3, greens, nsta=12, nxi=4, neta=3)
        d, _k = S.problem_desc([4], [3], xmin, xmax, zmin, obs, syn.MAX_SLIP, 96, mover=S.MOVER_HMC, gibbs_sweeps=3,
                               **kw)
        with pytest.raises(S.SmcfError, match="error %d" % code):
            S.problem_bytes(d, 4)
    d, _k = _desc(P=100000)
    with pytest.raises(S.SmcfError, match="nparticle_slip"):
        S.problem_bytes(d, 4)


def test_problem_bytes_hmc_split_bf16_counts_its_buffers():
    """An HMC problem in the split-BF16 form holds the split Q_g and two split position copies beside the FP64
    buffers: its sizing is larger than the FP64 form's by about ndim^2 x 4 + 2 x 32-row-padded ndim x P x 4 bytes per
    slot."""
    xmin, xmax, zmin, _ = syn.fault_frame()

    def greens(th, pts, uv, nx, ne):
        return O.Fault([nx], [ne], xmin, xmax, zmin).greens(th, pts, uv, 0, pts.shape[1] // 3)
    obs = syn.make_observations("c2", 3, greens, nsta=12, nxi=4, neta=3)
    sizes = {}
    for prec in (S.PRECISION_FP64, S.PRECISION_MIXED_BF16):
        d, _k = S.problem_desc([4], [3], xmin, xmax, zmin, obs, syn.MAX_SLIP, 96, mover=S.MOVER_HMC_ACCEPT,
                               precision=prec)
        sizes[prec] = S.problem_bytes(d, 4)
    # 4x3 patches: 5x4 nodes, DOFs without the bottom row and the two side columns: 3 x 3 = 9, ndim 18; 4 slots
    nd, P, C = 18, 96, 4
    qgx = ((nd + 15) // 16) * ((nd + 31) // 32) * 64 * 2 * 16 * C    # split Q_g fragments, 16 bytes each
    xx = ((nd + 31) // 32) * 32 // 8 * P * 2 * 16 * C                  # one split position copy
    r256 = lambda n: (n + 255) // 256 * 256
    assert sizes[S.PRECISION_MIXED_BF16] - sizes[S.PRECISION_FP64] == r256(qgx) + 2 * r256(xx)

