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
