"""InSAR observation covariance (calc_sigma_sar_mat, init.f90:82-119) on the host.

The device takes the inverse covariance blocks as input (smcf_problem_desc.sar_sinv), as the reference
takes sigma_sar_mat from main.f90; this helper computes them for Python callers with LAPACK (numpy).
"""
import numpy as np


def calc_sigma_sar_mat(nsar_index, obs_points, obs_sigma):
    """Inverse covariance of every InSAR path, column-major blocks concatenated in path order.

    Path p holds rows nsar_index[p] .. nsar_index[p+1]-1 (1-based).  C(i, j) = s_i s_j exp(-dist_ij / 10)
    with s_i = obs_sigma(nsar_index(p) + i), i = 1..n: the reference's sigma index is one row past the
    observation's own (init.f90:107-108), which this keeps, so the last row of a path takes the next row's
    sigma.  A path whose last index lies past nobs (the reference reads past the array) raises ValueError.
    """
    nsi = np.asarray(nsar_index, dtype=np.int64)
    pts = np.asarray(obs_points, dtype=np.float64)
    sig = np.asarray(obs_sigma, dtype=np.float64)
    blocks = []
    for p in range(len(nsi) - 1):
        a, n = int(nsi[p]) - 1, int(nsi[p + 1] - nsi[p])
        if n <= 0:
            continue
        if a + n >= sig.shape[0]:
            raise ValueError("InSAR path %d: sigma index %d past nobs %d" % (p + 1, a + n + 1, sig.shape[0]))
        xy = pts[:, a:a + n]
        dist = np.sqrt((xy[0][:, None] - xy[0][None, :]) ** 2 + (xy[1][:, None] - xy[1][None, :]) ** 2)
        s = sig[a + 1:a + n + 1]
        c = (s[:, None] * s[None, :]) * np.exp(-dist / 1e1)
        blocks.append(np.linalg.inv(c).ravel(order="F"))
    return np.concatenate(blocks) if blocks else np.zeros(1)
