"""Seeded synthetic inversion problems (observation sets, truth, priors).

The reference ships no data, so every workload is generated here from a seed:
GNSS 3-component stations uniform in a square box centred on the fault-trace
midpoint, a surface-breaking rectangular fault (strike 30 deg, length 40 km,
dip 45 deg, width 20 km), a Gaussian slip bump (peak 3 m) on the mesh nodes,
Gaussian noise.  Units follow the reference: km, m of slip, cm of displacement
(gfunc.f90:97-99).  Data are d = G(theta_true) s_true + noise, with G supplied
by the caller (device or any other implementation of calc_greens_func).
"""
import math

import numpy as np

# configuration sizes (patch grid, stations, noise sigma [cm], outer N, inner P, mover)
CONFIGS = {
    "c1": dict(nxi=8, neta=8, nsta=100, sigma=0.5, N=1000, P=1000, mover="hmc", sweeps=5),
    "c2": dict(nxi=16, neta=16, nsta=400, sigma=0.5, N=10000, P=2000, mover="gibbs", sweeps=5),
    "c3": dict(nxi=32, neta=16, nsta=1000, sigma=0.3, N=20000, P=2000, mover="hmc10", sweeps=5),
}

STRIKE_DEG = 30.0
LENGTH_KM = 40.0
DIP_DEG = 45.0
WIDTH_KM = 20.0
MAX_SLIP = 10.0
PEAK_SLIP = 3.0
BOX_KM = 100.0
# hyper-parameter prior ranges fixed by this project (the reference reads them from range.dat)
LOG_ALPHA2_RANGE = (-6.0, 0.0)
LOG_SIGMA_SAR2_RANGE = (-1.0, 1.0)
LOG_SIGMA_GNSS2_RANGE = (-1.0, 1.0)
LOG_ALPHA2_TRUE = -3.0


def fault_frame(strike=STRIKE_DEG, dip=DIP_DEG, length=LENGTH_KM, width=WIDTH_KM):
    """xmin, xmax, zmin and the true (ymin, ymax, dip) of a surface-breaking fault centred on the origin
    (defaults: the fault of the single-inversion configurations)."""
    s = math.radians(strike)
    dx, dy = length * math.sin(s), length * math.cos(s)
    xmin, xmax = -dx / 2, dx / 2
    ymin, ymax = -dy / 2, dy / 2
    zmin = -width * math.sin(math.radians(dip))
    return xmin, xmax, zmin, (ymin, ymax, dip)


def theta_true(frame=None):
    _, _, _, (ymin, ymax, dip) = frame or fault_frame()
    return np.array([ymin, ymax, dip, LOG_ALPHA2_TRUE, 0.0, 0.0])


def prior_range(frac=0.2, frame=None):
    """2 x 6 range (range.dat layout): +-frac of the true geometry, fixed hyper-parameter ranges."""
    _, _, _, (ymin, ymax, dip) = frame or fault_frame()
    r = np.zeros((2, 6))
    for k, v in enumerate((ymin, ymax, dip)):
        lo, hi = v * (1 - frac), v * (1 + frac)
        r[:, k] = (min(lo, hi), max(lo, hi))
    r[1, 2] = min(r[1, 2], 89.0)  # a steep fault's dip range stays below vertical
    r[:, 3] = LOG_ALPHA2_RANGE
    r[:, 4] = LOG_SIGMA_SAR2_RANGE
    r[:, 5] = LOG_SIGMA_GNSS2_RANGE
    return r


def stations(nsta, rng):
    xy = rng.uniform(-BOX_KM / 2, BOX_KM / 2, size=(2, nsta))
    nobs = 3 * nsta
    pts = np.zeros((2, nobs), order="F")
    uv = np.zeros((3, nobs), order="F")
    for i in range(nsta):
        for c in range(3):
            pts[:, 3 * i + c] = xy[:, i]
            uv[c, 3 * i + c] = 1.0
    return pts, uv


def truth_slip(nxi, neta, theta, frame=None):
    """Gaussian bump on the DOF nodes (both slip components >= 0), DOF order of discretize_fault."""
    xmin, xmax, zmin, _ = frame or fault_frame()
    ymin, ymax, dip = theta[0], theta[1], theta[2]
    lxi = math.hypot(xmax - xmin, ymax - ymin)
    leta = -zmin / math.sin(math.radians(dip))
    dxi, deta = lxi / nxi, leta / neta
    vals = []
    for j in range(1, neta + 2):
        for i in range(1, nxi + 2):
            if j == 1 or i == 1 or i == nxi + 1:
                continue
            xi = (i - 1) * dxi - lxi / 2
            eta = (j - 1) * deta - leta / 2
            a = PEAK_SLIP * math.exp(-(xi ** 2) / (2 * (lxi / 5) ** 2) - (eta ** 2) / (2 * (leta / 4) ** 2))
            vals += [a, a]
    return np.array(vals)


# line-of-sight unit vectors (east, north, up) of two InSAR tracks, ascending and descending
SAR_LOS = ((-0.62, -0.11, 0.78), (0.62, -0.11, 0.78))
SAR_SIGMA = 1.0  # cm


def sar_points(npath, nper, rng):
    """npath InSAR paths of nper pixels each, uniform in the box; rows path after path (SAR rows first)."""
    n = npath * nper
    pts = np.zeros((2, n), order="F")
    uv = np.zeros((3, n), order="F")
    pts[:] = rng.uniform(-BOX_KM / 2, BOX_KM / 2, size=(2, n))
    for p in range(npath):
        los = np.array(SAR_LOS[p % len(SAR_LOS)])
        uv[:, p * nper:(p + 1) * nper] = (los / np.linalg.norm(los))[:, None]
    return pts, uv


def make_observations(cfg_name, seed, greens, nsta=None, sigma=None, nxi=None, neta=None, npath=0, nsar_per_path=0,
                      frame=None):
    """Build the observation dict of config `cfg_name` (sizes overridable for small tests).

    greens(theta, obs_points, obs_unitvec, nxi, neta) -> G (nobs x ndim) at the true geometry; frame (fault_frame's
    output) sets another fault than the configurations' one.  npath > 0 adds npath InSAR paths of nsar_per_path pixels before the GNSS rows (the reference's row
    order), with noise correlated as the path covariance of calc_sigma_sar_mat, and stores the inverse
    covariance blocks under "sar_sinv"; greens is then called as greens(theta, pts, uv, nxi, neta, nsar).
    """
    if npath:
        return _make_observations_sar(cfg_name, seed, greens, nsta, sigma, nxi, neta, npath, nsar_per_path)
    cfg = dict(CONFIGS[cfg_name])
    nsta = nsta or cfg["nsta"]
    sigma = sigma if sigma is not None else cfg["sigma"]
    nxi = nxi or cfg["nxi"]
    neta = neta or cfg["neta"]
    rng = np.random.default_rng(np.random.SeedSequence([seed, 0x5A17]))
    pts, uv = stations(nsta, rng)
    th = theta_true(frame)
    G = greens(th, pts, uv, nxi, neta)
    s = truth_slip(nxi, neta, th, frame)
    d = G @ s + sigma * rng.standard_normal(G.shape[0])
    return dict(obs_points=pts, obs_unitvec=uv, obs_sigma=np.full(G.shape[0], sigma), dvec=d,
                nsar_total=0, ngnss=nsta, nsar_index=np.array([1, 1], np.int32), nxi=nxi, neta=neta,
                slip_true=s)


def _make_observations_sar(cfg_name, seed, greens, nsta, sigma, nxi, neta, npath, nper):
    from .sar import calc_sigma_sar_mat
    cfg = dict(CONFIGS[cfg_name])
    nsta = nsta or cfg["nsta"]
    sigma = sigma if sigma is not None else cfg["sigma"]
    nxi = nxi or cfg["nxi"]
    neta = neta or cfg["neta"]
    rng = np.random.default_rng(np.random.SeedSequence([seed, 0x5A17, 0x5A2]))
    spts, suv = sar_points(npath, nper, rng)
    gpts, guv = stations(nsta, rng)
    pts = np.asfortranarray(np.concatenate([spts, gpts], axis=1))
    uv = np.asfortranarray(np.concatenate([suv, guv], axis=1))
    nsar = npath * nper
    nsar_index = np.array([1 + p * nper for p in range(npath + 1)], np.int32)
    obs_sigma = np.concatenate([np.full(nsar, SAR_SIGMA), np.full(3 * nsta, sigma)])
    th = theta_true()
    G = greens(th, pts, uv, nxi, neta, nsar)
    s = truth_slip(nxi, neta, th)
    noise = np.concatenate([np.zeros(nsar), sigma * rng.standard_normal(3 * nsta)])
    sinv = calc_sigma_sar_mat(nsar_index, pts, obs_sigma)
    off = 0
    for p in range(npath):  # correlated path noise with the covariance the likelihood uses
        c = np.linalg.inv(sinv[off:off + nper * nper].reshape(nper, nper, order="F"))
        noise[p * nper:(p + 1) * nper] = np.linalg.cholesky((c + c.T) / 2) @ rng.standard_normal(nper)
        off += nper * nper
    d = G @ s + noise
    return dict(obs_points=pts, obs_unitvec=uv, obs_sigma=obs_sigma, dvec=d, nsar_total=nsar, ngnss=nsta,
                nsar_index=nsar_index, sar_sinv=sinv, nxi=nxi, neta=neta, slip_true=s)


def sample_thetas(B, seed, frac=0.2, frame=None):
    """B geometry particles drawn uniformly from the prior box (fault_sample_init_particles)."""
    rng = np.random.default_rng(np.random.SeedSequence([seed, 0x7E7A]))
    r = prior_range(frac, frame)
    return np.asfortranarray(r[0][:, None] + rng.uniform(size=(6, B)) * (r[1] - r[0])[:, None])


# The batching configuration: independent inversions whose sizes and faults are drawn from the seed
BATCH = dict(n=64, patches=(4, 24), nsta=(50, 800), strike=(0.0, 360.0), dip=(20.0, 80.0), length=(10.0, 60.0),
             N=(1000, 8000), P=(500, 2000), sigma=(0.2, 1.0))


def batch_instances(n=None, seed=0, **over):
    """n independent inversions of the batching configuration: per instance a patch grid (nxi, neta uniform in
    4..24), a station count (50..800), the true fault (strike 0..360 deg, dip 20..80 deg, length 10..60 km, width
    half the length, surface-breaking, centred on the origin), the noise (2..10 mm), outer N (1000..8000) and inner
    P (500..2000) particle counts.  Keyword arguments replace ranges of BATCH (e.g. patches=(4, 8))."""
    cfg = dict(BATCH, **over)
    n = n or cfg["n"]
    rng = np.random.default_rng(np.random.SeedSequence([seed, 0xBA7C]))
    out = []
    for k in range(n):
        lo, hi = cfg["patches"]
        nxi, neta = (int(v) for v in rng.integers(lo, hi + 1, size=2))
        nsta = int(rng.integers(cfg["nsta"][0], cfg["nsta"][1] + 1))
        strike = float(rng.uniform(*cfg["strike"]))
        dip = float(rng.uniform(*cfg["dip"]))
        length = float(rng.uniform(*cfg["length"]))
        sigma = float(rng.uniform(*cfg["sigma"]))
        N = int(rng.integers(cfg["N"][0], cfg["N"][1] + 1))
        P = int(rng.integers(cfg["P"][0], cfg["P"][1] + 1))
        out.append(dict(k=k, seed=int(seed) * 1000 + k, nxi=nxi, neta=neta, nsta=nsta, sigma=sigma, N=N, P=P,
                        frame=fault_frame(strike, dip, length, length / 2), strike=strike, dip=dip, length=length))
    return out
