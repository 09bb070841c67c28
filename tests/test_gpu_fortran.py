"""The Fortran drop-in: a caller written against the reference's module procedures
(gfunc::calc_greens_func, smc_fault::fault_calc_likelihood) linked to the libsmcf shim."""
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np
import pytest

import smc_inversion_fort_amd as S

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "smc_inversion_fort_amd", "_build", "drop_in")

CASES = {
    "p1_3x2": (1, [3], [2], 0, -10.0, 10.0, -14.142, [], [-17.3, 17.1, 45.0, -1.0, 0.2, 0.1]),
    "p1_8x8": (1, [8], [8], 0, -10.0, 10.0, -14.142, [], [-17.32, 17.32, 45.0, -2.0, 0.0, 0.0]),
    "p2_fixed": (2, [2, 3], [3, 2], 1, -8.0, 9.0, -11.0, [0.5], [-12.0, 13.0, 1.5, 55.0, 42.0, 0.0, 0.0, 0.0]),
}


def run_drop_in(name, golden, dvec, sigma, P, seed):
    nplane, nxi, neta, fix, xmin, xmax, zmin, xbend, theta = CASES[name]
    g = golden("greens")
    pts, uv, nsi = g[name + "__obs_points"], g[name + "__obs_unitvec"], g[name + "__nsar_index"]
    nobs = pts.shape[1]
    b = struct.pack("<2i", nplane, fix) + np.array(nxi, "<i4").tobytes() + np.array(neta, "<i4").tobytes()
    b += np.array([xmin, xmax, zmin], "<f8").tobytes()
    if nplane > 1:
        b += np.array(xbend, "<f8").tobytes()
    b += np.array(theta, "<f8").tobytes()
    b += struct.pack("<4i", nobs, int(nsi[-1] - 1), int(g[name + "__ngnss"]), len(nsi) - 1)
    b += np.asarray(nsi, "<i4").tobytes() + pts.astype("<f8").tobytes(order="F") + uv.astype("<f8").tobytes(order="F")
    b += np.asarray(sigma, "<f8").tobytes() + np.asarray(dvec, "<f8").tobytes()
    b += struct.pack("<i", P) + struct.pack("<d", 10.0) + struct.pack("<q", seed)
    with tempfile.TemporaryDirectory() as td:
        fi, fo = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        open(fi, "wb").write(b)
        r = subprocess.run([EXE, fi, fo], capture_output=True, text=True, timeout=240)
        assert r.returncode == 0, r.stdout + r.stderr
        raw = open(fo, "rb").read()
    G = np.frombuffer(raw[:-8], "<f8").reshape((nobs, -1), order="F")
    return G, struct.unpack("<d", raw[-8:])[0]


@pytest.mark.parametrize("name", list(CASES))
def test_fortran_calc_greens_func_golden(golden, name):
    print("case: fortran drop-in greens", name, flush=True)
    g = golden("greens")
    nobs = g[name + "__obs_points"].shape[1]
    G, _ = run_drop_in(name, golden, np.zeros(nobs), np.ones(nobs), 16, 1)
    Gr = g[name + "__gmat"]
    assert np.abs(G - Gr).max() <= 1e-10 * np.abs(Gr).max()


def test_fortran_fault_calc_likelihood_matches_c_abi(golden):
    print("case: fortran drop-in fault_calc_likelihood", flush=True)
    name = "p1_8x8"
    g = golden("greens")
    nobs = g[name + "__obs_points"].shape[1]
    rng = np.random.default_rng(0)
    Gr = g[name + "__gmat"]
    s = np.exp(-np.linspace(-2, 2, Gr.shape[1]) ** 2)
    dvec = Gr @ s + 0.5 * rng.standard_normal(nobs)
    sigma = np.full(nobs, 0.5)
    P, seed = 64, 12345
    _, neglog = run_drop_in(name, golden, dvec, sigma, P, seed)
    nplane, nxi, neta, fix, xmin, xmax, zmin, xbend, theta = CASES[name]
    ctx = S.Context(0)
    obs = dict(obs_points=g[name + "__obs_points"], obs_unitvec=g[name + "__obs_unitvec"], obs_sigma=sigma,
               dvec=dvec, nsar_total=0, ngnss=int(g[name + "__ngnss"]))
    # the drop-in's default mover is the reference's (HMC, ntau 50, ESJD tuning)
    pb = S.Problem(ctx, nxi, neta, xmin, xmax, zmin, obs, 10.0, P, mover=S.MOVER_HMC, max_batch=64)
    ref, _ = pb.fault_calc_likelihood(theta, [S.inner_key(seed, 0, 2, 0)])
    assert neglog == ref[0]
    pb.close()
    ctx.close()
