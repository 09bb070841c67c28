"""Write the golden fixtures of tests/golden/*.npz from the compiled reference.

Runs oracle/_ref/ref_driver (built by `make -C oracle ref` from the reference's own
Fortran sources) on seeded inputs and stores inputs + outputs.  The fixtures are
data only; this script and ref_driver.f90 are how they were made.  Needs the
reference tree, so it runs only where the reference is mounted:

    make -C oracle ref && python tests/golden/make_golden.py
"""
import os
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "..", "..", "oracle", "_ref", "ref_driver")


def run(mode, payload, env_threads=True):
    with tempfile.TemporaryDirectory() as td:
        fi, fo = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fi, "wb") as f:
            f.write(payload)
        env = dict(os.environ)
        if env_threads:
            env["OMP_NUM_THREADS"] = "1"
        subprocess.run([DRIVER, mode, fi, fo], check=True, env=env, stdout=subprocess.DEVNULL)
        with open(fo, "rb") as f:
            return f.read()


def ints(*v):
    return struct.pack("<%di" % len(v), *v)


def dbl(a):
    return np.asarray(a, dtype="<f8").tobytes(order="F")


def i32(a):
    return np.asarray(a, dtype="<i4").tobytes(order="F")


def gen_dc3d0(rng):
    cases = []
    for alpha in (2.0 / 3.0, 0.5):
        for dip in (30.0, 45.0, 60.0, 90.0, 89.9999999, 12.5):
            for depth in (0.7, 3.0, 11.0):
                for z in (0.0, -1.3):
                    for _ in range(3):
                        x, y = rng.uniform(-30, 30, size=2)
                        for pots in ((1, 0, 0, 0), (0, 1, 0, 0), (0, 0, 1, 0), (0, 0, 0, 1),
                                     (0.7, -1.9, 0.4, 1.3)):
                            cases.append([alpha, x, y, z, depth, dip] + list(pots))
    # snapping and special branches
    cases += [
        [2 / 3, 5e-7, 3.0, 0.0, 2.0, 45.0, 1, 1, 1, 1],      # |x| < EPS
        [2 / 3, 4.0, -3e-7, 0.0, 2.0, 45.0, 1, 1, 1, 1],     # |y| < EPS
        [2 / 3, 4.0, 2.0, -2.0 + 5e-7, 2.0, 60.0, 1, 1, 1, 1],  # |d| < EPS on the real source
        [2 / 3, 0.0, 0.0, -2.0, 2.0, 45.0, 1, 1, 1, 1],      # R = 0 -> IRET 1
        [2 / 3, 1.0, 1.0, 0.5, 2.0, 45.0, 1, 1, 1, 1],       # z > 0 -> IRET 2
        [2 / 3, 1.0, 1.0, 0.0, -1.0, 45.0, 1, 1, 1, 1],      # depth < 0 -> zeros, IRET untouched
        [2 / 3, 0.3, -0.2, 0.0, 0.05, 45.0, 1, 1, 0, 0],     # near field
        [2 / 3, 80.0, 95.0, 0.0, 15.0, 45.0, 1, 1, 0, 0],    # far field
    ]
    a = np.array(cases, dtype=np.float64)
    out = run("dc3d0", ints(len(a)) + a.tobytes())
    rec = np.frombuffer(out, dtype=np.dtype([("u", "<f8", 12), ("iret", "<i4")]))
    np.savez_compressed(os.path.join(HERE, "dc3d0.npz"), inputs=a, u=rec["u"].copy(),
                        iret=rec["iret"].copy())


FAULTS = [
    # name, nplane, nxi, neta, fix_xbend, xmin, xmax, zmin, xbend, theta
    ("p1_3x2", 1, [3], [2], 0, -10.0, 10.0, -14.142, [], [-17.3, 17.1, 45.0, -1.0, 0.2, 0.1]),
    ("p1_4x3", 1, [4], [3], 0, -9.0, 11.0, -12.0, [], [-15.0, 19.5, 38.0, 0.5, 0.0, -0.3]),
    ("p1_8x8", 1, [8], [8], 0, -10.0, 10.0, -14.142, [], [-17.32, 17.32, 45.0, -2.0, 0.0, 0.0]),
    ("p2_free", 2, [3, 2], [2, 2], 0, -10.0, 12.0, -10.0, [],
     [-14.0, 16.0, 1.0, 2.0, 50.0, 70.0, 0.1, 0.2, 0.3]),
    ("p2_fixed", 2, [2, 3], [3, 2], 1, -8.0, 9.0, -11.0, [0.5], [-12.0, 13.0, 1.5, 55.0, 42.0, 0.0, 0.0, 0.0]),
]


def fault_header(nplane, nxi, neta, fix, xmin, xmax, zmin, xbend, theta):
    b = ints(nplane, fix) + i32(nxi) + i32(neta) + dbl([xmin, xmax, zmin])
    if nplane > 1:
        b += dbl(xbend if xbend else [0.0] * (nplane - 1))
    return b + dbl(theta)


def gen_mesh():
    out = {}
    for name, nplane, nxi, neta, fix, xmin, xmax, zmin, xbend, theta in FAULTS:
        raw = run("mesh", fault_header(nplane, nxi, neta, fix, xmin, xmax, zmin, xbend, theta))
        nnode, ndof, npatch = struct.unpack_from("<3i", raw, 0)
        off = 12

        def take(dtype, shape):
            nonlocal off
            n = int(np.prod(shape))
            a = np.frombuffer(raw, dtype=dtype, count=n, offset=off).reshape(shape, order="F").copy()
            off += n * np.dtype(dtype).itemsize
            return a
        rec = dict(coor=take("<f8", (2, nnode)), cny=take("<i4", (4, npatch)), n2e=take("<i4", (4, nnode)),
                   n2es=take("<i4", (nnode,)), id_dof=take("<i4", (ndof,)),
                   luni=take("<f8", (nnode, nnode)), lmat=take("<f8", (2 * nnode, 2 * ndof)),
                   li=take("<i4", (9, 2 * nnode)), lv=take("<f8", (9, 2 * nnode)),
                   lti=take("<i4", (9, 2 * ndof)), ltv=take("<f8", (9, 2 * ndof)))
        for k, v in rec.items():
            out[name + "__" + k] = v
    np.savez_compressed(os.path.join(HERE, "mesh.npz"), **out)


def obs_set(rng, nsar_ls, ngnss):
    nsar = sum(nsar_ls)
    nobs = nsar + 3 * ngnss
    pts = np.zeros((2, nobs))
    uv = np.zeros((3, nobs))
    pts[:, :nsar] = rng.uniform(-40, 40, size=(2, nsar))
    los = rng.normal(size=(3, nsar))
    uv[:, :nsar] = los / np.linalg.norm(los, axis=0)
    st = rng.uniform(-45, 45, size=(2, ngnss))
    st[:, 0] = [0.0, 0.0]          # on the trace midpoint
    for i in range(ngnss):
        for c in range(3):
            pts[:, nsar + 3 * i + c] = st[:, i]
            uv[c, nsar + 3 * i + c] = 1.0
    nsar_index = np.cumsum([1] + list(nsar_ls)).astype(np.int32)
    return pts, uv, nsar_index


def gen_greens(rng):
    out = {}
    for (name, nplane, nxi, neta, fix, xmin, xmax, zmin, xbend, theta), (nsar_ls, ngnss) in zip(
            FAULTS, [([3, 2], 6), ([4], 7), ([0], 20), ([2], 5), ([0], 6)]):
        pts, uv, nsi = obs_set(rng, nsar_ls, ngnss)
        nobs = pts.shape[1]
        raw = run("greens", fault_header(nplane, nxi, neta, fix, xmin, xmax, zmin, xbend, theta)
                  + ints(nobs, int(sum(nsar_ls)), ngnss, len(nsar_ls)) + i32(nsi) + dbl(pts) + dbl(uv))
        nobs_r, ndof = struct.unpack_from("<2i", raw, 0)
        g = np.frombuffer(raw, "<f8", offset=8).reshape((nobs_r, 2 * ndof), order="F").copy()
        out[name + "__obs_points"] = pts
        out[name + "__obs_unitvec"] = uv
        out[name + "__nsar_index"] = nsi
        out[name + "__ngnss"] = np.array(ngnss)
        out[name + "__gmat"] = g
    np.savez_compressed(os.path.join(HERE, "greens.npz"), **out)


def gen_like(rng):
    m = np.load(os.path.join(HERE, "mesh.npz"))
    li, lv = m["p1_4x3__li"], m["p1_4x3__lv"]
    nnode = li.shape[1] // 2
    ndim = m["p1_4x3__lti"].shape[1]
    P = 25
    A = rng.normal(size=(ndim, ndim))
    H = A @ A.T + ndim * np.eye(ndim)
    H = H + 1e-3 * rng.normal(size=H.shape)      # not exactly symmetric, like a BLAS product
    b = rng.normal(size=ndim) * 10
    s = rng.uniform(0, 5, size=(ndim, P))
    a2 = np.full(2 * nnode, np.exp(-0.7))
    dsd, lss2, lsg2, nsar, ngnss = 321.5, 0.3, -0.4, 7, 11
    raw = run("like", ints(ndim, nnode, P, nsar, ngnss) + dbl([dsd, lss2, lsg2]) + dbl(H) + dbl(b) + dbl(s)
              + i32(li) + dbl(lv) + dbl(a2))
    out = np.frombuffer(raw, "<f8").reshape((2, P), order="F")
    np.savez_compressed(os.path.join(HERE, "like.npz"), H=H, b=b, s=s, li=li, lv=lv, alpha2=a2,
                        scal=np.array([dsd, lss2, lsg2, nsar, ngnss]), likelihood=out[0].copy(),
                        prior=out[1].copy())


def gen_gamma(rng):
    out = {}
    cases = [(0.0, rng.uniform(10, 1e4, 100)), (0.0, rng.exponential(50, 300) + 1e3),
             (0.37, rng.normal(5e3, 200, 64)), (0.999999, rng.uniform(0, 1, 50)),
             (0.0, np.concatenate([rng.uniform(1e5, 1e7, 200), [3.0]]))]
    for i, (gp, L) in enumerate(cases):
        raw = run("gamma", ints(len(L)) + dbl([gp]) + dbl(L))
        v = np.frombuffer(raw, "<f8")
        out["c%d__gprev" % i] = np.array(gp)
        out["c%d__L" % i] = L
        out["c%d__gamma" % i] = v[0]
        out["c%d__neglog_ev" % i] = v[1]
        out["c%d__w" % i] = v[2:].copy()
    np.savez_compressed(os.path.join(HERE, "gamma.npz"), **out)


def gen_moments(rng):
    P, nd = 200, 10
    w = rng.exponential(1.0, P) * (rng.uniform(size=P) > 0.2)
    X = rng.normal(2.0, 0.04, size=(nd, P))
    X[3] = rng.normal(1.0, 1.0, P)
    raw = run("moments", ints(P, nd) + dbl(w) + dbl(X))
    v = np.frombuffer(raw, "<f8")
    np.savez_compressed(os.path.join(HERE, "moments.npz"), w_in=w, X=X, w=v[:P].copy(),
                        mean=v[P:P + nd].copy(), cov_diag=v[P + nd:].copy())


def gen_resample(rng):
    out = {}
    for i, (P, seed) in enumerate([(10, 3), (100, 17), (1000, 1234), (2000, 99), (37, 5)]):
        w = rng.exponential(1.0, P) * (rng.uniform(size=P) > 0.3)
        if i == 4:
            w = np.zeros(P)
            w[[2, 9, 30]] = [0.25, 0.5, 0.25]
        w = w / w.sum()
        raw = run("resample", ints(P, seed) + dbl(w))
        hi, lo = struct.unpack_from("<2d", raw, 0)
        a = np.frombuffer(raw, "<i4", offset=16).copy()
        out["c%d__w" % i] = w
        out["c%d__u" % i] = np.array([hi, lo])
        out["c%d__assigned" % i] = a
    np.savez_compressed(os.path.join(HERE, "resample.npz"), **out)


def gen_init(rng):
    m = np.load(os.path.join(HERE, "mesh.npz"))
    lmat = m["p1_4x3__lmat"]
    li, lv = m["p1_4x3__li"], m["p1_4x3__lv"]
    ll = lmat.T @ lmat
    nd = ll.shape[0]
    nnode = li.shape[1] // 2
    P, seed, max_slip = 30, 4242, 10.0
    a2 = np.full(2 * nnode, np.exp(-1.0))
    raw = run("init", ints(nd, P, nnode, seed) + dbl([max_slip]) + dbl(ll) + dbl(a2) + i32(li) + dbl(lv))
    v = np.frombuffer(raw, "<f8")
    nu = 2 * nd * P
    np.savez_compressed(os.path.join(HERE, "init.npz"), ll=ll, alpha2=a2, li=li, lv=lv,
                        max_slip=np.array(max_slip), uniforms=v[:nu].copy(),
                        particles=v[nu:nu + nd * P].reshape((nd, P), order="F").copy(),
                        prior=v[nu + nd * P:].copy())


def gen_sort(rng):
    import ctypes
    lib = ctypes.CDLL(os.path.join(HERE, "..", "..", "oracle", "_ref", "libref_sort.so"))
    out = {}
    for i, n in enumerate([10, 17, 100, 1000, 10000, 20000]):
        arr = rng.poisson(1.0, n).astype(np.int32)
        if i == 5:
            arr = rng.choice([0, 1, 2, 50], size=n, p=[0.5, 0.3, 0.19, 0.01]).astype(np.int32)
        idx = np.arange(1, n + 1, dtype=np.int32)
        nn = ctypes.c_int(n)
        lib.qsort_(arr.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                   idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), ctypes.byref(nn))
        out["c%d__arr" % i] = arr
        out["c%d__index" % i] = idx
    np.savez_compressed(os.path.join(HERE, "sort.npz"), **out)


if __name__ == "__main__":
    rng = np.random.default_rng(20261203)
    gen_dc3d0(rng)
    gen_mesh()
    gen_greens(rng)
    gen_like(rng)
    gen_gamma(rng)
    gen_moments(rng)
    gen_resample(rng)
    gen_init(rng)
    gen_sort(rng)
    print("golden fixtures written to", HERE)
