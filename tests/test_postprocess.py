"""Post-processing of main.f90 (calc_slip_map, point_cloud) restated on the host: mesh against the oracle's
discretize_fault (itself pinned to the reference's golden mesh), point clouds against closed-form properties of
bilinear interpolation and of the fault plane.  No device work."""
import math

import numpy as np
import pytest

import oracle as O
from smc_inversion_fort_amd import postprocess as PP
from smc_inversion_fort_amd import synthetic as syn

CASES = [  # (nxi_ls, neta_ls, fix_xbend, xbend, theta)
    ([3], [2], False, (), [-17.3, 17.1, 45.0, -1.0, 0.2, 0.1]),
    ([3, 2], [2, 2], False, (), [-14.0, 16.0, 1.0, 2.0, 50.0, 70.0, 0.1, 0.2, 0.3]),
    ([2, 3], [3, 2], True, (0.5,), [-12.0, 13.0, 1.5, 55.0, 42.0, 0.0, 0.0, 0.0]),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_discretize_fault_matches_oracle(case):
    nxi, neta, fx, xb, th = CASES[case]
    xmin, xmax, zmin = -10.0, 10.0, -12.0
    coor, cny = PP.discretize_fault(np.array(th), nxi, neta, xmin, xmax, zmin, fx, xb)
    m = O.Fault(nxi, neta, xmin, xmax, zmin, fx, xb).mesh(th)
    np.testing.assert_array_equal(cny, m["cny"])
    np.testing.assert_allclose(coor, m["coor"], rtol=0, atol=1e-12)


def test_point_cloud_geometry_and_interpolation():
    """Per patch the nr x nr points average to the patch centre (symmetric abscissae), a slip field linear in
    the node coordinates is reproduced exactly (bilinear interpolation), every point lies on the plane through
    the fault centre with the fault's strike and dip, and a surface-breaking fault reaches z = 0 at its top edge
    (the points' z spans (zmin, 0))."""
    xmin, xmax, zmin, _ = syn.fault_frame()
    nxi, neta = [4], [3]
    th = syn.theta_true()
    coor, cny = PP.discretize_fault(th, nxi, neta, xmin, xmax, zmin)
    a, b, c = 0.3, -0.2, 1.5
    nnode = coor.shape[1]
    slip = np.zeros((2, nnode))
    slip[0] = a * coor[0] + b * coor[1] + c
    slip[1] = 2.0 - 0.1 * coor[0]
    pts = PP.point_cloud(th[:, None], slip.T.reshape(-1)[:, None], nxi, neta, xmin, xmax, zmin)[0]
    nr = 8
    assert pts.shape == (nxi[0] * neta[0] * nr * nr, 6)
    xf, yf, zf, s, d = PP.plane_frame(th, 0, nxi, neta, xmin, xmax, zmin)
    normal = np.array([math.cos(s) * math.sin(d), -math.sin(s) * math.sin(d), math.cos(d)])
    dist = (pts[:, :3] - np.array([xf, yf, zf])) @ normal
    assert np.abs(dist).max() < 1e-9
    assert pts[:, 2].min() > zmin and pts[:, 2].max() < 0.0
    for ie in range(cny.shape[1]):
        blk = pts[ie * nr * nr:(ie + 1) * nr * nr]
        nodes = cny[:, ie] - 1
        xi_c, eta_c = coor[0, nodes].mean(), coor[1, nodes].mean()
        ctr = np.array([xf - eta_c * math.cos(d) * math.cos(s) + xi_c * math.sin(s),
                        yf + eta_c * math.cos(d) * math.sin(s) + xi_c * math.cos(s), zf + eta_c * math.sin(d)])
        np.testing.assert_allclose(blk[:, :3].mean(axis=0), ctr, atol=1e-9)
        # recover (xi, eta) of each point from its position and check the interpolated slip
        rel = blk[:, :3] - np.array([xf, yf, zf])
        xi = rel[:, 0] * math.sin(s) + rel[:, 1] * math.cos(s)
        eta = rel[:, 2] / math.sin(d)
        np.testing.assert_allclose(blk[:, 3], a * xi + b * eta + c, atol=1e-9)
        np.testing.assert_allclose(blk[:, 4], 2.0 - 0.1 * xi, atol=1e-9)
        np.testing.assert_allclose(blk[:, 5], np.hypot(blk[:, 3], blk[:, 4]), atol=1e-12)


def test_point_cloud_stale_layout_reads_eight_values_per_plane():
    """layout="stale8" takes each plane's centre, strike and dip (degrees) from theta(8(p-1)+1 .. 8(p-1)+5), as
    the reference does, while the mesh still comes from discretize_fault reading the same vector in the current
    layout: every point is centre + xi u(strike) + eta v(strike, dip) for its interpolated (xi, eta)."""
    xmin, xmax, zmin, _ = syn.fault_frame()
    nxi, neta = [2], [2]
    th8 = np.array([-15.0, 16.0, 40.0, 30.0, 50.0, 0.0, 0.0, 0.0])  # as (ymin, ymax, dip, ...) for the mesh too
    xf, yf, zf, strike, dip = th8[:5]
    s, d = math.radians(strike), math.radians(dip)
    slip = np.linspace(0.0, 1.0, 18)[:, None]
    pts = PP.point_cloud(th8[:, None], slip, nxi, neta, xmin, xmax, zmin, layout="stale8")[0]
    coor, cny = PP.discretize_fault(th8, nxi, neta, xmin, xmax, zmin)
    r = PP.cloud_offsets(8)
    k = 0
    for ie in range(cny.shape[1]):
        nodes = cny[:, ie] - 1
        for r1 in r:
            for r2 in r:
                nv = np.array([(1 - r1) * (1 - r2), (1 + r1) * (1 - r2), (1 + r1) * (1 + r2), (1 - r1) * (1 + r2)]) / 4
                xi, eta = nv @ coor[0, nodes], nv @ coor[1, nodes]
                exp = [xf - eta * math.cos(d) * math.cos(s) + xi * math.sin(s),
                       yf + eta * math.cos(d) * math.sin(s) + xi * math.cos(s), zf + eta * math.sin(d)]
                np.testing.assert_allclose(pts[k, :3], exp, atol=1e-12)
                k += 1


def test_slip_map_records_round_trip(tmp_path):
    """calc_slip_map's records (2 nnode F12.5 values, node-major) read back as point_cloud reads them."""
    nxi, neta = [3], [2]
    nd = 2 * int(sum(1 for _ in PP.D.id_dof(nxi, neta)))
    means = np.random.default_rng(0).uniform(0, 5, size=(nd, 4))
    path = tmp_path / "mapslip.dat"
    with open(path, "w") as f:
        for i in range(4):
            sl = PP.node_slip(means[:, i], nxi, neta)
            f.write("".join(PP.D.fortran_f(sl[k, m], 12, 5) for m in range(sl.shape[1]) for k in range(2)) + "\n")
    back = PP.read_slip_map(str(path), 4, 12)
    for i in range(4):
        np.testing.assert_allclose(back[:, i].reshape(-1, 2).T, PP.node_slip(means[:, i], nxi, neta), atol=5e-6)
