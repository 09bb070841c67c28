"""Post-processing of main.f90 over the device library: calc_slip_map (main.f90:553-666) and point_cloud
(main.f90:668-830).

calc_slip_map reads geometry samples from an outer-SMC stage file, runs the inner slip SMC of each on the device
(smcf_slip_mean_batch, batched) and writes the plain mean of its final slip particles on every node.
point_cloud reads geometry samples and those mean slips and writes, per fault plane, an 8 x 8 cloud of points per
patch with their position and interpolated slip (host arithmetic; it needs no device).
"""
import math

import numpy as np

from . import driver as D


# ------------------------------------------------------------------ geometry (host restatement)
def get_geometry(iplane, theta, nxi_ls, neta_ls, xmin, xmax, zmin, fix_xbend=False, xbend=()):
    """get_geometry (init.f90:121-173), 0-based plane index: (xl, yl, xr, yr, dip_deg, dip_rad, lxi, leta)."""
    nplane = len(nxi_ls)
    ymin, ymax = theta[0], theta[1]
    ip = iplane + 1  # the reference's 1-based plane number
    if fix_xbend:
        xl, yl = (xmin, ymin) if ip == 1 else (xbend[ip - 2], theta[ip])
        xr, yr = (xmax, ymax) if ip == nplane else (xbend[ip - 1], theta[ip + 1])
        dip = theta[nplane + ip]
    else:
        xl, yl = (xmin, ymin) if ip == 1 else (theta[2 * ip - 2], theta[2 * ip - 1])
        xr, yr = (xmax, ymax) if ip == nplane else (theta[2 * ip], theta[2 * ip + 1])
        dip = theta[2 * nplane + ip - 1]
    dip_rad = dip * (4.0 * math.atan(1.0)) / 180.0
    lxi = math.sqrt((xr - xl) ** 2 + (yr - yl) ** 2)
    leta = -zmin / math.sin(dip_rad)
    return xl, yl, xr, yr, dip, dip_rad, lxi, leta


def discretize_fault(theta, nxi_ls, neta_ls, xmin, xmax, zmin, fix_xbend=False, xbend=()):
    """Node coordinates (2 x nnode) and patch connectivity (4 x npatch, 1-based node ids) of discretize_fault
    (init.f90:198-293)."""
    coor, cny = [], []
    offset = 0
    for ip, (nxi, neta) in enumerate(zip(nxi_ls, neta_ls)):
        _xl, _yl, _xr, _yr, _dip, _dr, lxi, leta = get_geometry(ip, theta, nxi_ls, neta_ls, xmin, xmax, zmin,
                                                               fix_xbend, xbend)
        dxi, deta = lxi / nxi, leta / neta
        for j in range(1, neta + 2):
            for i in range(1, nxi + 2):
                coor.append(((i - 1) * dxi - lxi / 2.0, (j - 1) * deta - leta / 2.0))
        for j in range(1, neta + 1):
            for i in range(1, nxi + 1):
                n1 = i + (nxi + 1) * (j - 1) + offset
                cny.append((n1, n1 + 1, n1 + 2 + nxi, n1 + 1 + nxi))
        offset += (nxi + 1) * (neta + 1)
    return np.array(coor).T, np.array(cny, np.int32).T


def plane_frame(theta, iplane, nxi_ls, neta_ls, xmin, xmax, zmin, fix_xbend=False, xbend=()):
    """(xf, yf, zf, strike_rad, dip_rad) of plane iplane as calc_greens_func places it (gfunc.f90:285-296): the
    plane's centre and orientation in the current geometry layout."""
    xl, yl, xr, yr, _dip, dip_rad, _lxi, _leta = get_geometry(iplane, theta, nxi_ls, neta_ls, xmin, xmax, zmin,
                                                              fix_xbend, xbend)
    strike_rad = math.atan2(xr - xl, yr - yl)
    xf = (xl + xr) / 2.0 - (zmin * math.cos(strike_rad)) / (2.0 * math.tan(dip_rad))
    yf = (yl + yr) / 2.0 + (zmin * math.sin(strike_rad)) / (2.0 * math.tan(dip_rad))
    return xf, yf, zmin / 2.0, strike_rad, dip_rad


# ------------------------------------------------------------------ sample files
def read_samples(path, n, ncol, offset_sample=10, stride=1):
    """Geometry samples of an outer-SMC stage file as calc_slip_map / point_cloud read them: offset_sample
    records skipped, then n samples, each the last of `stride` list-directed READs of ncol values."""
    r = D.Records(path)
    for _ in range(offset_sample):
        r.values(ncol)
    out = np.zeros((ncol, n))
    for i in range(n):
        for _ in range(stride):
            v = r.values(ncol)
        out[:, i] = [D._real(t) for t in v]
    return out


def node_slip(mean, nxi_ls, neta_ls):
    """2 x nnode slip of every node from a DOF-ordered slip vector (zero off the DOF nodes), main.f90:648-656."""
    ids = D.id_dof(nxi_ls, neta_ls)
    nnode = sum((a + 1) * (b + 1) for a, b in zip(nxi_ls, neta_ls))
    slip = np.zeros((2, nnode))
    slip[0, ids - 1] = mean[0::2]
    slip[1, ids - 1] = mean[1::2]
    return slip


# ------------------------------------------------------------------ calc_slip_map
def calc_slip_map(pb, thetas, keys, out_path, nxi_ls, neta_ls):
    """calc_slip_map (main.f90:553-666): for each geometry (column of thetas), the inner slip SMC on the device
    and the plain mean of its final particles, written as one record per geometry of 2 x nnode F12.5 values
    (node-major, strike then dip slip).  Returns the DOF-ordered means (ndim_slip x n) and -log Z."""
    thetas = np.asfortranarray(thetas, dtype=np.float64)
    n = thetas.shape[1]
    means = np.zeros((pb.ndim, n), order="F")
    nl = np.zeros(n)
    for a in range(0, n, pb.max_batch):
        b = min(n, a + pb.max_batch)
        means[:, a:b], nl[a:b] = pb.slip_mean(thetas[:, a:b], keys[a:b])
    with open(out_path, "w") as f:
        for i in range(n):
            slip = node_slip(means[:, i], nxi_ls, neta_ls)
            f.write("".join(D.fortran_f(slip[k, m], 12, 5) for m in range(slip.shape[1]) for k in range(2)) + "\n")
    return means, nl


def read_slip_map(path, n, nnode):
    """The records calc_slip_map writes, read back as point_cloud reads them (2 nnode values per READ)."""
    r = D.Records(path)
    return np.array([[D._real(t) for t in r.values(2 * nnode)] for _ in range(n)]).T


# ------------------------------------------------------------------ point_cloud
def cloud_offsets(nr=8):
    """The nr sample abscissae per patch axis of point_cloud (main.f90:681-685): -1 + 1/nr + 2(k - 1)/nr."""
    return np.array([-1.0 + 1.0 / nr + (k - 1) * 2.0 / nr for k in range(1, nr + 1)])


def point_cloud(thetas, node_slips, nxi_ls, neta_ls, xmin, xmax, zmin, fix_xbend=False, xbend=(), nr=8,
                layout="current"):
    """point_cloud (main.f90:668-830): for each geometry and patch, nr x nr points with global position and
    bilinearly interpolated slip: rows (x, y, z, u_xi, u_eta, |u|), per plane.

    The reference takes each plane's (xf, yf, zf, strike, dip, lxi, leta) from the geometry vector as eight values
    per plane (main.f90:735-743), a layout the outer SMC no longer produces (it writes the current one of
    get_geometry).  layout="stale8" reproduces that reading for vectors in the old layout; layout="current" places
    each plane as calc_greens_func does (gfunc.f90:285-296).  Node coordinates always come from discretize_fault."""
    nr_off = cloud_offsets(nr)
    nplane = len(nxi_ls)
    pts = [[] for _ in range(nplane)]
    for ipart in range(thetas.shape[1]):
        th = thetas[:, ipart]
        coor, cny = discretize_fault(th, nxi_ls, neta_ls, xmin, xmax, zmin, fix_xbend, xbend)
        slip = node_slips[:, ipart].reshape(-1, 2).T  # 2 x nnode, as reshape(mean_slip(:, i), (/2, nnode/))
        offset_patch = 0
        for ip in range(nplane):
            nxi, neta = nxi_ls[ip], neta_ls[ip]
            if layout == "stale8":
                xf, yf, zf, strike, dip = (th[8 * ip + k] for k in range(5))
                srad, drad = strike * (4.0 * math.atan(1.0)) / 180.0, dip * (4.0 * math.atan(1.0)) / 180.0
            else:
                xf, yf, zf, srad, drad = plane_frame(th, ip, nxi_ls, neta_ls, xmin, xmax, zmin, fix_xbend, xbend)
            for ie in range(offset_patch, offset_patch + nxi * neta):
                nodes = cny[:, ie] - 1
                xin, etan = coor[0, nodes], coor[1, nodes]
                uxin, uetan = slip[0, nodes], slip[1, nodes]
                for r1 in nr_off:
                    for r2 in nr_off:
                        nv = np.array([(1 - r1) * (1 - r2), (1 + r1) * (1 - r2), (1 + r1) * (1 + r2),
                                       (1 - r1) * (1 + r2)]) / 4.0
                        xi, eta, uxi, ueta = nv @ xin, nv @ etan, nv @ uxin, nv @ uetan
                        x = xf - eta * math.cos(drad) * math.cos(srad) + xi * math.sin(srad)
                        y = yf + eta * math.cos(drad) * math.sin(srad) + xi * math.cos(srad)
                        z = zf + eta * math.sin(drad)
                        pts[ip].append((x, y, z, uxi, ueta, math.sqrt(uxi ** 2 + ueta ** 2)))
            offset_patch += nxi * neta
    return [np.array(p) for p in pts]


def write_point_cloud(points, path_prefix):
    """One file per plane, <prefix><plane number>, each point a record of six F12.5 values (main.f90:804-828)."""
    for ip, p in enumerate(points):
        with open("%s%d" % (path_prefix, ip + 1), "w") as f:
            for row in p:
                f.write("".join(D.fortran_f(v, 12, 5) for v in row) + "\n")
