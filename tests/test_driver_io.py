"""Run files of the reference (main.f90:99-307, 447-560; init.f90:5-80; smc_slip.f90:1632-1660) read and written
by smc_inversion_fort_amd.driver.  The reference ships no data files; these inputs are written in the layout its
READ statements expect."""
import numpy as np
import pytest

from smc_inversion_fort_amd import driver as D


def _write(path, text):
    path.write_text(text)
    return str(path)


SETTING_2PLANES = """number of fault planes
2
fix-xbend
.true.
xbend
0.5

number of samples for fault
4000
number of samples for slip
2000

observation path
data/obs.csv
number of SAR paths
2
output directory path
output/run1/

max slip(m)
10.0

nxi
3, 2
neta
2 2

xmin
-10.0
xmax
12.5
zmin
-10d0

GSMC
F
"""


def test_settings_two_planes(tmp_path):
    st = D.read_settings(_write(tmp_path / "setting.dat", SETTING_2PLANES))
    assert st["nplane"] == 2 and st["fix_xbend"] is True and st["xbend"] == [0.5]
    assert st["nparticle_fault"] == 4000 and st["nparticle_slip"] == 2000
    assert st["observation_path"] == "data/obs.csv" and st["npath"] == 2
    assert st["output_dir"] == "output/run1/"
    assert st["max_slip"] == 10.0 and st["nxi_ls"] == [3, 2] and st["neta_ls"] == [2, 2]
    assert (st["xmin"], st["xmax"], st["zmin"]) == (-10.0, 12.5, -10.0)
    assert st["gsmc"] is False


def test_settings_one_plane_reads_an_empty_xbend_record(tmp_path):
    text = SETTING_2PLANES.replace("2\nfix-xbend", "1\nfix-xbend").replace("0.5\n", "\n", 1) \
        .replace("3, 2\n", "16\n").replace("2 2\n", "16\n").replace(".true.", "false")
    st = D.read_settings(_write(tmp_path / "setting.dat", text))
    assert st["nplane"] == 1 and st["fix_xbend"] is False and st["xbend"] == []
    assert st["nxi_ls"] == [16] and st["neta_ls"] == [16] and st["npath"] == 2


def test_list_directed_records(tmp_path):
    p = _write(tmp_path / "x.dat", "1 2\n\n 3, 4 5 / 9\n2*7.5 'a b'\nlast\n")
    r = D.Records(p)
    assert r.values(3) == ["1", "2", "3"]        # items continue over records; blank records skipped
    assert r.values(3) == ["7.5", "7.5", "a b"]  # the next READ starts at the next record
    r = D.Records(p)
    assert r.values(2) == ["1", "2"]             # the rest of the record is dropped
    assert r.values(4) == ["3", "4", "5", None]  # '/' ends the READ
    assert r.values(3) == ["7.5", "7.5", "a b"]  # repeat count, quoted string
    assert r.line() == "last"


OBS = """x, y, ex, ey, ez, dlos, sigma, type
1.0, 2.0, -0.6, -0.1, 0.78, 0.5, 1.0, InSAR_asc
3.0, 4.0, -0.6, -0.1, 0.78, 0.6, 1.1, InSAR_asc
5.0, 6.0, 0.6, -0.1, 0.78, 0.7, 1.2, InSAR_dsc

7.0, 8.0, 1.0, 0.0, 0.0, 0.1, 0.5, GNSS
7.0, 8.0, 0.0, 1.0, 0.0, 0.2, 0.5, GNSS
7.0, 8.0, 0.0, 0.0, 1.0, 0.3, 0.5, GNSS
"""


def test_observation_paths_and_gnss(tmp_path):
    o = D.read_observation(_write(tmp_path / "obs.csv", OBS), 2)
    assert o["nsar_ls"] == [2, 1] and o["nsar_index"].tolist() == [1, 3, 4]
    assert o["nsar_total"] == 3 and o["ngnss"] == 1
    np.testing.assert_array_equal(o["obs_points"][:, 2], [5.0, 6.0])
    np.testing.assert_array_equal(o["obs_unitvec"][:, 3], [1.0, 0.0, 0.0])
    np.testing.assert_array_equal(o["dvec"], [0.5, 0.6, 0.7, 0.1, 0.2, 0.3])
    np.testing.assert_array_equal(o["obs_sigma"], [1.0, 1.1, 1.2, 0.5, 0.5, 0.5])
    with pytest.raises(ValueError):
        D.read_observation(str(tmp_path / "obs.csv"), 1)


def test_range_and_value_files(tmp_path):
    rng = D.read_range(_write(tmp_path / "range.dat", "ymin\n-20, -14\nymax\n14 20\ndip\n36.0 54.0\n"), 3)
    np.testing.assert_array_equal(rng, [[-20, 14, 36], [-14, 20, 54]])
    v = D.read_values(_write(tmp_path / "peak.dat", "1.5\n2.5 9\n3.5\n"), 3, one_per_read=True)
    np.testing.assert_array_equal(v, [1.5, 2.5, 3.5])
    v = D.read_values(_write(tmp_path / "q.dat", "1 2\n3\n4\n"), 4)
    np.testing.assert_array_equal(v, [1, 2, 3, 4])


def test_fortran_f_editing():
    assert D.fortran_f(0.5, 12, 5) == "     0.50000"
    assert D.fortran_f(-1.25, 12, 5) == "    -1.25000"
    assert D.fortran_f(1e10, 20, 5) == "   10000000000.00000"
    assert D.fortran_f(1e10, 12, 5) == "*" * 12


@pytest.mark.parametrize("name", ["p1_3x2", "p1_4x3", "p1_8x8", "p2_free", "p2_fixed"])
def test_id_dof_matches_reference_golden(golden, name):
    from test_oracle_golden import FAULTS
    _, nxi, neta = FAULTS[name][:3]
    np.testing.assert_array_equal(D.id_dof(nxi, neta), golden("mesh")[name + "__id_dof"])


def test_slip_particle_file_layout(tmp_path):
    nxi, neta = [3], [2]
    ids = D.id_dof(nxi, neta)
    nd, P = 2 * len(ids), 3
    X = np.arange(nd * P, dtype=float).reshape(nd, P, order="F") / 10.0
    p = str(tmp_path / "slip.dat")
    D.write_slip_particles(p, X, np.array([1.0, 2.0, 3.0]), np.array([0.1, 0.2, 0.3]), nxi, neta)
    lines = open(p).read().splitlines()
    nnode = 4 * 3
    assert len(lines) == P and all(len(ln) == 12 * (2 * nnode + 2) for ln in lines)
    vals = [float(lines[1][12 * k:12 * (k + 1)]) for k in range(2 * nnode + 2)]
    n0 = ids[0] - 1  # first DOF node: its strike and dip slip are the particle's first two coordinates
    assert vals[2 * n0] == pytest.approx(X[0, 1]) and vals[2 * n0 + 1] == pytest.approx(X[1, 1])
    assert vals[0] == 0.0 and vals[-2:] == [2.0, 0.2]
