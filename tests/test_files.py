"""Files around the time loop: the Nu wisdom file and its Nr field
(NuWisdom.cpp, WisdomNrField.cpp; the reference's "NetCDF ascii format",
NetCDF_ReaderAscii.cpp) and the ascii station files (PointwiseIOAscii.cpp).

Pins: the fixtures under tests/golden follow the ascii layout documented with
the reference's reader (comment lines, other variables, values split over lines
and tabs, a 3-column variant); the interpolation is checked against a
brute-force nearest-neighbour search in numpy; a wisdom written from a mesh and
read back gives that mesh's Nr field again (exact hits); station files read back
with numpy reproduce the FP32 traces exactly and the reference's time column.
"""
import os

import numpy as np
import pytest

import axisem3d_amd as ax
from axisem3d_amd import _abi

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_read_golden_wisdom_files():
    r4 = ax.read_wisdom(os.path.join(GOLDEN, "wisdom_ascii_4col.txt"))
    np.testing.assert_array_equal(r4, [[0., 6371000., 3, 10], [1000.5, 6e6, 7, 10], [2.5e6, -3e6, 12, 20],
                                       [3e6, 1e6, 5, 20], [6e6, 2e5, 0, 4]])
    r3 = ax.read_wisdom(os.path.join(GOLDEN, "wisdom_ascii_3col.txt"))
    np.testing.assert_array_equal(r3, [[10., 20., 4, 4], [30., 40., 9, 9]])


def test_wisdom_round_trip_and_errors(tmp_path):
    rng = np.random.default_rng(1)
    rows = np.column_stack([rng.uniform(0, 6.4e6, 500), rng.uniform(-6.4e6, 6.4e6, 500),
                            rng.integers(0, 200, 500), rng.integers(0, 200, 500)]).astype(np.float64)
    p = str(tmp_path / "w.txt")
    ax.write_wisdom(p, rows)
    np.testing.assert_array_equal(ax.read_wisdom(p), rows)
    with open(p) as f:
        lines = f.read().splitlines()
    assert lines[0] == "@@ NETCDF ASCII FORMAT @@"
    bad = str(tmp_path / "bad.txt")
    with open(bad, "w") as f:
        f.write("\n".join(lines[:-1]) + "\n")  # no @VAR_END
    with pytest.raises(RuntimeError, match="@VAR_END"):
        ax.read_wisdom(bad)
    with open(bad, "w") as f:
        f.write("\n".join(lines[1:]) + "\n")  # no header line
    with pytest.raises(RuntimeError, match="Error identifying"):
        ax.read_wisdom(bad)
    with open(bad, "w") as f:
        f.write("@@ NETCDF ASCII FORMAT @@\n@VAR_START axisem3d_wisdom\n2 5\n" + " 1" * 10 + "\n"
                "@VAR_END axisem3d_wisdom\n")
    with pytest.raises(RuntimeError, match="3 or 4 columns"):
        ax.read_wisdom(bad)


def _brute_nu(rows, q, k=4):
    d = np.hypot(rows[:, 0] - q[0], rows[:, 1] - q[1])
    idx = np.argsort(d, kind="stable")[:k]
    if d[idx].min() < 1e-10:
        return int(rows[idx[np.argmin(d[idx])], 2])
    w = 1. / d[idx]
    return int(np.round((rows[idx, 2] * w).sum() / w.sum()))


def test_wisdom_nu_matches_brute_force():
    rng = np.random.default_rng(2)
    n = 3000
    rows = np.column_stack([rng.uniform(0, 6.4e6, n), rng.uniform(-6.4e6, 6.4e6, n),
                            rng.integers(0, 300, n), np.full(n, 300)]).astype(np.float64)
    # queries inside, outside the bounding box and on the rows themselves
    q = np.concatenate([np.column_stack([rng.uniform(-1e6, 7.4e6, 400), rng.uniform(-7.4e6, 7.4e6, 400)]),
                        rows[:50, :2]])
    got = ax.wisdom_nu(rows, q)
    exp = np.array([_brute_nu(rows, x) for x in q])
    # ties of the inverse-distance mean at exactly .5 are measure-zero
    assert (got == exp).all(), np.flatnonzero(got != exp)
    np.testing.assert_array_equal(got[-50:], rows[:50, 2])
    # fewer rows than samples: all rows
    few = rows[:3]
    assert ax.wisdom_nu(few, [[1e6, 2e6]])[0] == _brute_nu(few, [1e6, 2e6], k=3)


def test_wisdom_nr_field_reproduces_the_mesh(small_layers, tmp_path):
    """A mesh's own Nu written as wisdom and read back by the NU_TYPE wisdom
    field (every point an exact hit) gives the same odd Nr (lucky numbers off,
    so that Nr = 2 Nu + 1 up to the upper limit); a factor scales it."""
    kw = dict(layers=small_layers, nelem_theta=10, nstep=10, lucky=0)
    a = ax.ShellMesh(ax.shell_config(0, nu_mode=1, nu_min=3, nu_max=40, **kw))
    nr_a = a.array("point_nr")
    rows = ax.wisdom_rows(a, nr_a // 2)
    p = str(tmp_path / "wisdom.txt")
    ax.write_wisdom(p, rows)
    b = ax.ShellMesh(ax.shell_config(0, wisdom=ax.read_wisdom(p), **kw))
    nr_b = b.array("point_nr")
    odd = nr_a % 2 == 1
    assert odd.sum() > 0.5 * nr_a.size
    np.testing.assert_array_equal(nr_b[odd], nr_a[odd])
    assert (nr_b[~odd] >= nr_a[~odd]).all()
    # NU_WISDOM_REUSE_FACTOR
    h = ax.ShellMesh(ax.shell_config(0, wisdom=rows, wisdom_factor=0.5, **kw))
    assert h.array("point_nr").sum() < 0.7 * nr_b.sum()
    # a finer mesh of the same shell interpolates between the rows
    f = ax.ShellMesh(ax.shell_config(0, wisdom=rows, layers=small_layers, nelem_theta=16, nstep=10, lucky=0))
    nu_f = f.array("point_nr") // 2
    assert nu_f.min() >= 1 and nu_f.max() <= nr_a.max()


def test_station_ascii_files(tmp_path):
    rng = np.random.default_rng(3)
    nstep, nrec = 37, 4
    s = rng.standard_normal((nstep, nrec, 3)).astype(np.float32) * 1e-3
    dt, shift = 0.731, 12.5
    paths = ax.write_stations(str(tmp_path), s, dt, shift, names=["II.AAK", "IU.ANMO", "G.CAN", "XX.X1"])
    assert [os.path.basename(x) for x in paths] == ["II.AAK.RTZ.ascii", "IU.ANMO.RTZ.ascii", "G.CAN.RTZ.ascii",
                                                    "XX.X1.RTZ.ascii"]
    t = np.zeros(nstep)
    acc = -shift
    for i in range(nstep):
        t[i] = acc
        acc += dt
    for r, path in enumerate(paths):
        d = np.loadtxt(path)
        assert d.shape == (nstep, 4)
        np.testing.assert_array_equal(d[:, 0], t)
        np.testing.assert_array_equal(d[:, 1:].astype(np.float32), s[:, r, :])
    # record interval: step indices 0, 3, 6, ...
    p3 = ax.write_stations(str(tmp_path / "i3"), s, dt, shift, components="ENZ", interval=3)
    d = np.loadtxt(p3[1])
    np.testing.assert_array_equal(d[:, 0], t[::3])
    np.testing.assert_array_equal(d[:, 1:].astype(np.float32), s[::3, 1, :])
    assert p3[1].endswith("AX.S0001.ENZ.ascii")
    # strain (6 columns) and curl files
    st = rng.standard_normal((nstep, 2, 6)).astype(np.float32)
    ps = ax.write_stations(str(tmp_path / "st"), st, dt, shift, kind="strain")
    assert ps[0].endswith("AX.S0000.RTZ.strain.ascii")
    d = np.loadtxt(ps[1])
    assert d.shape == (nstep, 7)
    np.testing.assert_array_equal(d[:, 1:].astype(np.float32), st[:, 1, :])
