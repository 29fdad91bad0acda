"""NetCDF classic files: the Nu wisdom (NuWisdom::writeToFile / readFromFile,
NuWisdom.cpp:19-79) and the assembled station file axisem3d_synthetics.nc
(PointwiseIONetCDF.cpp:272-380), written and read by the library's own
classic-format code (no NetCDF library in the image).

Pins: scipy.io.netcdf_file, an independent implementation of the classic
format, reads every file the library writes (names, dimensions, types,
attributes, values) and writes the files the library reads, in both the
32-bit (CDF-1) and 64-bit offset (CDF-2) variants, with fixed and record
(unlimited) dimensions; the station file's time column equals the ascii
station file's, and its series equal the traces.
"""

import numpy as np
import pytest
from scipy.io import netcdf_file

import axisem3d_amd as ax


def _scipy(path):
    return netcdf_file(path, "r", mmap=False)


def test_wisdom_netcdf_read_by_scipy_and_back(tmp_path):
    rng = np.random.default_rng(11)
    for n in (300, 4, 1):
        rows = np.column_stack([rng.uniform(0, 6.4e6, n), rng.uniform(-6.4e6, 6.4e6, n),
                                rng.integers(0, 200, n), rng.integers(0, 200, n)]).astype(np.float64)
        p = str(tmp_path / ("w%d.nc" % n))
        ax.write_wisdom(p, rows, netcdf=True)
        with open(p, "rb") as f:
            assert f.read(4) == b"CDF\x02"
        with _scipy(p) as nc:
            v = nc.variables["axisem3d_wisdom"]
            assert v.dimensions == ("ncdim_%d" % n, "ncdim_4")
            assert v.data.dtype == np.dtype(">f8")
            np.testing.assert_array_equal(v.data, rows)
        np.testing.assert_array_equal(ax.read_wisdom(p), rows)
        np.testing.assert_array_equal(ax.read_netcdf(p, "axisem3d_wisdom"), rows)
    # no rows: the record dimension with no records
    p = str(tmp_path / "w0.nc")
    ax.write_wisdom(p, np.zeros((0, 4)), netcdf=True)
    assert ax.read_wisdom(p).shape == (0, 4)
    with _scipy(p) as nc:
        assert nc.variables["axisem3d_wisdom"].shape == (0, 4)


@pytest.mark.parametrize("version", [1, 2])
def test_wisdom_written_by_scipy(tmp_path, version):
    """Classic files of both variants from another writer, 3 and 4 columns,
    float and double, another variable first, and a record dimension."""
    rng = np.random.default_rng(12)
    r3 = np.column_stack([rng.uniform(0, 1e6, 50), rng.uniform(-1e6, 1e6, 50), rng.integers(1, 90, 50)])
    p = str(tmp_path / "s.nc")
    with netcdf_file(p, "w", version=version) as nc:
        nc.createDimension("a", 7)
        nc.createDimension("rows", 50)
        nc.createDimension("cols", 3)
        other = nc.createVariable("other", "i4", ("a",))
        other[:] = np.arange(7)
        nc.title = "wisdom from scipy"
        v = nc.createVariable("axisem3d_wisdom", "f4", ("rows", "cols"))
        v.units = "m"
        v[:] = r3
    exp = np.column_stack([r3.astype(np.float32).astype(np.float64), r3[:, 2]])
    np.testing.assert_array_equal(ax.read_wisdom(p), exp)
    np.testing.assert_array_equal(ax.read_netcdf(p, "other"), np.arange(7))
    # record (unlimited) dimension, two record variables interleaved
    r4 = np.column_stack([r3, r3[:, 2] + 1])
    q = str(tmp_path / "rec.nc")
    with netcdf_file(q, "w", version=version) as nc:
        nc.createDimension("n", None)
        nc.createDimension("c", 4)
        nc.createDimension("k", 3)
        a = nc.createVariable("axisem3d_wisdom", "f8", ("n", "c"))
        b = nc.createVariable("tag", "i4", ("n", "k"))
        a[:] = r4
        b[:] = np.arange(150).reshape(50, 3)
    np.testing.assert_array_equal(ax.read_wisdom(q), r4)
    np.testing.assert_array_equal(ax.read_netcdf(q, "tag"), np.arange(150).reshape(50, 3))
    # one record variable alone (no record padding)
    q1 = str(tmp_path / "rec1.nc")
    with netcdf_file(q1, "w", version=version) as nc:
        nc.createDimension("n", None)
        v = nc.createVariable("x", "f4", ("n",))
        v[:] = np.arange(5, dtype=np.float32) + 0.5
    np.testing.assert_array_equal(ax.read_netcdf(q1, "x"), np.arange(5) + 0.5)


def test_netcdf_errors(tmp_path):
    p = str(tmp_path / "w.nc")
    ax.write_wisdom(p, np.ones((3, 4)), netcdf=True)
    with pytest.raises(RuntimeError, match="variable: nothing"):
        ax.read_netcdf(p, "nothing")
    h = str(tmp_path / "h5.nc")
    with open(h, "wb") as f:
        f.write(b"\x89HDF\r\n\x1a\n" + bytes(64))
    with pytest.raises(RuntimeError, match="NetCDF-4"):
        ax.read_wisdom(h)
    with pytest.raises(RuntimeError, match="not a NetCDF classic"):
        ax.read_netcdf(h, "axisem3d_wisdom")
    t = str(tmp_path / "trunc.nc")
    with open(p, "rb") as f:
        data = f.read()
    with open(t, "wb") as f:
        f.write(data[:-20])
    with pytest.raises(RuntimeError, match="truncated"):
        ax.read_wisdom(t)
    # a 2-column variable is no wisdom
    q = str(tmp_path / "c2.nc")
    with netcdf_file(q, "w", version=2) as nc:
        nc.createDimension("r", 2)
        nc.createDimension("c", 2)
        nc.createVariable("axisem3d_wisdom", "f8", ("r", "c"))[:] = 1.
    with pytest.raises(RuntimeError, match="3 or 4 columns"):
        ax.read_wisdom(q)


def test_station_netcdf_file(tmp_path):
    rng = np.random.default_rng(13)
    nstep, nrec, dt, shift = 41, 3, 0.37, 5.5
    disp = rng.standard_normal((nstep, nrec, 3)).astype(np.float32)
    strain = rng.standard_normal((nstep, 2, 6)).astype(np.float32)
    curl = rng.standard_normal((nstep, 1, 3)).astype(np.float32)
    names = ["II.AAK", "IU.ANMO", "G.CAN"]
    geo = np.array([[42.6, 74.5, 0.], [34.9, -106.5, 1820.], [-35.3, 149.0, 0.]])
    src = (12.5, -33.25, 25e3)
    p = str(tmp_path / "axisem3d_synthetics.nc")
    for interval in (1, 4):
        ax.write_stations_netcdf(p, disp, dt, shift, names=names, interval=interval, rec_geo=geo, src_geo=src,
                                 strain=strain, strain_rec=[2, 0], curl=curl, curl_rec=[1])
        asc = ax.write_stations(str(tmp_path / ("a%d" % interval)), disp, dt, shift, names=names,
                                interval=interval)
        t_ascii = np.loadtxt(asc[0])[:, 0]
        with _scipy(p) as nc:
            assert list(nc.variables) == ["time_points", "II.AAK.RTZ", "IU.ANMO.RTZ", "G.CAN.RTZ",
                                          "G.CAN.RTZ.strain", "II.AAK.RTZ.strain", "IU.ANMO.RTZ.curl"]
            tp = nc.variables["time_points"]
            assert tp.data.dtype == np.dtype(">f8") and tp.dimensions == ("ncdim_%d" % len(t_ascii),)
            np.testing.assert_array_equal(tp.data, t_ascii)
            for r, nm in enumerate(names):
                v = nc.variables[nm + ".RTZ"]
                assert v.data.dtype == np.dtype(">f4")
                assert v.dimensions == ("ncdim_%d" % len(t_ascii), "ncdim_3")
                np.testing.assert_array_equal(v.data, disp[::interval, r, :])
                assert (v.latitude, v.longitude, v.depth) == tuple(geo[r])
            s = nc.variables["G.CAN.RTZ.strain"]
            assert s.dimensions[1] == "ncdim_6" and s.depth == geo[2, 2]
            np.testing.assert_array_equal(s.data, strain[::interval, 0, :])
            np.testing.assert_array_equal(nc.variables["II.AAK.RTZ.strain"].data, strain[::interval, 1, :])
            c = nc.variables["IU.ANMO.RTZ.curl"]
            np.testing.assert_array_equal(c.data, curl[::interval, 0, :])
            assert c.latitude == geo[1, 0]
            assert (nc.source_latitude, nc.source_longitude, nc.source_depth) == src
        np.testing.assert_array_equal(ax.read_netcdf(p, "IU.ANMO.RTZ"), disp[::interval, 1, :])
    # components and default names; no strain or curl
    ax.write_stations_netcdf(p, disp, dt, shift, components="ENZ")
    with _scipy(p) as nc:
        assert list(nc.variables) == ["time_points", "AX.S0000.ENZ", "AX.S0001.ENZ", "AX.S0002.ENZ"]
    # no recorded steps: the record dimension with no records
    ax.write_stations_netcdf(p, disp[:0], dt, shift)
    with _scipy(p) as nc:
        assert nc.variables["AX.S0001.RTZ"].shape == (0, 3)
    with pytest.raises(ValueError):
        ax.write_stations_netcdf(p, disp, dt, shift, names=names[:2])
    with pytest.raises(RuntimeError, match="strain station index"):
        ax.write_stations_netcdf(p, disp, dt, shift, strain=strain, strain_rec=[0, 3])
