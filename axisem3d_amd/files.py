"""Files the reference reads or writes around the time loop.

* Nu wisdom: rows (s, z, learned Nu, original Nu) of every GLL point, the
  output of Domain::dumpWisdom (SOLVER/src/core/domain/Domain.cpp:404-438) and
  the input of the NU_TYPE wisdom Nr field (WisdomNrField.cpp:20-24), in the
  reference's "NetCDF ascii format" alternative that NuWisdom::readFromFile
  accepts (NuWisdom.cpp:47-79).
* Station traces: one ascii file NET.STA.<components>.ascii per station with a
  line "t c0 c1 c2" per recorded step (PointwiseIOAscii.cpp:25-35, 96-102).
"""
import ctypes as C
import os

import numpy as np

from . import _abi

_f64p = C.POINTER(C.c_double)


def _lib():
    from . import lib
    return lib()


def _check(rc):
    if rc != 0:
        raise RuntimeError(_lib().axg_last_error().decode())


def write_wisdom(path, rows):
    """Write rows [n][4] (s, z, learned Nu, original Nu) as a wisdom file."""
    rows = np.ascontiguousarray(np.asarray(rows, dtype=np.float64).reshape(-1, 4))
    _check(_lib().axg_wisdom_write(os.fsencode(path), rows.shape[0], rows.ctypes.data_as(_f64p)))


def read_wisdom(path):
    """Rows [n][4] of a wisdom file (a 3-column file repeats its third column)."""
    n = C.c_int64(0)
    _check(_lib().axg_wisdom_read(os.fsencode(path), None, 0, C.byref(n)))
    rows = np.zeros((max(n.value, 1), 4), dtype=np.float64)
    _check(_lib().axg_wisdom_read(os.fsencode(path), rows.ctypes.data_as(_f64p), n.value, C.byref(n)))
    return rows[:n.value]


def wisdom_nu(rows, sz, num_samples=4):
    """NuWisdom::getNu (NuWisdom.cpp:82-98) at points sz [m][2]."""
    rows = np.ascontiguousarray(np.asarray(rows, dtype=np.float64).reshape(-1, 4))
    sz = np.ascontiguousarray(np.asarray(sz, dtype=np.float64).reshape(-1, 2))
    out = np.zeros(max(sz.shape[0], 1), dtype=np.int32)
    _check(_lib().axg_wisdom_get_nu(rows.shape[0], rows.ctypes.data_as(_f64p), num_samples, sz.shape[0],
                                    sz.ctypes.data_as(_f64p), out.ctypes.data_as(_abi.i32p)))
    return out[:sz.shape[0]]


def wisdom_rows(mesh, learned_nu):
    """Domain::dumpWisdom rows of a mesh: its points' (s, z), the learned Nu of
    each point (Domain.wisdom()) and its Nu."""
    sz = mesh.array("point_sz").reshape(-1, 2)
    nu = mesh.array("point_nr") // 2
    learned = np.asarray(learned_nu).reshape(-1)
    if learned.shape[0] != sz.shape[0]:
        raise ValueError("wisdom_rows: one learned Nu per point expected")
    return np.column_stack([sz[:, 0], sz[:, 1], learned.astype(np.float64), nu.astype(np.float64)])


def write_stations(outdir, seismograms, dt, t_shift, names=None, components="RTZ", interval=1, kind=""):
    """One ascii file per station from seismograms [nstep][nrec][ncol] (the
    layout of Domain.seismograms(), or of Domain.strain_curl() with kind
    "strain" / "curl", whose files the reference names NET.STA.RTZ.strain.ascii
    and NET.STA.RTZ.curl.ascii); names are "NET.STA" strings (default AX.S0000,
    ...).  Returns the paths."""
    s = np.ascontiguousarray(seismograms, dtype=np.float32)
    nstep, nrec, ncol = s.shape
    if kind:
        components = "RTZ." + kind
    if names is None:
        names = ["AX.S%04d" % i for i in range(nrec)]
    if len(names) != nrec:
        raise ValueError("write_stations: one name per station expected")
    os.makedirs(outdir, exist_ok=True)
    paths = []
    base = s.ctypes.data_as(_abi.f32p)
    for r, name in enumerate(names):
        path = os.path.join(outdir, "%s.%s.ascii" % (name, components))
        ptr = C.cast(C.addressof(base.contents) + 4 * ncol * r, _abi.f32p)
        _check(_lib().axg_write_station_ascii(os.fsencode(path), nstep, interval, t_shift, dt, ncol, ptr,
                                              ncol * nrec))
        paths.append(path)
    return paths
