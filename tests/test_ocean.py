"""Ocean load on the solid surface: MassOcean1D / MassOcean3D
(SOLVER/src/core/point/mass/MassOcean1D.cpp, MassOcean3D.cpp), built where
GLLPoint::release builds them (SOLVER/src/preloop/mesh/GLLPoint.cpp:55-72).

CPU pins of the builder and the oracle:
* the ocean rests on the solid points of the outer surface, and its mass is
  that of a water shell: sum of the surface areas per unit azimuth = 2 R^2
  (the integral of R sin(theta) R dtheta over [0, pi]);
* MassOcean1D divides the radial part of the force by m + m_o and the rest by
  m, whatever the rotation the reference writes it with;
* MassOcean3D agrees with a brute-force DFT of the same product on the slices,
  with the FFTW conventions (C2R unscaled, R2C times 1/Nr);
* a MassOcean1D mesh rewritten as MassOcean3D rows gives the same traces, and
  rows with a zero ocean mass give the traces of the plain mesh.
GPU (marked gpu): the same operators and traces against the oracle.
"""
import numpy as np
import pytest

import axisem3d_amd as ax
from axisem3d_amd import _abi
from tests.helpers import OceanRows, lift_ocean3d, random_field, rel_l2

WATER = 1027.


@pytest.fixture(scope="module")
def refcpu():
    from oracle import refcpu as r
    r.lib()
    return r


def ocean_mesh(layers, depth=4000., var=0., dv_amp=0., nu=(3, 12), nstep=60, nelem_theta=8, **kw):
    cfg = ax.shell_config(0, layers=layers, nelem_theta=nelem_theta, nu_mode=1, nu_min=nu[0], nu_max=nu[1],
                          nstep=nstep, nrec=6, dv_amp=dv_amp, dv_corr_len=800e3, ocean_depth=depth,
                          ocean_var=var, nthread=1, **kw)
    return ax.ShellMesh(cfg)


def solid_blocks(mesh, field):
    """{point: complex array [nu+1][3]} of a solid field."""
    nr, fl = mesh.array("point_nr"), mesh.array("point_flags")
    out, pos = {}, 0
    for p in range(mesh.d.npoint):
        if not fl[p] & _abi.PT_SOLID:
            continue
        n = (nr[p] // 2 + 1) * 3
        blk = field[2 * pos:2 * (pos + n)].reshape(-1, 3, 2)
        out[p] = blk[..., 0] + 1j * blk[..., 1]
        pos += n
    return out


def accel_of(dom, mesh, stiff):
    """Accel after one Newmark update from rest with the given stiff."""
    for w in (_abi.FIELD_SOLID_DISPL, _abi.FIELD_SOLID_VELOC, _abi.FIELD_SOLID_ACCEL):
        dom.set_field(w, np.zeros(mesh.field_size(w)))
    dom.set_field(_abi.FIELD_SOLID_STIFF, stiff)
    dom.updateNewmark(mesh.d.dt)
    return dom.get_field(_abi.FIELD_SOLID_ACCEL)


def test_ocean_rows_on_the_surface(small_layers):
    m = ocean_mesh(small_layers, depth=4000.)
    R = small_layers[-1]["r_top"]
    sz = m.array("point_sz").reshape(-1, 2)
    r = np.hypot(sz[:, 0], sz[:, 1])
    fl = m.array("point_flags")
    off = m.array("ocean_off")
    rows = np.diff(off)
    top = (np.abs(r - R) < 1e-6 * R) & ((fl & _abi.PT_SOLID) != 0)
    assert top.sum() == 4 * m.d.nelem // (sum(l["nelem_radial"] for l in small_layers)) + 1
    assert (rows[top] == 1).all() and (rows[~top] == 0).all()
    o = m.array("ocean").reshape(-1, 4).astype(np.float64)
    iz, ir, st, ct = o.T
    assert (iz < ir).all()
    np.testing.assert_allclose(st ** 2 + ct ** 2, 1., atol=1e-6)
    np.testing.assert_allclose(ct, sz[top, 1] / r[top], atol=1e-6)
    # the solid mass is the point's own Mass1D
    ms = m.array("inv_mass_solid")[m.array("mass_solid_off")[:-1][top]]
    np.testing.assert_allclose(ir, ms, rtol=1e-6)
    # ocean mass = 1027 x depth x area; the areas per unit azimuth sum to 2 R^2
    mo = 1. / iz - 1. / ir
    area = mo / (WATER * 4000.)
    assert abs(area.sum() / (2. * R * R) - 1.) < 1e-4
    # deeper water, more mass, the same areas
    m2 = ocean_mesh(small_layers, depth=8000.)
    o2 = m2.array("ocean").reshape(-1, 4).astype(np.float64)
    area2 = (1. / o2[:, 0] - 1. / o2[:, 1]) / (WATER * 8000.)
    np.testing.assert_allclose(area2, area, rtol=1e-4)
    # no ocean: no rows
    m0 = ocean_mesh(small_layers, depth=0.)
    assert m0.array("ocean_off").size == 0


def test_varying_depth_gives_massocean3d(small_layers):
    m = ocean_mesh(small_layers, depth=4000., var=0.4)
    nr, sz = m.array("point_nr"), m.array("point_sz").reshape(-1, 2)
    rows = np.diff(m.array("ocean_off"))
    on = rows > 0
    axis = np.abs(sz[:, 0]) < 1e-3
    # one row (MassOcean1D) where all slices coincide: on the axis or Nr = 1
    assert (rows[on & (axis | (nr == 1))] == 1).all()
    assert (rows[on & ~axis & (nr > 1)] == nr[on & ~axis & (nr > 1)]).all()
    assert (on & ~axis & (nr > 1)).any()
    o = m.array("ocean").reshape(-1, 4).astype(np.float64)
    off = m.array("ocean_off")
    for p in np.flatnonzero(on & ~axis & (nr > 1))[:5]:
        blk = o[off[p]:off[p + 1]]
        n = np.linalg.norm(blk[:, 1:], axis=1)
        assert np.ptp(n) > 0 and np.ptp(blk[:, 0]) == 0
        # unit normal times sqrt(mo / (m (m + mo))): radial direction
        np.testing.assert_allclose(blk[:, 2], 0., atol=1e-12)
        np.testing.assert_allclose(blk[:, 3] / n, sz[p, 1] / np.hypot(*sz[p]), atol=1e-6)
    with pytest.raises(RuntimeError):
        ocean_mesh(small_layers, depth=4000., var=1.5)


def test_oracle_ocean1d_splits_normal_and_tangential(small_layers, refcpu):
    m = ocean_mesh(small_layers, depth=4000.)
    dom = refcpu.RefDomain(m, double=True)
    s = random_field(m, _abi.FIELD_SOLID_STIFF, seed=4, scale=1.)
    a = solid_blocks(m, accel_of(dom, m, s))
    sb = solid_blocks(m, s)
    off, o = m.array("ocean_off"), m.array("ocean").reshape(-1, 4).astype(np.float64)
    moff, minv = m.array("mass_solid_off"), m.array("inv_mass_solid").astype(np.float64)
    nr, fl = m.array("point_nr"), m.array("point_flags")
    checked = 0
    for p, blk in sb.items():
        if fl[p] & _abi.PT_AXIAL:
            continue
        blk = blk.copy()
        if nr[p] % 2 == 0:
            blk[-1] = 0.
        if off[p + 1] > off[p]:
            iz, ir, st, ct = o[off[p]]
            n = np.array([st, 0., ct])
            sn = blk @ n
            exp = np.outer(sn, n) * iz + (blk - np.outer(sn, n)) * ir
            checked += 1
        else:
            exp = blk * minv[moff[p]]
        # the FP32 rows hold sin and cos with sin^2 + cos^2 = 1 to ~1e-7
        np.testing.assert_allclose(a[p], exp, rtol=1e-6, atol=1e-6 * np.abs(exp).max())
    assert checked > 10


def _c2r_brute(X, nr):
    """x_n = sum_k Z_k exp(+2 pi i k n / Nr) over the Hermitian extension Z of X."""
    Z = np.zeros(nr, complex)
    nu = nr // 2
    Z[:nu + 1] = X[:nu + 1]
    for k in range(1, (nr + 1) // 2):
        Z[nr - k] = np.conj(X[k])
    if nr % 2 == 0:
        Z[nu] = X[nu].real
    Z[0] = X[0].real
    n = np.arange(nr)
    return np.real(np.exp(2j * np.pi * np.outer(n, np.arange(nr)) / nr) @ Z)


def _r2c_brute(x, nr):
    n = np.arange(nr)
    return (np.exp(-2j * np.pi * np.outer(np.arange(nr // 2 + 1), n) / nr) @ x) / nr


def test_oracle_ocean3d_matches_brute_force_dft(small_layers, refcpu):
    m = ocean_mesh(small_layers, depth=4000., var=0.4, nu=(3, 9))
    dom = refcpu.RefDomain(m, double=True)
    s = random_field(m, _abi.FIELD_SOLID_STIFF, seed=5, scale=1.)
    a = solid_blocks(m, accel_of(dom, m, s))
    sb = solid_blocks(m, s)
    off, o = m.array("ocean_off"), m.array("ocean").reshape(-1, 4).astype(np.float64)
    nr, fl = m.array("point_nr"), m.array("point_flags")
    checked = 0
    for p, blk in sb.items():
        if off[p + 1] - off[p] <= 1 or fl[p] & _abi.PT_AXIAL:
            continue
        n_ = nr[p]
        blk = blk.copy()
        blk[0] = blk[0].real
        if n_ % 2 == 0:
            blk[-1] = 0.
        q = o[off[p]:off[p + 1]]
        x = np.stack([_c2r_brute(blk[:, c], n_) for c in range(3)], 1)  # [Nr][3]
        fn = (x * q[:, 1:]).sum(1)
        y = x * q[:, :1] - fn[:, None] * q[:, 1:]
        exp = np.stack([_r2c_brute(y[:, c], n_) for c in range(3)], 1)
        exp[0] = exp[0].real
        if n_ % 2 == 0:
            exp[-1] = 0.
        np.testing.assert_allclose(a[p], exp, rtol=1e-9, atol=1e-12 * np.abs(exp).max())
        checked += 1
    assert checked > 10


def test_oracle_lifted_and_zero_ocean(small_layers, refcpu):
    m1 = ocean_mesh(small_layers, depth=6000., nstep=200)
    runs = {}
    for name, mm in (("1d", m1), ("3d", lift_ocean3d(m1))):
        d = refcpu.RefDomain(mm, double=True)
        d.run(1, 200, 0)
        runs[name] = d.seismograms()
    assert np.abs(runs["1d"]).max() > 0
    assert rel_l2(runs["3d"], runs["1d"]) < 1e-6
    # rows of zero ocean mass (1/(m+0) = 1/m) against the mesh without rows
    o = m1.array("ocean").reshape(-1, 4).copy()
    o[:, 0] = o[:, 1]
    m0 = OceanRows(m1, m1.array("ocean_off"), o)
    plain = OceanRows(m1, np.zeros(m1.d.npoint + 1, np.int64), np.zeros((0, 4)))
    s = []
    for mm in (m0, plain):
        d = refcpu.RefDomain(mm, double=True)
        d.run(1, 200, 0)
        s.append(d.seismograms())
    assert rel_l2(s[0], s[1]) < 1e-6  # sin^2 + cos^2 of the FP32 rows is 1 to ~1e-7
    # the ocean changes the traces
    assert rel_l2(runs["1d"], s[1]) > 1e-4


# ---------------------------------------------------------------------------
# GPU against the oracle
# ---------------------------------------------------------------------------

@pytest.mark.gpu
@pytest.mark.parametrize("var", [0., 0.4])
def test_ocean_newmark_operator_gpu(small_layers, refcpu, var):
    print("case ocean newmark var=%g" % var, flush=True)
    m = ocean_mesh(small_layers, depth=4000., var=var, dv_amp=0.05 if var else 0.)
    g = ax.Domain(m)
    r = refcpu.RefDomain(m, double=True)
    s = random_field(m, _abi.FIELD_SOLID_STIFF, seed=6, scale=1.)
    ag, ar = accel_of(g, m, s), accel_of(r, m, s)
    g.close()
    assert np.abs(ar).max() > 0
    assert rel_l2(ag, ar) < 1e-6


@pytest.mark.gpu
def test_ocean_seismograms_gpu(small_layers, refcpu):
    print("case ocean seismograms 1-D", flush=True)
    m = ocean_mesh(small_layers, depth=6000., nstep=300)
    s = []
    for mm in (m, lift_ocean3d(m)):
        g = ax.Domain(mm)
        g.resetZero()
        g.run(1, 300, 100, use_graph=True)
        s.append(g.seismograms())
        g.close()
    r = refcpu.RefDomain(m, double=False)
    r.run(1, 300, 100)
    sr = r.seismograms()
    assert np.abs(sr).max() > 0
    assert rel_l2(s[0], sr) < 1e-4
    assert rel_l2(s[1], s[0]) < 1e-5


@pytest.mark.gpu
def test_ocean3d_seismograms_gpu(small_layers, refcpu):
    print("case ocean seismograms 3-D", flush=True)
    m = ocean_mesh(small_layers, depth=5000., var=0.4, dv_amp=0.05, material=_abi.MAT_TRANSVERSELY_ISO,
                   nstep=300)
    assert (np.diff(m.array("ocean_off")) > 1).any()
    g = ax.Domain(m)
    g.resetZero()
    g.run(1, 300, 100, use_graph=True)
    sg = g.seismograms()
    g.close()
    r = refcpu.RefDomain(m, double=False)
    r.run(1, 300, 100)
    sr = r.seismograms()
    assert np.abs(sr).max() > 0
    assert rel_l2(sg, sr) < 1e-4
