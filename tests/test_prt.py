"""Particle relabelling of undulated elements, PRT_1D (SOLVER/src/core/element/
prt/PRT_1D.cpp) in the 1-D solid and fluid element paths: the 3 x 3 gradient
(Gradient::computeGrad9, Gradient.cpp:84-152), its rotation
(CrdTransTIsoSolid.cpp:44-84), the relabelling to the undulated strain and back,
and Gradient::computeQuad9 (Gradient.cpp:154-204).

Oracle pins: with X = identity the relabelled path is the plain one (grad9 ->
6-strain equals grad6, quad9 of a symmetric stress equals quad6); with random
X the operator stays self-adjoint under the Parseval weights (quad9 is the
adjoint of grad9 and undulatedToSpherical the transpose of
sphericalToUndulated) and keeps rigid translations in its null space.
GPU (marked gpu): operator and traces against the oracle.
"""
import numpy as np
import pytest

import axisem3d_amd as ax
from axisem3d_amd import _abi
from tests.helpers import WithPRT, random_field, random_prt, rel_l2


@pytest.fixture(scope="module")
def refcpu():
    from oracle import refcpu as r
    r.lib()
    return r


def _mesh(layers, material=_abi.MAT_TRANSVERSELY_ISO, nstep=10, **kw):
    cfg = ax.shell_config(0, layers=layers, nelem_theta=kw.pop("nelem_theta", 8), material=material, nu_mode=1,
                          nu_min=2, nu_max=9, nstep=nstep, nrec=6, nthread=1, **kw)
    return ax.ShellMesh(cfg)


def _identity(n):
    X = np.zeros((n, 4, 25), dtype=np.float32)
    X[:, 0] = 1.
    X[:, 3] = 1.
    return X


def _stiff(dom, m, us, uf=None):
    dom.set_field(_abi.FIELD_SOLID_DISPL, us)
    dom.set_field(_abi.FIELD_SOLID_STIFF, np.zeros_like(us))
    nf = m.field_size(_abi.FIELD_FLUID_DISPL)
    if nf:
        dom.set_field(_abi.FIELD_FLUID_DISPL, np.zeros(nf) if uf is None else uf)
        dom.set_field(_abi.FIELD_FLUID_STIFF, np.zeros(nf))
    dom.computeStiff()
    out = dom.get_field(_abi.FIELD_SOLID_STIFF)
    return (out, dom.get_field(_abi.FIELD_FLUID_STIFF)) if nf else (out, None)


def _weights(mesh, solid=True):
    nr, fl = mesh.array("point_nr"), mesh.array("point_flags")
    bit, nc = (_abi.PT_SOLID, 3) if solid else (_abi.PT_FLUID, 1)
    out = []
    for p in range(mesh.d.npoint):
        if fl[p] & bit:
            for a in range(nr[p] // 2 + 1):
                out += [1. if a == 0 else 2.] * (2 * nc)
    return np.array(out)


@pytest.mark.parametrize("material", [_abi.MAT_ISOTROPIC, _abi.MAT_TRANSVERSELY_ISO, _abi.MAT_ANISOTROPIC])
def test_identity_relabelling_is_the_plain_path(small_layers, refcpu, material):
    m = _mesh(small_layers, material=material)
    ne = m.d.nelem
    mp = WithPRT(m, list(range(ne)), _identity(ne))
    u = random_field(m, _abi.FIELD_SOLID_DISPL, 3, scale=1.)
    a, _ = _stiff(refcpu.RefDomain(m, double=True), m, u)
    b, _ = _stiff(refcpu.RefDomain(mp, double=True), mp, u)
    assert np.abs(a).max() > 0
    # a relabelled isotropic element also rotates to (R, T, Z) and back, with the
    # descriptor's FP32 sin/cos (orthonormal to ~1e-7); the others rotate either way
    tol = 1e-7 if material == _abi.MAT_ISOTROPIC else 1e-12
    assert rel_l2(b, a) < tol


def test_identity_relabelling_fluid(fluid_layers, refcpu):
    m = _mesh(fluid_layers)
    fl = m.array("elem_flags")
    ne = m.d.nelem
    mp = WithPRT(m, list(range(ne)), _identity(ne))
    assert (fl & _abi.EL_FLUID).any()
    us = random_field(m, _abi.FIELD_SOLID_DISPL, 4, scale=1.)
    uf = random_field(m, _abi.FIELD_FLUID_DISPL, 5, scale=1.)
    a = _stiff(refcpu.RefDomain(m, double=True), m, us, uf)
    b = _stiff(refcpu.RefDomain(mp, double=True), mp, us, uf)
    # the relabelled fluid elements rotate with the FP32 sin/cos (see above)
    assert rel_l2(b[0], a[0]) < 1e-12 and rel_l2(b[1], a[1]) < 1e-7 and np.abs(a[1]).max() > 0


@pytest.mark.parametrize("material", [_abi.MAT_ISOTROPIC, _abi.MAT_ANISOTROPIC])
def test_relabelled_operator_self_adjoint_rigid_null(small_layers, refcpu, material):
    m = _mesh(small_layers, material=material)
    ne = m.d.nelem
    mp = WithPRT(m, list(range(ne)), random_prt(ne, 7))
    dom = refcpu.RefDomain(mp, double=True)
    w = _weights(m)
    u = random_field(m, _abi.FIELD_SOLID_DISPL, 1, scale=1.)
    v = random_field(m, _abi.FIELD_SOLID_DISPL, 2, scale=1.)
    Ku, Kv = -_stiff(dom, mp, u)[0], -_stiff(dom, mp, v)[0]
    a, b = np.dot(w * u, Kv), np.dot(w * v, Ku)
    assert abs(a - b) < 1e-10 * np.linalg.norm(Ku) * np.linalg.norm(v)
    # the relabelling changes the operator
    K0 = -_stiff(refcpu.RefDomain(m, double=True), m, u)[0]
    assert rel_l2(Ku, K0) > 1e-3
    # translation along z (alpha 0) and along x (alpha 1) stay in the null space
    nr, fl = m.array("point_nr"), m.array("point_flags")
    tz, tx = [], []
    for p in range(m.d.npoint):
        if fl[p] & _abi.PT_SOLID:
            for al in range(nr[p] // 2 + 1):
                tz += [0., 0., 0., 0., 1. if al == 0 else 0., 0.]
                tx += [0.5, 0., 0., 0.5, 0., 0.] if al == 1 else [0.] * 6
    for t in (np.array(tz), np.array(tx)):
        r = np.linalg.norm(_stiff(dom, mp, t)[0]) / (np.linalg.norm(Ku) / np.linalg.norm(u) * np.linalg.norm(t))
        assert r < 1e-11, r


def test_relabelled_fluid_operator_self_adjoint(fluid_layers, refcpu):
    m = _mesh(fluid_layers)
    ne = m.d.nelem
    mp = WithPRT(m, list(range(ne)), random_prt(ne, 8))
    dom = refcpu.RefDomain(mp, double=True)
    w = _weights(m, solid=False)
    zs = np.zeros(m.field_size(_abi.FIELD_SOLID_DISPL))
    u = random_field(m, _abi.FIELD_FLUID_DISPL, 1, scale=1.)
    v = random_field(m, _abi.FIELD_FLUID_DISPL, 2, scale=1.)
    Ku, Kv = -_stiff(dom, mp, zs, u)[1], -_stiff(dom, mp, zs, v)[1]
    a, b = np.dot(w * u, Kv), np.dot(w * v, Ku)
    assert abs(a - b) < 1e-10 * np.linalg.norm(Ku) * np.linalg.norm(v)
    assert np.dot(w * u, Ku) > 0.


# ---------------------------------------------------------------------------
# GPU against the oracle
# ---------------------------------------------------------------------------

@pytest.mark.gpu
@pytest.mark.parametrize("material,att", [(_abi.MAT_ISOTROPIC, _abi.ATT_NONE), (_abi.MAT_TRANSVERSELY_ISO, _abi.ATT_FULL),
                                          (_abi.MAT_ANISOTROPIC, _abi.ATT_NONE)])
def test_prt_operator_gpu(fluid_layers, refcpu, material, att):
    print("case prt operator material=%d att=%d" % (material, att), flush=True)
    m = _mesh(fluid_layers, material=material, attenuation=att)
    ne = m.d.nelem
    half = list(range(0, ne, 2))  # relabelled and plain elements side by side
    mp = WithPRT(m, half, random_prt(len(half), 9))
    g = ax.Domain(mp)
    r = refcpu.RefDomain(mp, double=True)
    us = random_field(m, _abi.FIELD_SOLID_DISPL, 11)
    uf = random_field(m, _abi.FIELD_FLUID_DISPL, 12)
    sg, fg = _stiff(g, mp, us, uf)
    sr, fr = _stiff(r, mp, us, uf)
    g.close()
    assert np.abs(sr).max() > 0 and np.abs(fr).max() > 0
    assert rel_l2(sg, sr) < 1e-5 and rel_l2(fg, fr) < 1e-5


@pytest.mark.gpu
def test_prt_seismograms_gpu(fluid_layers, refcpu):
    print("case prt seismograms", flush=True)
    m = _mesh(fluid_layers, nstep=300, nelem_theta=10)
    ne = m.d.nelem
    mp = WithPRT(m, list(range(ne)), random_prt(ne, 10, amp=0.05))
    g = ax.Domain(mp)
    g.resetZero()
    g.run(1, 300, 100, use_graph=True)
    sg = g.seismograms()
    g.close()
    r = refcpu.RefDomain(mp, double=False)
    r.run(1, 300, 100)
    sr = r.seismograms()
    assert np.abs(sr).max() > 0
    assert rel_l2(sg, sr) < 1e-4


# ---------------------------------------------------------------------------
# the relabelling built by the shell builder (Relabelling.cpp)
# ---------------------------------------------------------------------------

def _scaled(layers, f):
    return [dict(l, r_bottom=l["r_bottom"] * f, r_top=l["r_top"] * f) for l in layers]


def test_uniform_stretch_relabelling_equals_a_larger_sphere(fluid_layers, refcpu):
    """deltaR = c r relabels the shell into the shell of radii (1 + c) r: X0 = X3 =
    1 / (1 + c), X1 = 0, Jacobian (1 + c)^3.  Masses, time step, solid-fluid
    normals, source and the stiffness operator of the relabelled mesh equal those
    of the mesh built (1 + c) larger, and so do the station traces (oracle)."""
    c = 0.05
    kw = dict(nelem_theta=8, material=_abi.MAT_ISOTROPIC, nu_mode=1, nu_min=2, nu_max=6, nstep=150, nrec=5,
              rec_depth=0., nthread=1)
    b = ax.ShellMesh(ax.shell_config(0, layers=fluid_layers, undulation=1, undulation_amp=c, src_depth=200e3, **kw))
    a = ax.ShellMesh(ax.shell_config(0, layers=_scaled(fluid_layers, 1 + c), src_depth=200e3 * (1 + c), **kw))
    assert b.d.nelem == a.d.nelem and (np.diff(b.array("prt_off")) == 100).all()
    X = b.array("prt_pool").reshape(-1, 4, 25).astype(np.float64)
    np.testing.assert_allclose(X[:, 0], 1 / (1 + c), rtol=1e-6)
    np.testing.assert_allclose(X[:, 3], 1 / (1 + c), rtol=1e-6)
    assert np.abs(X[:, 1:3]).max() < 1e-6
    assert abs(b.d.dt / a.d.dt - 1) < 1e-9
    for name in ("inv_mass_solid", "inv_mass_fluid"):
        np.testing.assert_allclose(b.array(name), a.array(name), rtol=1e-6)
    np.testing.assert_allclose(b.array("sf_normal"), a.array("sf_normal"), rtol=1e-5, atol=1e-6 * np.abs(a.array("sf_normal")).max())
    np.testing.assert_allclose(b.array("src_force"), a.array("src_force"), rtol=1e-5,
                               atol=1e-6 * np.abs(a.array("src_force")).max())
    u = random_field(a, _abi.FIELD_SOLID_DISPL, 21, scale=1.)
    uf = random_field(a, _abi.FIELD_FLUID_DISPL, 22, scale=1.)
    ka = _stiff(refcpu.RefDomain(a, double=True), a, u, uf)
    kb = _stiff(refcpu.RefDomain(b, double=True), b, u, uf)
    assert rel_l2(kb[0], ka[0]) < 1e-6 and rel_l2(kb[1], ka[1]) < 1e-6
    s = []
    for m in (a, b):
        d = refcpu.RefDomain(m, double=True)
        d.run(1, 150, 0)
        s.append(d.seismograms())
    assert np.abs(s[0]).max() > 0
    assert rel_l2(s[1], s[0]) < 1e-5


def test_topography_relabelling(small_layers, refcpu):
    """A surface topography tapering with depth relabels the elements above the
    taper depth only; X1 follows the colatitude derivative of the undulation
    (X1 = -dZ/dR / (J0 J3)), X2 = 0; the relabelled operator stays
    self-adjoint and keeps rigid translations in its null space."""
    R = small_layers[-1]["r_top"]
    m = ax.ShellMesh(ax.shell_config(0, layers=small_layers, nelem_theta=8, nu_mode=1, nu_min=2, nu_max=6, nstep=10,
                                     undulation=2, undulation_amp=8e3, undulation_order=3, undulation_depth=250e3,
                                     nthread=1))
    off = m.array("prt_off")
    has = np.diff(off) == 100
    ep = m.array("elem_points").reshape(-1, 25)
    sz = m.array("point_sz").reshape(-1, 2)
    rmax = np.hypot(sz[ep, 0], sz[ep, 1]).max(1)
    assert has.any() and (~has).any()
    assert (rmax[has] > R - 250e3).all() and (rmax[~has] <= R - 250e3 + 1.).all()
    X = m.array("prt_pool").reshape(-1, 4, 25).astype(np.float64)
    assert np.abs(X[:, 2]).max() == 0 and np.abs(X[:, 1]).max() > 1e-4
    # at the surface, dZ/dR = (1/r) d/dtheta (A cos(3 theta)) = -3 A sin(3 theta) / R
    pts = ep[has]
    r = np.hypot(sz[pts, 0], sz[pts, 1])
    th = np.arccos(np.clip(sz[pts, 1] / r, -1, 1))
    top = np.abs(r - R) < 1.
    dZ = 8e3 * np.cos(3 * th)
    J0 = 1 + dZ / r
    dZdZ = 8e3 * np.cos(3 * th) / 250e3  # linear taper in r
    J3 = 1 + dZdZ
    J1 = -3 * 8e3 * np.sin(3 * th) / r
    np.testing.assert_allclose(X[:, 0][top], (1 / J0)[top], rtol=1e-5)
    np.testing.assert_allclose(X[:, 3][top], (1 / J3)[top], rtol=2e-3)
    np.testing.assert_allclose(X[:, 1][top], (-J1 / (J0 * J3))[top], rtol=0.05, atol=1e-5)
    dom = refcpu.RefDomain(m, double=True)
    w = _weights(m)
    u = random_field(m, _abi.FIELD_SOLID_DISPL, 1, scale=1.)
    v = random_field(m, _abi.FIELD_SOLID_DISPL, 2, scale=1.)
    Ku, Kv = -_stiff(dom, m, u)[0], -_stiff(dom, m, v)[0]
    assert abs(np.dot(w * u, Kv) - np.dot(w * v, Ku)) < 1e-10 * np.linalg.norm(Ku) * np.linalg.norm(v)
    assert np.dot(w * u, Ku) > 0


@pytest.mark.gpu
def test_topography_seismograms_gpu(fluid_layers, refcpu):
    """Relabelled solid and fluid elements (the taper reaches through the fluid
    layer, so the solid-fluid boundaries are undulated too): GPU traces against
    the oracle's, and a uniform stretch against the larger sphere on the GPU."""
    print("case topography seismograms", flush=True)
    m = ax.ShellMesh(ax.shell_config(0, layers=fluid_layers, nelem_theta=10, nu_mode=1, nu_min=2, nu_max=8,
                                     nstep=300, nrec=6, undulation=2, undulation_amp=20e3, undulation_order=2,
                                     undulation_depth=5500e3))
    fl = m.array("elem_flags")
    has = np.diff(m.array("prt_off")) == 100
    assert has[(fl & _abi.EL_FLUID) != 0].any()
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
    c = 0.05
    kw = dict(nelem_theta=8, nu_mode=1, nu_min=2, nu_max=6, nstep=200, nrec=5, rec_depth=0.)
    b = ax.ShellMesh(ax.shell_config(0, layers=fluid_layers, undulation=1, undulation_amp=c, src_depth=200e3, **kw))
    a = ax.ShellMesh(ax.shell_config(0, layers=_scaled(fluid_layers, 1 + c), src_depth=200e3 * (1 + c), **kw))
    s = []
    for mm in (a, b):
        d = ax.Domain(mm)
        d.resetZero()
        d.run(1, 200, 0, use_graph=True)
        s.append(d.seismograms())
        d.close()
    assert rel_l2(s[1], s[0]) < 1e-4


# ---------------------------------------------------------------------------
# PRT_3D: relabelling on 3-D elements, X per slice (PRT_3D.cpp)
# ---------------------------------------------------------------------------

def _mesh3d(layers, material=_abi.MAT_TRANSVERSELY_ISO, nstep=10, fluid3d=True, **kw):
    cfg = ax.shell_config(0, layers=layers, nelem_theta=kw.pop("nelem_theta", 8), material=material, nu_mode=1,
                          nu_min=2, nu_max=9, nstep=nstep, nrec=6, dv_amp=0.05, dv_fluid=1 if fluid3d else 0,
                          nthread=1, **kw)
    return ax.ShellMesh(cfg)


def _x3d(m, elems, seed=None, amp=0.1):
    """Per-slice X of each element: identity, or random near identity."""
    nr = m.array("point_nr")
    ep = m.array("elem_points").reshape(-1, 25)
    out = []
    for i, e in enumerate(elems):
        rows = int(nr[ep[e]].max())
        if seed is None:
            X = np.zeros((rows, 4, 25), np.float32)
            X[:, 0] = 1.
            X[:, 3] = 1.
        else:
            X = random_prt(1, seed + i, amp, rows=max(rows, 2))[0][:rows]
        out.append(X)
    return out


@pytest.mark.parametrize("material", [_abi.MAT_ISOTROPIC, _abi.MAT_TRANSVERSELY_ISO])
def test_identity_prt3d_is_the_plain_path(fluid_layers, refcpu, material):
    m = _mesh3d(fluid_layers, material=material)
    fl = m.array("elem_flags")
    el = list(np.flatnonzero(fl & _abi.EL_3D))
    assert (fl[el] & _abi.EL_FLUID).any() and (~fl[el] & _abi.EL_FLUID).any()
    mp = WithPRT(m, el, _x3d(m, el))
    us = random_field(m, _abi.FIELD_SOLID_DISPL, 3, scale=1.)
    uf = random_field(m, _abi.FIELD_FLUID_DISPL, 4, scale=1.)
    a = _stiff(refcpu.RefDomain(m, double=True), m, us, uf)
    b = _stiff(refcpu.RefDomain(mp, double=True), mp, us, uf)
    tol = 1e-7 if material == _abi.MAT_ISOTROPIC else 1e-12
    assert rel_l2(b[0], a[0]) < tol and rel_l2(b[1], a[1]) < 1e-7


def test_prt3d_operator_self_adjoint(small_layers, refcpu):
    m = _mesh3d(small_layers, material=_abi.MAT_ANISOTROPIC)
    el = list(range(m.d.nelem))
    mp = WithPRT(m, el, _x3d(m, el, seed=30))
    dom = refcpu.RefDomain(mp, double=True)
    w = _weights(m)
    u = random_field(m, _abi.FIELD_SOLID_DISPL, 1, scale=1.)
    v = random_field(m, _abi.FIELD_SOLID_DISPL, 2, scale=1.)
    Ku, Kv = -_stiff(dom, mp, u)[0], -_stiff(dom, mp, v)[0]
    assert abs(np.dot(w * u, Kv) - np.dot(w * v, Ku)) < 1e-10 * np.linalg.norm(Ku) * np.linalg.norm(v)
    assert np.dot(w * u, Ku) > 0
    K0 = -_stiff(refcpu.RefDomain(m, double=True), m, u)[0]
    assert rel_l2(Ku, K0) > 1e-3


@pytest.mark.gpu
@pytest.mark.parametrize("att", [_abi.ATT_NONE, _abi.ATT_FULL])
def test_prt3d_operator_gpu(fluid_layers, refcpu, att):
    print("case prt3d operator att=%d" % att, flush=True)
    m = _mesh3d(fluid_layers, attenuation=att)
    fl = m.array("elem_flags")
    el = list(np.flatnonzero(fl & _abi.EL_3D))[::2]
    mp = WithPRT(m, el, _x3d(m, el, seed=40))
    g = ax.Domain(mp)
    r = refcpu.RefDomain(mp, double=True)
    us = random_field(m, _abi.FIELD_SOLID_DISPL, 11)
    uf = random_field(m, _abi.FIELD_FLUID_DISPL, 12)
    sg, fg = _stiff(g, mp, us, uf)
    sr, fr = _stiff(r, mp, us, uf)
    g.close()
    assert rel_l2(sg, sr) < 1e-5 and rel_l2(fg, fr) < 1e-5


@pytest.mark.gpu
def test_prt3d_seismograms_gpu(fluid_layers, refcpu):
    print("case prt3d seismograms", flush=True)
    m = _mesh3d(fluid_layers, nstep=300, nelem_theta=10, rec_depth=4000e3)
    fl = m.array("elem_flags")
    el = list(np.flatnonzero(fl & _abi.EL_3D))
    mp = WithPRT(m, el, _x3d(m, el, seed=50, amp=0.05))
    g = ax.Domain(mp)
    g.resetZero()
    g.run(1, 300, 100, use_graph=True)
    sg = g.seismograms()
    g.close()
    r = refcpu.RefDomain(mp, double=False)
    r.run(1, 300, 100)
    sr = r.seismograms()
    assert np.abs(sr).max() > 0
    assert rel_l2(sg, sr) < 1e-4


def test_builder_prt3d_on_a_3d_model(small_layers, refcpu):
    """An undulated 3-D model: relabelled 3-D elements carry PRT_3D rows (the
    same X on every slice of an azimuth-independent undulation), which equal
    the PRT_1D rows of the same undulation on the 1-D model."""
    kw = dict(layers=small_layers, nelem_theta=8, nu_mode=1, nu_min=2, nu_max=6, nstep=10, nthread=1,
              undulation=2, undulation_amp=8e3, undulation_order=3, undulation_depth=250e3)
    m3 = ax.ShellMesh(ax.shell_config(0, dv_amp=0.05, **kw))
    m1 = ax.ShellMesh(ax.shell_config(0, **kw))
    f3, off3, off1 = m3.array("elem_flags"), m3.array("prt_off"), m1.array("prt_off")
    nr = m3.array("point_nr")[m3.array("elem_points").reshape(-1, 25)].max(1)
    rows = np.diff(off3) // 100
    has = rows > 0
    assert (has == (np.diff(off1) > 0)).all() and (f3[has] & _abi.EL_3D).any()
    assert (rows[has & ((f3 & _abi.EL_3D) != 0)] == nr[has & ((f3 & _abi.EL_3D) != 0)]).all()
    p3, p1 = m3.array("prt_pool"), m1.array("prt_pool")
    for e in np.flatnonzero(has)[:10]:
        x1 = p1[off1[e]:off1[e] + 100]
        x3 = p3[off3[e]:off3[e + 1]].reshape(-1, 100)
        assert (x3 == x1).all()


@pytest.mark.gpu
def test_builder_prt3d_seismograms_gpu(fluid_layers, refcpu):
    print("case builder prt3d seismograms", flush=True)
    m = ax.ShellMesh(ax.shell_config(0, layers=fluid_layers, nelem_theta=10, nu_mode=1, nu_min=2, nu_max=8,
                                     nstep=300, nrec=6, dv_amp=0.05, dv_fluid=1, material=_abi.MAT_TRANSVERSELY_ISO,
                                     undulation=2, undulation_amp=20e3, undulation_order=2, undulation_depth=5500e3))
    rows = np.diff(m.array("prt_off")) // 100
    assert (rows > 1).any()
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
