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
