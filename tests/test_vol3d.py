"""3-D volumetric models (Volumetric3D::get3dProperties; the reference-type
algebra of Material.cpp:145-209) in the shell builder.

CPU pins, each against a mesh built without the model:
* REFERENCE_1D with value 0 is the 1-D model itself, bit for bit;
* REFERENCE_3D with a constant value v on VS is the model with vs (1 + v);
* REFERENCE_PERT with v on VP and VS over the built-in random perturbation
  dv (a REFERENCE_1D perturbation) is the perturbation dv (1 + v);
* ABSOLUTE sets the value (a model returning the layer's own values changes
  nothing) and a local density blob makes its elements 3-D and its points
  Mass3D.
GPU (marked gpu): traces of a model with a density and velocity blob against
the oracle.
"""
import numpy as np
import pytest

import axisem3d_amd as ax
from axisem3d_amd import _abi
from tests.helpers import rel_l2


def _mesh(layers, **kw):
    cfg = dict(layers=layers, nelem_theta=8, nu_mode=1, nu_min=2, nu_max=8, nstep=10, nrec=4, nthread=1)
    cfg.update(kw)
    return ax.ShellMesh(ax.shell_config(0, **cfg))


def _same(a, b, names=("mat_pool", "inv_mass_solid", "inv_mass_fluid", "mass_solid_off"), rtol=0.):
    for n in names:
        x, y = a.array(n), b.array(n)
        assert x.shape == y.shape, n
        if rtol == 0.:
            assert np.array_equal(x, y), n
        else:
            np.testing.assert_allclose(x, y, rtol=rtol, err_msg=n)


def test_reference_1d_zero_is_the_1d_model(small_layers):
    base = _mesh(small_layers, material=_abi.MAT_TRANSVERSELY_ISO)
    m = _mesh(small_layers, material=_abi.MAT_TRANSVERSELY_ISO,
              vol3d=[([_abi.PROP_VP, _abi.PROP_VS, _abi.PROP_RHO], [_abi.REF_REFERENCE_1D] * 3,
                      lambda r, t, p: (0., 0., 0.))])
    _same(base, m)


def test_reference_3d_scales_vs(small_layers):
    v = 0.1
    scaled = [dict(l, vs_bottom=l["vs_bottom"] * (1 + v), vs_top=l["vs_top"] * (1 + v)) for l in small_layers]
    a = _mesh(scaled)
    b = _mesh(small_layers, vol3d=[([_abi.PROP_VS], [_abi.REF_REFERENCE_3D], lambda r, t, p: (v,))])
    _same(a, b, rtol=1e-6)


def test_reference_perturb_scales_the_perturbation(small_layers):
    v = 0.5
    kw = dict(material=_abi.MAT_ISOTROPIC, nu_min=3, nu_max=6, dv_corr_len=800e3)
    a = _mesh(small_layers, dv_amp=0.05 * (1 + v), **kw)
    b = _mesh(small_layers, dv_amp=0.05, vol3d=[([_abi.PROP_VP, _abi.PROP_VS], [_abi.REF_REFERENCE_PERT] * 2,
                                                  lambda r, t, p: (v, v))], **kw)
    assert (a.array("elem_flags") & _abi.EL_3D).any()
    _same(a, b, names=("mat_pool",), rtol=1e-5)


def _blob(R, r0, t0, rad, drho, dvs):
    """A ball of radius rad at (r0, t0, phi = 0) with rho and vs raised (ABSOLUTE
    values relative to a reference density and velocity)."""
    x0 = np.array([r0 * np.sin(t0), 0., r0 * np.cos(t0)])

    def fn(r, t, p):
        x = np.array([r * np.sin(t) * np.cos(p), r * np.sin(t) * np.sin(p), r * np.cos(t)])
        d = np.linalg.norm(x - x0)
        if d > rad:
            return None
        f = 0.5 * (1 + np.cos(np.pi * d / rad))
        return (4000. * (1 + drho * f), 5000. * (1 + dvs * f))
    return [([_abi.PROP_RHO, _abi.PROP_VS], [_abi.REF_ABSOLUTE] * 2, fn)]


def test_absolute_blob_makes_3d_elements_and_mass3d(small_layers):
    R = small_layers[-1]["r_top"]
    base = _mesh(small_layers)
    m = _mesh(small_layers, vol3d=_blob(R, R - 300e3, 0.6, 400e3, 0.2, 0.1))
    fl = m.array("elem_flags")
    assert (fl & _abi.EL_3D).any() and not (fl & _abi.EL_3D).all()
    nm = np.diff(m.array("mass_solid_off"))
    nr = m.array("point_nr")
    assert ((nm == 1) | (nm == nr)).all() and (nm > 1).any()
    # elements outside the blob keep their material
    mo, bo = m.array("mat_off"), base.array("mat_off")
    mp, bp = m.array("mat_pool"), base.array("mat_pool")
    plain = np.flatnonzero((fl & _abi.EL_3D) == 0)
    assert plain.size > 0
    for e in plain:
        assert np.array_equal(mp[mo[e]:mo[e + 1]], bp[bo[e]:bo[e + 1]]), e


@pytest.mark.gpu
def test_blob_seismograms_gpu(small_layers):
    from oracle.refcpu import RefDomain
    print("case volumetric blob", flush=True)
    R = small_layers[-1]["r_top"]
    m = _mesh(small_layers, nstep=300, nelem_theta=10, material=_abi.MAT_TRANSVERSELY_ISO,
              vol3d=_blob(R, R - 300e3, 0.4, 500e3, 0.2, 0.1))
    assert (np.diff(m.array("mass_solid_off")) > 1).any()
    g = ax.Domain(m)
    g.resetZero()
    g.run(1, 300, 100, use_graph=True)
    sg = g.seismograms()
    g.close()
    r = RefDomain(m, double=False)
    r.run(1, 300, 100)
    sr = r.seismograms()
    assert np.abs(sr).max() > 0
    assert rel_l2(sg, sr) < 1e-4
