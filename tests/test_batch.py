"""Independent simulations batched in one context (axg_create_batch): every
simulation's stations and learned Nu equal its own single-context run bit for
bit (each keeps its dt, source and STF; colouring and launch order per point
are unchanged by the merge).  One simulation carries an ocean load
(MassOcean1D and MassOcean3D points), the others none."""
import numpy as np
import pytest

import axisem3d_amd as ax
from axisem3d_amd import _abi


def _meshes(small_layers, fluid_layers, nstep):
    return [
        ax.ShellMesh(ax.shell_config(0, layers=small_layers, nelem_theta=10, nu_const=4, nstep=nstep, seed=3)),
        ax.ShellMesh(ax.shell_config(1, layers=fluid_layers, nelem_theta=12, material=_abi.MAT_TRANSVERSELY_ISO,
                                     dv_amp=0.05, nu_mode=1, nu_min=4, nu_max=40, nstep=nstep, seed=4)),
        ax.ShellMesh(ax.shell_config(1, layers=small_layers, nelem_theta=8, material=_abi.MAT_ISOTROPIC,
                                     dv_amp=0.05, nu_mode=1, nu_min=6, nu_max=20, nstep=nstep, seed=5,
                                     src_depth=300e3, ocean_depth=4000., ocean_var=0.3)),
    ]


def test_batch_rejects_different_step_counts(small_layers):
    a = ax.ShellMesh(ax.shell_config(0, layers=small_layers, nelem_theta=8, nstep=10))
    b = ax.ShellMesh(ax.shell_config(0, layers=small_layers, nelem_theta=8, nstep=11))
    with pytest.raises(RuntimeError, match="share the number of steps"):
        ax.DomainBatch([a, b])


@pytest.mark.gpu
def test_batch_equals_single_runs(small_layers, fluid_layers):
    print("case batch", flush=True)
    nstep = 150
    ms = _meshes(small_layers, fluid_layers, nstep)
    assert len({m.d.dt for m in ms}) == 3
    b = ax.DomainBatch(ms)
    b.setLearning(1e-2, 5)
    b.resetZero()
    b.run(1, nstep, 50, use_graph=True)
    for k, m in enumerate(ms):
        d = ax.Domain(m)
        d.setLearning(1e-2, 5)
        d.resetZero()
        d.run(1, nstep, 50, use_graph=True)
        s1, sb = d.seismograms(), b.seismograms(k)
        assert np.abs(s1).max() > 0
        assert np.array_equal(sb, s1), (k, float(np.abs(sb - s1).max()))
        assert np.array_equal(b.wisdom(k), d.wisdom())
        d.close()
    # per-call stepping of a batch takes dt = 0 (each simulation's own)
    with pytest.raises(RuntimeError, match="own dt"):
        b.updateNewmark(1.0)
    b.updateNewmark(0.0)
