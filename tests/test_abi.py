"""The C-ABI library loads on a machine without a GPU and exports every
function include/axg.h declares (no compute calls are made here)."""
import ctypes
import os
import re

import axisem3d_amd as ax

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "axg.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(axg_[a-z_0-9]+)\s*\(", txt, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = ax.lib()
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_no_gpu_fallback_message():
    # creating a domain without a device must fail loudly, not fall back to a CPU path
    import torch  # noqa: F401  (torch only reports whether a device exists)
    if torch.cuda.is_available():
        return
    m = ax.ShellMesh(which=0, nelem_theta=4, nstep=5)
    try:
        ax.Domain(m)
    except RuntimeError as e:
        assert "HIP" in str(e)
    else:
        raise AssertionError("axg_create succeeded without a GPU")
