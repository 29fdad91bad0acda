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


def test_cpp_facade_driver_links_and_fails_loudly_without_gpu():
    """The C++ Domain / Newmark facade (include/axg_domain.hpp) compiles into a
    program linked against libaxg; without a device it reports the HIP error
    and exits non-zero instead of running anything on the CPU."""
    import subprocess
    import torch
    exe = os.path.join(ROOT, "axisem3d_amd", "_build", "axg_solve")
    assert os.path.exists(exe)
    if torch.cuda.device_count() > 0:
        return
    r = subprocess.run([exe, "0", "5", "4"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1
    assert "HIP" in r.stderr
