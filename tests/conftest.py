import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


@pytest.fixture(scope="session")
def small_layers():
    """Two-layer thin shell used by the fast parity cases (solid only)."""
    return [
        dict(r_bottom=5371e3, r_top=5971e3, nelem_radial=3, fluid=0, vp_bottom=10.0e3, vp_top=9.0e3,
             vs_bottom=5.5e3, vs_top=4.9e3, rho_bottom=4.2e3, rho_top=3.7e3, q_mu=0., q_kappa=57823.),
        dict(r_bottom=5971e3, r_top=6371e3, nelem_radial=3, fluid=0, vp_bottom=8.9e3, vp_top=5.8e3,
             vs_bottom=4.8e3, vs_top=3.2e3, rho_bottom=3.5e3, rho_top=2.6e3, q_mu=0., q_kappa=57823.),
    ]


@pytest.fixture(scope="session")
def fluid_layers():
    """Solid / fluid / solid shell with both boundary orientations."""
    return [
        dict(r_bottom=1000e3, r_top=1221.5e3, nelem_radial=1, fluid=0, vp_bottom=11.3e3, vp_top=11.0e3,
             vs_bottom=3.7e3, vs_top=3.5e3, rho_bottom=13.1e3, rho_top=12.8e3, q_mu=0., q_kappa=57823.),
        dict(r_bottom=1221.5e3, r_top=3480e3, nelem_radial=3, fluid=1, vp_bottom=10.3e3, vp_top=8.0e3,
             vs_bottom=0., vs_top=0., rho_bottom=12.2e3, rho_top=9.9e3, q_mu=0., q_kappa=57823.),
        dict(r_bottom=3480e3, r_top=6371e3, nelem_radial=4, fluid=0, vp_bottom=11.0e3, vp_top=5.8e3,
             vs_bottom=6.2e3, vs_top=3.2e3, rho_bottom=5.5e3, rho_top=2.6e3, q_mu=0., q_kappa=57823.),
    ]
