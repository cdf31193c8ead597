"""Shared test setup: import paths, the `gpu` marker and common workloads."""
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("gr-ray-tracer_amd", "oracle", os.path.join("tests", "harness")):
    p = os.path.join(REPO, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle():
    import grt_oracle
    grt_oracle.lib()
    return grt_oracle


@pytest.fixture(scope="session")
def gpu():
    """Skip-free GPU guard: a GPU-marked test fails loudly without a device."""
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import grt_amd
    grt_amd.load()
    return torch.device("cuda:0")


def c1_fields(rho_floor=1e-6, nr=64, nth=32, nph=32, spin=0.9):
    """The first reference configuration's grid: FM torus, a=0.9, r_in=6,
    r_pressure_max=12, r in [1.5, 100]."""
    from grt_amd.fm_torus import fishbone_moncrief
    return fishbone_moncrief(nr=nr, ntheta=nth, nphi=nph, r_min=1.5, r_max=100.0, mass=1.0,
                             spin=spin, r_in=6.0, r_pressure_max=12.0, rho_floor=rho_floor)


def c1_params(spin=0.9, theta=np.pi / 3, **kw):
    """Image mode, camera at r=1000, +-25 M field of view, RK45 rtol=atol=1e-8,
    absorption and emission on (coefficients 1), emitting normalisation."""
    d = dict(a_bh=spin, m_bh=1.0, r_term_min=1.0 + np.sqrt(1.0 - spin * spin),
             r_term_max=1500.0, atol=1e-8, rtol=1e-8, integrator=1, mode=0,
             absorption_enabled=1, emission_enabled=1, absorption_coefficient=1.0,
             emission_coefficient=1.0, superphoton_count_normalization=1e8,
             camera_distance=1000.0, camera_theta=theta, camera_phi=0.0,
             camera_width=50.0, camera_height=50.0)
    d.update(kw)
    return d


def pixel_screen(nx, ny=None, width=50.0, height=50.0):
    ny = ny or nx
    u = -width / 2 + (np.arange(nx) + 0.5) * width / nx
    v = -height / 2 + (np.arange(ny) + 0.5) * height / ny
    U, V = np.meshgrid(u, v, indexing="xy")
    return U.ravel().copy(), V.ravel().copy()
