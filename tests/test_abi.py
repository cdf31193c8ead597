"""The C-ABI library: it loads, exports every entry point declared in
include/grt.h, carries the reference's defaults (utils.cpp:4-37,
simulation_options.hpp:13-32, photon_generation.hpp:40-51) and rejects
invalid configuration with the reference's messages (toml_config.hpp:166-277)
before touching a device.  No kernel runs here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import REPO, gpu_available


def declared_functions():
    names = set()
    for h in ("grt.h",):
        text = open(os.path.join(REPO, "include", h)).read()
        names |= set(re.findall(r"\b(grt_[a-z_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    import grt_amd
    L = grt_amd.load()
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n
    assert set(names) <= set(grt_amd.EXPORTED)
    assert L.grt_abi_version() == 2


def test_geometry_recording_build_exports_the_same_entry_points():
    """The build for gradient companions that read the sample's position, time
    or magnetic field (GRT_GRADIENT_READS_GEOMETRY=1) is the same C ABI."""
    import grt_amd
    L = grt_amd.load(grt_amd.GEOMETRY_LIB_PATH)
    for n in declared_functions():
        assert hasattr(L, n), n
    assert L.grt_abi_version() == grt_amd.load().grt_abi_version()
    assert b"gfx950" in open(grt_amd.GEOMETRY_LIB_PATH, "rb").read()


def test_library_is_in_tree_and_built_for_gfx950():
    import grt_amd
    assert os.path.dirname(grt_amd.LIB_PATH) == os.path.join(REPO, "gr-ray-tracer_amd")
    blob = open(grt_amd.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


def test_reference_defaults():
    import grt_amd
    p = grt_amd.default_params()
    assert (p.m_bh, p.a_bh) == (1.0, 1.0)
    assert (p.camera_theta, p.camera_phi) == (np.pi / 4, np.pi / 4)
    assert (p.camera_width, p.camera_height, p.camera_distance) == (20.0, 20.0, 10000.0)
    assert (p.r_term_min, p.r_term_max) == (1.0, 15000.0)
    assert (p.atol, p.rtol) == (1e-8, 1e-8)
    assert (p.safety, p.min_step_scale, p.max_step_scale) == (0.9, 0.01, 10.0)
    assert p.integrator == grt_amd.GRT_RK45 and p.mode == grt_amd.MODE_IMAGE
    assert (p.albedo, p.seed) == (1.0, 12345)
    assert p.superphoton_count_normalization == 1.0 and p.energy_per_cell_erg == 1.0
    assert (p.power_law_slope, p.nu_min_hz, p.nu_max_hz) == (2.0, 1e9, 1e20)
    assert p.custom_frequency_hz == 2.30e11
    assert p.precision == grt_amd.PRECISION_FP64


@pytest.mark.parametrize("field,value,message", [
    ("a_bh", 1.5, "black_hole.spin must satisfy |spin| <= mass"),
    ("integrator", 7, "integration.integrator must be"),
    ("mode", 3, "simulation.mode must be 'image' or 'disk'"),
    ("camera_width", 0.0, "camera distance, width, and height must be positive"),
    ("albedo", 1.5, "scattering.albedo must be finite and in [0, 1]"),
    ("absorption_coefficient", -1.0, "radiative_transfer.absorption_coefficient"),
    ("superphoton_count_normalization", 0.0, "photons.superphoton_count_normalization"),
    ("density_g_cm3_per_code", -2.0, "all [units] conversion factors"),
])
def test_invalid_params_rejected(field, value, message):
    import grt_amd
    L = grt_amd.load()
    p = grt_amd.default_params(**{field: value})
    h = C.c_void_p()
    rc = L.grt_create(C.byref(p), C.byref(h))
    assert rc == -1 and not h.value
    assert message in L.grt_last_error(None).decode()


def test_power_law_bounds_rejected():
    import grt_amd
    L = grt_amd.load()
    p = grt_amd.default_params(generator=grt_amd.GEN_POWER_LAW, nu_min_hz=1e12, nu_max_hz=1e9)
    h = C.c_void_p()
    assert L.grt_create(C.byref(p), C.byref(h)) == -1
    assert "0 < photons.nu_min_hz < photons.nu_max_hz" in L.grt_last_error(None).decode()


@pytest.mark.skipif(gpu_available(), reason="checks the no-device error path")
def test_create_without_device_fails_cleanly():
    import grt_amd
    with pytest.raises(grt_amd.GrtError, match="HIP device"):
        grt_amd.Context(grt_amd.default_params())


def test_null_arguments_rejected():
    import grt_amd
    L = grt_amd.load()
    assert L.grt_trace(None, None, 0, 1, 1.0, 0, None) == -1
    assert L.grt_set_grid(None, 1, 1, 1, *([None] * 7)) == -1
    h = C.c_void_p()
    assert L.grt_create(None, C.byref(h)) == -1
