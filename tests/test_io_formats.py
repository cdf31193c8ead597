"""The file formats either side of the path (grt_io.cpp), on the CPU.

Config cases restate the reference's test_simulation_options.cpp:47-120
(keys, defaults, error texts); the .npy and .npz cases write and read with
numpy as the independent codec.
"""
import os

import numpy as np
import pytest


@pytest.fixture(scope="module")
def io():
    import grt_amd.io as io
    return io


def write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_reference_config_case(io, tmp_path):
    # test_simulation_options.cpp:54-90
    path = write(tmp_path, "runtime.toml",
                 '[simulation]\nmode = "disk"\nname = "my-run"\n'
                 '[input]\nvacuum = true\nnumpy_dir = "/data/snapshot"\n'
                 '[photons]\ncount = 100000000000\n'
                 '[integration]\nintegrator = "rk4"\n'
                 '[scattering]\nenabled = true\noptical_depth = 0.75\nalbedo = 0.25\n'
                 'seed = 987654321\n'
                 '[output]\ndirectory = "/tmp/grraytracer-output"\n'
                 'variables = ["id", "frequency"]\n')
    p, run = io.load_config(path)
    assert run["vacuum"] and p.mode == 1 and run["name"] == "my-run"
    assert run["numpy_dir"] == "/data/snapshot"
    assert run["output_dir"] == "/tmp/grraytracer-output"
    assert p.scattering_enabled == 1 and p.albedo == 0.25 and p.seed == 987654321
    assert p.integrator == 0 and run["photon_count"] == 100000000000
    assert run["output_variables"] == "id,frequency"


def test_run_config_dict_round_trip(io, tmp_path):
    """The dict load_config returns becomes the RunConfig the integrate entry
    point takes, field for field; a partial dict keeps the struct's zeros."""
    path = write(tmp_path, "rt.toml",
                 '[simulation]\nname = "rt"\n[integration]\nmax_steps = 123\n'
                 'termination_percent = 0.5\n[output]\ninterval = 7\nstride = 3\n'
                 'directory = "out"\nvariables = "id, x1"\n')
    _, run = io.load_config(path)
    rc = io.run_config_struct(run)
    for name, _ in io.RunConfig._fields_:
        v = getattr(rc, name)
        v = v.decode() if isinstance(v, bytes) else v
        want = run[name]
        assert (bool(v) if name == "vacuum" else v) == want, name
    assert (rc.max_steps, rc.output_interval, rc.output_stride) == (123, 7, 3)
    assert io.run_config_struct(rc) is rc
    part = io.run_config_struct(dict(name="x", max_steps=5))
    assert part.name == b"x" and part.max_steps == 5 and part.output_interval == 0


def test_defaults_and_derived_values(io, tmp_path):
    # utils.cpp:4-37, simulation_options.hpp:13-32, toml_config.hpp:222-223
    path = write(tmp_path, "d.toml", "# only a comment\n[black_hole]\nmass = 2.0\nspin = 1.2\n"
                                     "[camera]\ndistance = 500 # trailing comment\n")
    p, run = io.load_config(path)
    assert p.m_bh == 2.0 and p.a_bh == 1.2
    assert p.r_term_min == pytest.approx(2.0 + np.sqrt(4.0 - 1.44))
    assert p.r_term_max == 750.0
    assert p.camera_theta == pytest.approx(np.pi / 4) and p.camera_phi == pytest.approx(np.pi / 4)
    assert p.camera_width == 20.0 and p.camera_height == 20.0
    assert run["photon_count"] == 20 and run["max_steps"] == 10000
    assert run["termination_percent"] == 0.9 and run["initial_step"] == 0.1
    assert p.atol == 1e-8 and p.rtol == 1e-8 and p.integrator == 1 and p.mode == 0
    assert run["output_interval"] == 1 and run["output_stride"] == 1
    assert (run["image_nx"], run["image_ny"], run["spectrum_bins"]) == (256, 256, 128)
    assert run["spectrum_min_hz"] == 1e8 and run["spectrum_max_hz"] == 1e22
    assert run["name"] == "output" and run["output_dir"] == "./output"
    assert run["output_variables"] == "id,frequency,x0,x1,x2,x3,k0,k1,k2,k3,I,Q,U,V,phase"
    assert p.superphoton_count_normalization == 1.0 and p.generator == 0
    assert p.custom_frequency_hz == 2.30e11 and p.power_law_slope == 2.0
    assert p.scattering_enabled == 0 and p.albedo == 1.0 and p.seed == 12345


@pytest.mark.parametrize("text,message", [
    ("[photons]\ncount = 100000000001\n", "photons.count must be in [1, 100000000000]"),
    ("[input]\nvacuum = yes\n", "must be true or false"),
    ("[simulation]\nmode = \"pinhole\"\n", "must be 'image' or 'disk'"),
    ("[integration]\nintegrator = \"euler\"\n", "integration.integrator must be"),
    ("[black_hole]\nmass = 1.0\nspin = 1.5\n", "naked singularity"),
    ("[black_hole]\nmass = -1\n", "black_hole.mass must be positive"),
    ("[photons]\ngenerator = \"laser\"\n", "photons.generator must be"),
    ("[output]\ninterval = 0\n", "output.interval must be positive"),
    ("[output]\nstride = -2\n", "must not be negative"),
    ("[output]\nvariables = [\"id\", \"bogus\"]\n", "unknown output variable: bogus"),
    ("[output]\nspectrum_min_hz = 10\nspectrum_max_hz = 5\n", "0 < spectrum_min_hz < spectrum_max_hz"),
    ("[scattering]\nalbedo = 1.5\n", "scattering.albedo must be finite and in [0, 1]"),
    ("[integration]\ntermination_percent = 0\n", "termination_percent must be in (0, 1]"),
    ("[camera]\nwidth = 0\n", "camera distance, width, and height must be positive"),
    ("no equals sign here\n", "Malformed config line"),
])
def test_config_errors(io, tmp_path, text, message):
    import grt_amd
    path = write(tmp_path, "bad.toml", text)
    with pytest.raises(grt_amd.GrtError, match=None) as e:
        io.load_config(path)
    assert message in str(e.value)


def test_missing_config_file(io, tmp_path):
    import grt_amd
    with pytest.raises(grt_amd.GrtError, match="Failed to open config file"):
        io.load_config(str(tmp_path / "absent.toml"))


def test_npy_reader(io, tmp_path):
    import grt_amd
    a = np.random.default_rng(0).normal(size=(3, 4, 5, 4))
    np.save(tmp_path / "a.npy", a)
    assert np.array_equal(io.read_npy(str(tmp_path / "a.npy")), a)
    np.save(tmp_path / "f.npy", np.asfortranarray(a[..., 0]))
    np.save(tmp_path / "i.npy", np.arange(6))
    with pytest.raises(grt_amd.GrtError, match="float64"):
        io.read_npy(str(tmp_path / "i.npy"))
    b = np.asfortranarray(np.ones((3, 4)))
    np.save(tmp_path / "fo.npy", b)
    with pytest.raises(grt_amd.GrtError, match="C-ordered"):
        io.read_npy(str(tmp_path / "fo.npy"))


def test_numpy_bundle_discovery(io, tmp_path):
    # test_simulation_options.cpp:105-120 and load_python_arrays.hpp:82-180
    import grt_amd
    with pytest.raises(grt_amd.GrtError, match="does not exist"):
        io.load_numpy_fields(None, str(tmp_path / "absent"))
    grid = tmp_path / "grid"
    grid.mkdir()
    np.save(grid / "r.npy", np.ones((2, 3, 4)))
    with pytest.raises(grt_amd.GrtError, match="theta.npy"):
        io.load_numpy_fields(None, str(grid))
    for name in ("theta", "phi", "rho", "Tgas"):
        np.save(grid / f"{name}.npy", np.ones((2, 3, 4)))
    np.save(grid / "vel.npy", np.ones((2, 3, 4, 4)))
    np.save(grid / "mag.npy", np.ones((2, 3, 4, 3)))
    with pytest.raises(grt_amd.GrtError, match="four components"):
        io.load_numpy_fields(None, str(grid))
    np.save(grid / "mag.npy", np.ones((2, 3, 4, 4)))
    np.save(grid / "rho.npy", np.ones((2, 3, 5)))
    with pytest.raises(grt_amd.GrtError, match="incompatible dimensions"):
        io.load_numpy_fields(None, str(grid))
    np.save(grid / "rho.npy", np.ones((2, 3, 4)))
    io.load_numpy_fields(None, str(grid))


def test_image_products_archive(io, tmp_path):
    rng = np.random.default_rng(1)
    prod = dict(image=rng.normal(size=(4, 3, 5)), spectrum_frequency_hz=np.logspace(8, 22, 7),
                spectrum_I=rng.normal(size=7), arrived_count=42)
    path = str(tmp_path / "out" / "run_image_products.npz")
    io.write_image_products_npz(path, prod)
    z = np.load(path)
    assert sorted(z.files) == sorted(["image_I", "image_Q", "image_U", "image_V",
                                      "spectrum_frequency_hz", "spectrum_I", "arrived_count"])
    for q, k in enumerate("IQUV"):
        assert np.array_equal(z[f"image_{k}"], prod["image"][q])
    assert np.array_equal(z["spectrum_I"], prod["spectrum_I"])
    assert z["arrived_count"].dtype == np.uint64 and z["arrived_count"].tolist() == [42]
    assert not os.path.exists(path + ".tmp")


def test_photon_archive_selection_and_stride(io, tmp_path):
    import grt_amd
    n = 11
    ph = grt_amd.empty_photons(n)
    rng = np.random.default_rng(2)
    for name, dt in grt_amd.PHOTON_FIELDS:
        if dt == np.float64:
            ph[name][:] = rng.normal(size=n)
    ph["id"][:] = np.arange(100, 100 + n, dtype=np.uint64)
    ph["terminate"][:] = rng.integers(0, 2, n)
    ph["phase"][:] = rng.integers(0, 5, n)
    path = str(tmp_path / "run_step_7_rank0.npz")
    io.write_photons_npz(path, ph, "[\"id\", 'terminate', x1, all]", 3)
    z = np.load(path)
    assert sorted(z.files) == sorted(name for name, _ in grt_amd.PHOTON_FIELDS)
    for name, _ in grt_amd.PHOTON_FIELDS:
        assert np.array_equal(z[name], ph[name][::3]), name
    io.write_photons_npz(path, ph, None, 1)
    z = np.load(path)
    assert z.files == ["id", "frequency", "x0", "x1", "x2", "x3", "k0", "k1", "k2", "k3",
                       "I", "Q", "U", "V", "phase"]
