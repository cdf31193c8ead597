"""The reference program's flow on one GPU (grt_amd.io.run): TOML config ->
seven-file .npy bundle -> image camera or disk emission -> traced loop ->
archives with the reference's names and members."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bundle(tmp_path):
    from grt_amd.fm_torus import fishbone_moncrief
    f = fishbone_moncrief(nr=32, ntheta=16, nphi=16, r_min=1.5, r_max=60.0, mass=1.0, spin=0.9,
                          r_in=6.0, r_pressure_max=12.0)
    d = tmp_path / "grid"
    d.mkdir()
    for k, v in f.items():
        np.save(d / f"{k}.npy", v)
    return f, str(d)


CONFIG = """
[simulation]
mode = "{mode}"
name = "t"
[input]
numpy_dir = "{grid}"
[black_hole]
mass = 1.0
spin = 0.9
[camera]
distance = 1000.0
theta = 1.0471975511965976
phi = 0.0
width = 50.0
height = 50.0
[photons]
count = 4096
superphoton_count_normalization = {norm}
[integration]
max_steps = 20000
termination_percent = 1.0
initial_step = 0.1
[radiative_transfer]
absorption_enabled = true
emission_enabled = true
absorption_coefficient = 0.05
emission_coefficient = 1.0
[output]
directory = "{out}"
interval = 1000000
image_nx = 64
image_ny = 64
variables = "all"
"""


def test_image_run_writes_reference_archives(gpu, tmp_path):
    import grt_amd.io as io
    f, grid = bundle(tmp_path)
    cfg = tmp_path / "image.toml"
    cfg.write_text(CONFIG.format(mode="image", grid=grid, out=tmp_path / "out", norm=1e8))
    res = io.run(str(cfg))
    assert res["photons"] == 4096 and res["arrived_count"] > 1000
    z = np.load(res["step_archive"])
    assert len(z["x0"]) == 4096 and set(np.unique(z["phase"])) <= {3, 4}
    p = np.load(res["image_products"])
    arrived = z["phase"] == 3
    assert p["arrived_count"].tolist() == [int(arrived.sum())]
    assert p["image_I"].shape == (64, 64)
    assert p["image_I"].sum() == pytest.approx(z["I"][arrived].sum(), rel=1e-12)
    assert os.path.basename(res["step_archive"]).startswith("t_step_")


def test_disk_run(gpu, tmp_path):
    import grt_amd.io as io
    f, grid = bundle(tmp_path)
    cfg = tmp_path / "disk.toml"
    cfg.write_text(CONFIG.format(mode="disk", grid=grid, out=tmp_path / "out", norm=4.0))
    res = io.run(str(cfg))
    expect = int(np.maximum(0, (f["rho"] * 4.0).astype(np.int64)).sum())
    assert res["photons"] == expect > 0
    z = np.load(res["step_archive"])
    assert len(z["id"]) == expect and z["terminate"].all()
    assert "image_products" not in res
