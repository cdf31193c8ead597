"""The Fishbone-Moncrief input generator (restating
create_example_data.py:54-248) checked against what the physics fixes:
u.u = -1 in Boyer-Lindquist, constant specific angular momentum
-u_phi/u_t = l(r_pmax) (Bardeen-Press-Teukolsky Keplerian value) inside the
torus, h u_t = const (h from the polytrope), zero angular momentum outside,
peak density 1, the documented grid axes, and the .npy round trip."""
import os

import numpy as np
import pytest


@pytest.fixture(scope="module")
def torus():
    from grt_amd.fm_torus import fishbone_moncrief
    return fishbone_moncrief(nr=48, ntheta=24, nphi=8, r_min=1.5, r_max=100.0, spin=0.9,
                             r_in=6.0, r_pressure_max=12.0)


def bl(r, th, a):
    s2 = np.sin(th) ** 2
    sig = r * r + a * a * np.cos(th) ** 2
    dl = r * r - 2 * r + a * a
    A = (r * r + a * a) ** 2 - dl * a * a * s2
    return -(1 - 2 * r / sig), -2 * r * a * s2 / sig, A * s2 / sig


def test_axes_and_shapes(torus):
    r, th, ph = torus["r"], torus["theta"], torus["phi"]
    assert r.shape == (48, 24, 8) and torus["vel"].shape == (48, 24, 8, 4)
    assert all(v.dtype == np.float64 and v.flags.c_contiguous for v in torus.values())
    assert np.allclose(np.log(r[:, 0, 0]), np.linspace(np.log(1.5), np.log(100.0), 48))
    assert np.allclose(th[0, :, 0], np.linspace(1e-3, np.pi - 1e-3, 24))
    assert np.allclose(ph[0, 0, :], np.arange(8) * 2 * np.pi / 8)
    assert np.all(r == r[:, :1, :1]) and np.all(th == th[:1, :, :1])


def test_four_velocity_physics(torus):
    a = 0.9
    r, th = torus["r"], torus["theta"]
    ut, uph = torus["vel"][..., 0], torus["vel"][..., 3]
    gtt, gtp, gpp = bl(r, th, a)
    norm = gtt * ut * ut + 2 * gtp * ut * uph + gpp * uph * uph
    assert np.allclose(norm, -1.0, atol=1e-10)
    inside = ~((torus["rho"] == 1e-6) & (torus["Tgas"] == 1e-6))   # floor cells
    u_t = gtt * ut + gtp * uph
    u_ph = gtp * ut + gpp * uph
    rp = 12.0
    ell = (rp ** 2 - 2 * a * np.sqrt(rp) + a * a) / (rp ** 1.5 - 2 * np.sqrt(rp) + a)
    assert np.allclose(-u_ph[inside] / u_t[inside], ell, rtol=1e-10)
    out = ~inside                                               # ZAMO atmosphere: u_phi = 0
    assert np.all(np.abs(u_ph[out]) <= 1e-12 * np.abs(gtp[out] * ut[out]) + 1e-15)
    # h u_t = const with h - 1 proportional to rho^(gamma - 1) (polytrope):
    # 1/u_t is an exact affine function of w = rho^(gamma - 1) over the torus
    g = 4.0 / 3.0
    w = torus["rho"][inside] ** (g - 1)
    y = 1.0 / u_t[inside]
    coef = np.polyfit(w, y, 1)
    assert np.abs(np.polyval(coef, w) - y).max() < 1e-10 * np.abs(y).max()
    assert np.all(torus["Tgas"][inside] > 0)
    assert np.allclose(torus["Tgas"][inside], torus["rho"][inside] ** (g - 1), rtol=1e-12)
    assert torus["rho"].max() == 1.0
    assert np.all(torus["rho"][~inside] == 1e-6) and np.all(torus["Tgas"][~inside] == 1e-6)
    assert np.all(torus["mag"] == 0.0)


def test_save_load_roundtrip(tmp_path, torus):
    from grt_amd.fm_torus import load_bundle, save_bundle
    save_bundle(torus, str(tmp_path))
    assert sorted(os.listdir(tmp_path)) == sorted(
        n + ".npy" for n in ("r", "theta", "phi", "rho", "Tgas", "vel", "mag"))
    back = load_bundle(str(tmp_path))
    for k, v in torus.items():
        assert np.array_equal(back[k], v)


def test_invalid_arguments():
    from grt_amd.fm_torus import fishbone_moncrief
    with pytest.raises(ValueError):
        fishbone_moncrief(nr=1)
    with pytest.raises(ValueError):
        fishbone_moncrief(spin=1.5, mass=1.0)
    with pytest.raises(ValueError):
        fishbone_moncrief(r_in=10.0, r_pressure_max=8.0)
