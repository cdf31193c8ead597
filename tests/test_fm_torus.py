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
    with pytest.raises(ValueError):
        fishbone_moncrief(r_min=1.0, spin=0.9)           # inside the horizon
    with pytest.raises(ValueError):
        fishbone_moncrief(gamma=1.0)


def test_equals_the_reference_generator():
    """The reference's own generator (create_example_data.py,
    create_fishbone_moncrief_torus) on four small grids -- mass 2, no floor,
    gamma 5/3, an open equipotential, Schwarzschild with the default r_min --
    stored by tests/golden/make_fm_torus_fixtures.py: the same torus cells,
    and every value to the rounding of the two derivations (the density
    carries the cancellation h - 1 near the torus edge)."""
    from grt_amd.fm_torus import fishbone_moncrief
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    from make_fm_torus_fixtures import CASES, FIELDS
    ref = np.load(os.path.join(os.path.dirname(__file__), "golden", "fm_torus_reference.npz"))
    tol = dict(r=1e-15, theta=0.0, phi=0.0, rho=1e-11, Tgas=1e-11, vel=1e-14, mag=0.0)
    for name, kw in CASES.items():
        got = fishbone_moncrief(**kw)
        floor = kw.get("rho_floor", 1e-6)
        assert np.array_equal(got["rho"] == floor, ref[f"{name}/rho"] == floor), name
        for k in FIELDS:
            want = ref[f"{name}/{k}"]
            assert got[k].shape == want.shape, (name, k)
            assert np.allclose(got[k], want, rtol=tol[k], atol=0.0), (name, k)


def test_two_derivations_of_the_keplerian_angular_momentum():
    """-u_phi/u_t of the prograde circular orbit from its angular velocity
    1/(r^1.5 + a) equals the closed form of Bardeen, Press & Teukolsky
    (1972, eqs. 2.12-2.13), L/E = (r^2 - 2 a sqrt r + a^2) / (r^1.5 - 2 sqrt r + a)."""
    from grt_amd.fm_torus import keplerian_angular_momentum
    for a in (0.0, 0.5, 0.9, 0.94, 0.998):
        for r in (4.0, 6.0, 10.0, 12.0, 30.0):
            bpt = (r * r - 2 * a * np.sqrt(r) + a * a) / (r ** 1.5 - 2 * np.sqrt(r) + a)
            assert keplerian_angular_momentum(r, a) == pytest.approx(bpt, rel=1e-13)


@pytest.mark.parametrize("spin,mass", [(0.9, 1.0), (0.5, 1.0), (1.88, 2.0)])
def test_kerr_schild_four_velocity_is_normalised(oracle, spin, mass):
    """The stored (u^t, 0, 0, u^phi) taken to Cartesian Kerr-Schild components
    by the oracle's spherical-to-Cartesian Jacobian (fluid_frame.hpp:29-49)
    and the position map (:16-27) is a unit timelike vector of the
    Kerr-Schild metric, g(u, u) = -1, in and outside the torus."""
    import ctypes as C
    from grt_amd.fm_torus import fishbone_moncrief
    f = fishbone_moncrief(nr=20, ntheta=12, nphi=6, r_min=1.3 * mass + 1.0, r_max=80.0 * mass,
                          mass=mass, spin=spin, r_in=6.0 * mass, r_pressure_max=12.0 * mass)
    L = oracle.lib()
    worst = 0.0
    for idx in np.ndindex(f["r"].shape):
        r, th, ph = f["r"][idx], f["theta"][idx], f["phi"][idx]
        X = np.array([0.0, (r * np.cos(ph) - spin * np.sin(ph)) * np.sin(th),
                      (r * np.sin(ph) + spin * np.cos(ph)) * np.sin(th), r * np.cos(th)])
        u = np.zeros(4)
        L.orc_spherical_to_cartesian_vector(r, th, ph, spin, oracle.dptr(f["vel"][idx].copy()),
                                            oracle.dptr(u))
        g = oracle.metric(X, spin, mass)
        worst = max(worst, abs(u @ g @ u + 1.0))
    assert worst < 1e-12
