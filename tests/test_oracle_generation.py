"""Pins of the oracle's photon generation and disk emission against what the
mathematics and the reference's own test fix.

* Thermal generator (orc_generate_photon, photon_generation.hpp:101-124):
  x = h nu / k T drawn from the Planck law 15/pi^4 x^3/(e^x - 1) (KS test
  on 1e6 draws against the closed-form CDF, mean pi^4 / (30 zeta(3))), with
  the exact SI value of k/h.
* Power-law generator (:128-143): KS against the analytic CDF of
  nu^-slope on [nu_min, nu_max] for slope 2.5 and the log-uniform slope 1.
* Disk emission (orc_init_disk, initialize_photons.hpp:32-62, :173-327):
  the reference's test_photon_generation.cpp:11-98 case (one cell, 64
  packets, ids from 5e10, null momenta, sum of I = 8, power-law frequencies
  in range); the packet-to-cell assignment against a brute-force expansion
  of the per-cell counts on a multi-cell torus; positions against the
  oblate Kerr-Schild map inverted with the oracle's Boyer-Lindquist radius;
  fluid-frame isotropy of the emitted directions.
* The spherical-to-Cartesian Jacobian (fluid_frame.hpp:29-49) against
  central differences of the position map (:16-27).

A plausible mistake -- 1/n^3 mixture weights, a Gamma(3) stage, a wrong k/h,
the power-law exponent off by one, mu drawn on [0, 1], a sign slip in the
azimuthal column of the Jacobian, a packet assigned to the next cell --
fails one of these.
"""
import ctypes as C

import numpy as np
import pytest
from scipy import stats

from conftest import c1_fields

K_OVER_H = 1.380649e-23 / 6.62607015e-34     # SI-exact constants, Hz / K


def planck_cdf(x):
    """15/pi^4 int_0^x t^3 / (e^t - 1) dt."""
    x = np.asarray(x, dtype=np.float64)
    out = np.empty_like(x)
    small = x < 0.05
    xs = x[small]
    out[small] = 15 / np.pi ** 4 * (xs ** 3 / 3 - xs ** 4 / 8 + xs ** 5 / 60 - xs ** 7 / 5040)
    xl = x[~small][:, None]
    n = np.arange(1, 801, dtype=np.float64)[None, :]
    terms = np.exp(-n * xl) * (xl ** 3 / n + 3 * xl ** 2 / n ** 2 + 6 * xl / n ** 3 + 6 / n ** 4)
    out[~small] = 1.0 - 15 / np.pi ** 4 * terms.sum(axis=1)
    return out


def _cdf_table():
    grid = np.concatenate([np.linspace(0.0, 0.05, 200, endpoint=False),
                           np.linspace(0.05, 80.0, 40000)])
    return grid, planck_cdf(grid)


def test_planck_cdf_helper():
    """The test's own CDF: monotone, 0 -> 0, inf -> 1, and both branches meet."""
    g, F = _cdf_table()
    assert F[0] == 0.0 and abs(F[-1] - 1.0) < 1e-15 and np.all(np.diff(F) >= 0)
    assert planck_cdf(np.array([0.0499999]))[0] == pytest.approx(
        planck_cdf(np.array([0.0500001]))[0], abs=1e-9)


def test_thermal_generator_draws_the_planck_law(oracle):
    T = 3.7e9
    p = oracle.make_params(generator=0, energy_per_cell_erg=6.0)
    fl = dict(density_g_cm3=1.0, temperature_k=T)
    nu, en = oracle.sample_generator(p, fl, 3, 1_000_000)
    assert np.all(nu > 0) and np.all(np.isfinite(nu))
    assert np.all(en == 2.0)                       # E_cell / packets
    x = nu / (K_OVER_H * T)
    g, F = _cdf_table()
    ks = stats.kstest(x, lambda v: np.interp(v, g, F))
    print("Planck KS", ks)
    assert ks.pvalue > 1e-3
    # Moments of the energy spectrum x^3/(e^x - 1) (equal-energy packets):
    # <x^k> = Gamma(4+k) zeta(4+k) / (Gamma(4) zeta(4)), so <x> = 4 zeta(5)/zeta(4)
    # = 3.832 and <x^2> = 20 zeta(6)/zeta(4).  (pi^4 / (30 zeta(3)) = 2.701 is
    # the mean of the number spectrum x^2/(e^x - 1), which a Gamma(3) mixture
    # with 1/n^3 weights would draw.)
    zeta4, zeta5, zeta6 = np.pi ** 4 / 90, 1.0369277551433699, np.pi ** 6 / 945
    sem = x.std() / np.sqrt(len(x))
    assert abs(x.mean() - 4 * zeta5 / zeta4) < 5 * sem
    assert abs(x.mean() - np.pi ** 4 / (30 * 1.2020569031595942)) > 100 * sem
    assert np.mean(x ** 2) == pytest.approx(20 * zeta6 / zeta4, rel=5e-3)
    # a non-positive temperature emits nothing
    nu0, _ = oracle.sample_generator(p, dict(density_g_cm3=1.0, temperature_k=0.0), 3, 10)
    assert np.all(nu0 == 0.0)


@pytest.mark.parametrize("slope", [2.5, 1.0, 1.0 + 5e-13, -0.5])
def test_power_law_generator_cdf(oracle, slope):
    lo, hi = 1e9, 1e12
    p = oracle.make_params(generator=1, power_law_slope=slope, nu_min_hz=lo, nu_max_hz=hi,
                           energy_per_cell_erg=5.0)
    nu, en = oracle.sample_generator(p, dict(density_g_cm3=1.0, temperature_k=1.0), 4, 300_000)
    assert np.all((nu >= lo) & (nu <= hi)) and np.all(en == 1.25)
    if abs(slope - 1.0) < 1e-9:
        cdf = lambda v: np.log(v / lo) / np.log(hi / lo)
    else:
        e = 1.0 - slope
        cdf = lambda v: (v ** e - lo ** e) / (hi ** e - lo ** e)
    ks = stats.kstest(nu, cdf)
    print("power law", slope, ks)
    assert ks.pvalue > 1e-3


def test_custom_generator(oracle):
    p = oracle.make_params(generator=2, custom_frequency_hz=4.2e11, energy_per_cell_erg=3.0)
    nu, en = oracle.sample_generator(p, dict(density_g_cm3=1.0, temperature_k=1.0), 6, 5)
    assert np.all(nu == 4.2e11) and np.all(en == 0.5)


def one_cell(rho=2.0):
    f = {k: np.full((1, 1, 1), v) for k, v in
         dict(r=10.0, theta=np.pi / 2, phi=0.25, rho=rho, Tgas=1.0).items()}
    vel = np.zeros((1, 1, 1, 4))
    vel[..., 0] = 1.0
    f["vel"] = vel
    f["mag"] = np.zeros((1, 1, 1, 4))
    return f


def test_reference_photon_generation_case(oracle):
    """test_photon_generation.cpp:11-98: one cell at r = 10, theta = pi/2,
    phi = 0.25 with rho = 2, T = 1e10 K, u = (1, 0, 0, 0); normalisation 32
    emits 64 packets, energy 8 per cell, ids from 5e10; a = M = 1 (the
    reference's globals, utils.cpp:10-11)."""
    fields = one_cell()
    G = oracle.GridHolder(fields, 1.0, 1.0, units=dict(temperature_k_per_code=1e10))
    p = oracle.make_params(a_bh=1.0, m_bh=1.0, generator=0, superphoton_count_normalization=32.0,
                           energy_per_cell_erg=8.0)
    assert oracle.disk_count(p, G) == 64
    off = 50_000_000_000
    ph = oracle.init_disk(p, G, 0.1, id_offset=off)
    assert np.array_equal(ph["id"], off + np.arange(64, dtype=np.uint64))
    assert np.all(ph["frequency"] > 0) and np.all(np.isfinite(ph["frequency"]))
    assert not ph["terminate"].any() and np.all(ph["phase"] == 0)
    assert abs(ph["I"].sum() - 8.0) < 1e-12
    for i in range(64):
        s = np.array([ph[k][i] for k in ("x0", "x1", "x2", "x3", "k0", "k1", "k2", "k3")])
        assert abs(oracle.observables(s, 1.0, 1.0)["norm"]) < 1e-10
    # power law, ids from 1000: every frequency in [nu_min, nu_max]
    p = oracle.make_params(a_bh=1.0, m_bh=1.0, generator=1, superphoton_count_normalization=32.0,
                           energy_per_cell_erg=8.0, nu_min_hz=1e9, nu_max_hz=1e12,
                           power_law_slope=2.5)
    ph = oracle.init_disk(p, G, 0.1, id_offset=1000)
    assert np.all((ph["frequency"] >= 1e9) & (ph["frequency"] <= 1e12))
    assert ph["id"][0] == 1000


@pytest.fixture(scope="module")
def torus():
    return c1_fields(nr=24, nth=12, nph=12)


def test_packets_follow_the_prefix_sum(oracle, torus):
    """Packet i belongs to the cell whose prefix-sum range holds i: the
    brute-force expansion of max(0, int(rho * norm)) per cell in C order.
    Each packet sits at its cell's Kerr-Schild position and carries
    E_cell / packets of that cell."""
    a, norm, ecell = 0.9, 7.0, 3.0
    G = oracle.GridHolder(torus, a, 1.0)
    p = oracle.make_params(a_bh=a, m_bh=1.0, superphoton_count_normalization=norm,
                           energy_per_cell_erg=ecell, mode=1)
    counts = np.maximum(0, np.trunc(torus["rho"].ravel() * norm)).astype(np.int64)
    assert counts.sum() > 500 and (counts > 0).sum() > 20
    cell = np.repeat(np.arange(counts.size), counts)
    ph = oracle.init_disk(p, G, 0.1)
    assert len(ph["x0"]) == counts.sum() == oracle.disk_count(p, G)
    r = torus["r"].ravel()[cell]
    th = torus["theta"].ravel()[cell]
    phi = torus["phi"].ravel()[cell]
    x = (r * np.cos(phi) - a * np.sin(phi)) * np.sin(th)
    y = (r * np.sin(phi) + a * np.cos(phi)) * np.sin(th)
    z = r * np.cos(th)
    for got, want in ((ph["x1"], x), (ph["x2"], y), (ph["x3"], z)):
        assert np.allclose(got, want, rtol=0, atol=1e-13 * r.max())
    ok = ph["phase"] == 0
    assert ok.all()
    assert np.allclose(ph["I"], ecell / counts[cell], rtol=1e-15)
    # the oblate map inverts through the oracle's Boyer-Lindquist radius
    rr = np.array([oracle.lib().orc_compute_r(a1, b1, c1, a) for a1, b1, c1 in
                   zip(ph["x1"], ph["x2"], ph["x3"])])
    assert np.allclose(rr, r, rtol=1e-13)
    assert np.allclose(ph["x3"] / rr, np.cos(th), atol=1e-13)
    az = np.mod(np.arctan2(rr * ph["x2"] - a * ph["x1"], rr * ph["x1"] + a * ph["x2"]),
                2 * np.pi)
    dphi = np.abs(np.angle(np.exp(1j * (az - phi))))
    assert dphi.max() < 1e-12


def test_emission_is_isotropic_in_the_fluid_frame(oracle, torus):
    """Each packet's direction in its cell's fluid tetrad, n_i = (p.e_i) /
    (-p.e_0), is a unit vector with cos(polar angle) uniform on [-1, 1] and
    azimuth uniform on [0, 2 pi) (KS), and -p.e_0 equals the stored
    emission-frame energy."""
    a = 0.9
    G = oracle.GridHolder(torus, a, 1.0)
    p = oracle.make_params(a_bh=a, m_bh=1.0, superphoton_count_normalization=60.0, mode=1)
    ph = oracle.init_disk(p, G, 0.1)
    n = len(ph["x0"])
    assert n > 5000
    counts = np.maximum(0, np.trunc(torus["rho"].ravel() * 60.0)).astype(np.int64)
    cell = np.repeat(np.arange(counts.size), counts)
    L = oracle.lib()
    mu = np.empty(n)
    az = np.empty(n)
    for i in range(n):
        c = cell[i]
        r, th, phc = (torus[k].ravel()[c] for k in ("r", "theta", "phi"))
        u = np.zeros(4)
        L.orc_spherical_to_cartesian_vector(r, th, phc, a,
                                            oracle.dptr(torus["vel"].reshape(-1, 4)[c].copy()),
                                            oracle.dptr(u))
        X = np.array([ph["x0"][i], ph["x1"][i], ph["x2"][i], ph["x3"][i]])
        e = np.zeros(16)
        assert L.orc_build_fluid_tetrad(oracle.dptr(X), oracle.dptr(u), a, 1.0, oracle.dptr(e))
        e = e.reshape(4, 4)
        pc = np.array([ph[k][i] for k in ("k0", "k1", "k2", "k3")])
        proj = e @ pc
        E = -proj[0]
        assert E == pytest.approx(ph["emission_frame_energy"][i], rel=1e-12)
        nv = proj[1:] / E
        assert np.linalg.norm(nv) == pytest.approx(1.0, abs=1e-10)
        mu[i] = nv[2]
        az[i] = np.mod(np.arctan2(nv[1], nv[0]), 2 * np.pi)
    assert stats.kstest(mu, stats.uniform(-1, 2).cdf).pvalue > 1e-3
    assert stats.kstest(az, stats.uniform(0, 2 * np.pi).cdf).pvalue > 1e-3


def test_spherical_to_cartesian_jacobian_is_the_derivative_of_the_position_map(oracle):
    """c = J s with J = d(x, y, z)/d(r, theta, phi) of the oblate Kerr-Schild
    position map x = (r cos phi - a sin phi) sin theta, y = (r sin phi +
    a cos phi) sin theta, z = r cos theta; J from central differences."""
    rng = np.random.default_rng(11)

    def position(q, a):
        r, th, ph = q
        return np.array([(r * np.cos(ph) - a * np.sin(ph)) * np.sin(th),
                         (r * np.sin(ph) + a * np.cos(ph)) * np.sin(th), r * np.cos(th)])

    L = oracle.lib()
    for _ in range(200):
        a = rng.uniform(-1, 1)
        q = np.array([rng.uniform(1.5, 80), rng.uniform(0.01, np.pi - 0.01),
                      rng.uniform(0, 2 * np.pi)])
        s = rng.normal(size=4)
        J = np.zeros((3, 3))
        for j in range(3):
            h = 1e-5 * (q[0] if j == 0 else 1.0)
            dq = np.zeros(3)
            dq[j] = h
            J[:, j] = (position(q + dq, a) - position(q - dq, a)) / (2 * h)
        c = np.zeros(4)
        L.orc_spherical_to_cartesian_vector(q[0], q[1], q[2], a, oracle.dptr(s.copy()),
                                            oracle.dptr(c))
        assert c[0] == s[0]
        assert np.allclose(c[1:], J @ s[1:], rtol=0, atol=1e-8 * (1 + np.abs(J).max() *
                                                                  np.abs(s).max()))
