"""Pins the oracle's scattering (scattering.hpp:253-313), Philox RNG, camera
(integrate_cart_ks.hpp:96-130), image initialisation
(initialize_photons.hpp:82-170) and observation products
(observation.hpp:22-84):
- Philox4x32-10 against the Random123 known-answer vectors;
- scattering statistics in the fluid frame (isotropy, energy, null norm,
  Stokes factors, Thomson probability) and the reference's one-cell case
  (test_scattering.cpp:10-116, minus its Kokkos-stream-specific draw);
- flat space (M = 0): a ray crosses the camera plane at the analytic point;
- the reference's image-product case (test_output_stride.cpp:19-45) and the
  (n - 1)-bin rule.
"""
import ctypes as C

import numpy as np


def test_philox_known_answers(oracle):
    # Random123 kat_vectors, philox4x32 with 10 rounds
    assert oracle.philox([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert oracle.philox([0xffffffff] * 4, [0xffffffff] * 2) == \
        [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert oracle.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
                         [0xa4093822, 0x299f31d0]) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_rng_uniform(oracle):
    rng = oracle.Rng(id=42, draw=0, stream=0, seed=1234)
    L = oracle.lib()
    d = np.array([L.orc_rng_drand(C.byref(rng)) for _ in range(20000)])
    assert rng.draw == 20000
    assert d.min() >= 0.0 and d.max() < 1.0
    assert abs(d.mean() - 0.5) < 0.01 and abs(d.var() - 1 / 12) < 0.003


def one_cell(oracle, a):
    # test_scattering.cpp:11-33: r = 10, theta = pi/2, phi = 0, rho = 1, T = 2, u = (1,0,0,0)
    shape = (1, 1, 1)
    vel = np.zeros(shape + (4,))
    vel[..., 0] = 1.0
    f = dict(r=np.full(shape, 10.0), theta=np.full(shape, np.pi / 2), phi=np.zeros(shape),
             rho=np.ones(shape), Tgas=np.full(shape, 2.0), vel=vel, mag=np.zeros(shape + (4,)))
    return oracle.GridHolder(f, a, 1.0)


def test_scattering_statistics(oracle):
    a, m = 1.0, 1.0
    G = one_cell(oracle, a)
    L = oracle.lib()
    pos = np.array([0.0, 10.0, a, 0.0])       # spherical_kerr_schild_to_cartesian(10, pi/2, 0)
    ok, p = oracle.null_momentum(pos, [-1.0, 0.0, 0.0], a, m)
    s0 = np.array([*pos, *p])
    fl = oracle.Fluid()
    e = np.zeros(16)
    assert L.orc_sample_fluid(oracle.dptr(s0), C.byref(G.grid), C.byref(fl), oracle.dptr(e))
    E = e.reshape(4, 4)
    g = oracle.metric(pos, a, m)
    eloc0 = -(s0[4:] @ E[0])
    mus, phis, scattered = [], [], 0
    albedo = 0.5
    n = 4000
    dl = 2.0               # optical depth sigma_T rho/m_p dl ~ 0.8 per step
    tau = 6.6524587158e-25 * (1.0 / 1.67262192369e-24) * dl
    for i in range(n):
        s = s0.copy()
        st = np.array([2.0, 1.0, 1.0, 1.0])
        rng = oracle.Rng(id=i, draw=0, stream=0, seed=42)
        hit = L.orc_maybe_scatter(oracle.dptr(s), oracle.dptr(st), 1, albedo, dl, 1.0,
                                  C.byref(fl), oracle.dptr(e), a, m, C.byref(rng))
        if not hit:
            assert list(st) == [2.0, 1.0, 1.0, 1.0] and np.array_equal(s, s0)
            continue
        scattered += 1
        assert np.allclose(st, [2.0 * albedo, 0.5 * albedo, 0.5 * albedo, 0.25 * albedo])
        gi = oracle.inverse_metric(pos, a, m)
        k = gi @ s[4:]
        assert abs(s[4:] @ k) < 1e-10 * (np.abs(s[4:]) ** 2).sum()
        eloc = -(s[4:] @ E[0])
        assert abs(eloc - eloc0) < 1e-10 * eloc0
        # direction in the fluid frame: n_i = e_i . p / E_loc
        nvec = np.array([E[q] @ s[4:] for q in (1, 2, 3)]) / eloc
        assert abs(np.linalg.norm(nvec) - 1.0) < 1e-9
        mus.append(nvec[2])
        phis.append(np.arctan2(nvec[1], nvec[0]) % (2 * np.pi))
    frac = scattered / n
    prob = 1 - np.exp(-tau)
    assert abs(frac - prob) < 4 * np.sqrt(prob * (1 - prob) / n)
    # isotropy: mu uniform in [-1, 1], phi uniform in [0, 2 pi)  (KS tests)
    from scipy import stats
    assert stats.kstest(mus, "uniform", args=(-1, 2)).pvalue > 1e-3
    assert stats.kstest(phis, "uniform", args=(0, 2 * np.pi)).pvalue > 1e-3
    del g


def test_reference_one_cell_sampling(oracle):
    # test_scattering.cpp:38-44,74: the ray at the single cell samples it
    a = 1.0
    G = one_cell(oracle, a)
    pos = np.array([0.0, 10.0, a, 0.0])
    ok, p = oracle.null_momentum(pos, [-1.0, 0.0, 0.0], a, 1.0)
    s = np.array([*pos, *p])
    fl = oracle.Fluid()
    e = np.zeros(16)
    assert oracle.lib().orc_sample_fluid(oracle.dptr(s), C.byref(G.grid), C.byref(fl),
                                         oracle.dptr(e))
    assert fl.density_g_cm3 == 1.0 and fl.temperature_k == 2.0
    # vacuum grid marked unavailable is never sampled (test_scattering.cpp:110-116)
    G.grid.available = 0
    assert not oracle.lib().orc_sample_fluid(oracle.dptr(s), C.byref(G.grid), C.byref(fl),
                                             oracle.dptr(e))


def test_flat_space_camera_arrival(oracle):
    """M = 0: a forward ray launched along the camera normal crosses the plane
    at its own screen offset, with frequency unchanged (static emitter)."""
    th, ph, d = 1.0, 0.4, 100.0
    p = oracle.make_params(a_bh=0.0, m_bh=0.0, r_term_min=0.0, r_term_max=1e6, mode=0,
                           camera_distance=d, camera_theta=th, camera_phi=ph,
                           camera_width=10.0, camera_height=10.0)
    n = np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
    t = np.array([np.cos(th) * np.cos(ph), np.cos(th) * np.sin(ph), -np.sin(th)])
    f = np.array([-np.sin(ph), np.cos(ph), 0.0])
    u0, v0 = 1.25, -3.5
    x0 = 20.0 * n + u0 * t + v0 * f
    ph_ = oracle.empty_photons(1)
    for q, key in enumerate(("x1", "x2", "x3")):
        ph_[key][0] = x0[q]
    ph_["k0"][0] = -1.0
    ph_["k1"][0], ph_["k2"][0], ph_["k3"][0] = n
    ph_["dlambda"][0] = 1.0
    ph_["phase"][0] = 2              # ImageForward
    ph_["frequency"][0] = 5e11
    ph_["emission_frame_energy"][0] = 1.0
    ph_["I"][0] = 3.0
    st = oracle.trace(p, None, ph_, 100)
    assert ph_["phase"][0] == 3 and ph_["terminate"][0] == 1
    assert abs(ph_["theta_disp"][0] - u0) < 1e-12 and abs(ph_["phi_disp"][0] - v0) < 1e-12
    assert ph_["frequency"][0] == 5e11 and ph_["I"][0] == 3.0
    assert n @ np.array([ph_["x1"][0], ph_["x2"][0], ph_["x3"][0]]) >= d
    assert st["accepted_steps"] >= 2


def test_image_init_geometry(oracle):
    p = oracle.make_params(a_bh=0.9, m_bh=1.0, camera_distance=1000.0, camera_theta=np.pi / 3,
                           camera_phi=0.2)
    su = np.array([-25.0, 0.0, 3.5])
    sv = np.array([10.0, 0.0, -7.0])
    ph = oracle.init_image(p, su, sv, 0.1, id_offset=7)
    th, phi = np.pi / 3, 0.2
    n = np.array([np.sin(th) * np.cos(phi), np.sin(th) * np.sin(phi), np.cos(th)])
    X = np.stack([ph["x1"], ph["x2"], ph["x3"]], axis=1)
    assert np.allclose(X @ n, 1000.0, rtol=1e-14)
    assert list(ph["id"]) == [7, 8, 9] and np.all(ph["phase"] == 1)
    assert np.all(ph["dlambda"] == 100.0)
    for i in range(3):
        gi = oracle.inverse_metric(np.array([0.0, *X[i]]), 0.9, 1.0)
        k = np.array([ph["k0"][i], ph["k1"][i], ph["k2"][i], ph["k3"][i]])
        K = gi @ k
        assert abs(k @ K) < 1e-12 and np.allclose(K[1:], -n, atol=1e-12)


def test_reference_image_products(oracle):
    # test_output_stride.cpp:19-45
    ph = oracle.empty_photons(2)
    ph["frequency"][:] = 1e12
    ph["I"][:] = [2.0, 100.0]
    ph["phase"][:] = [3, 4]
    res = oracle.image_products(ph, 4, 4, 8, 1e9, 1e15, 20.0, 20.0)
    assert res["arrived_count"] == 1 and res["image"][0].sum() == 2.0
    # observation_bin: floor(clamp((v - lo)/(hi - lo)) * (bins - 1))
    B = oracle.lib().orc_observation_bin
    assert B(10.0, -10.0, 10.0, 4) == 3
    assert B(9.9, -10.0, 10.0, 4) == 2       # only u = +w/2 reaches the last bin
    assert B(-50.0, -10.0, 10.0, 4) == 0 and B(0.0, -10.0, 10.0, 1) == 0
    assert B(0.0, -10.0, 10.0, 5) == 2
