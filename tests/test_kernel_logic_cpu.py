"""The tracing kernel's per-lane logic (grt_device.hpp, compiled for the CPU
by tests/harness) against the oracle on the first reference configuration
and on variants exercising every branch of the functor: RK4, the general
(non-separable) cell search, image mode with scattering enabled (where it
never fires), disk-mode packets with multiple scattering (statistical),
horizon capture, 12-failure termination.

Compiled for the CPU (the only fused operations are the explicit fma calls
that the oracle makes too) the kernel logic reproduces the oracle to the last
bit for positions, momenta, Stokes and step sizes; the fluid tetrad
uses the factored metric, so the emission-frame energy (and the redshifted
frequency) agrees to ~1e-15 relative.
"""
import numpy as np
import pytest

from conftest import c1_fields, c1_params, pixel_screen

STATE = ["x0", "x1", "x2", "x3", "k0", "k1", "k2", "k3", "I", "Q", "U", "V", "dlambda",
         "theta_disp", "phi_disp"]


UNIT_KEYS = ("length_cm_per_code", "time_s_per_code", "density_g_cm3_per_code",
             "temperature_k_per_code", "four_velocity_cm_s_per_code", "magnetic_gauss_per_code")


def split_units(kw):
    """(oracle params kwargs, oracle grid units) from a grt_params kwargs dict."""
    return ({k: v for k, v in kw.items() if k not in UNIT_KEYS},
            {k: v for k, v in kw.items() if k in UNIT_KEYS})


def run(oracle, fields, kw, ph0, max_steps=10000, force_general=False):
    import grt_amd
    import harness
    A = {k: v.copy() for k, v in ph0.items()}
    B = {k: v.copy() for k, v in ph0.items()}
    okw, units = split_units(kw)
    G = oracle.GridHolder(fields, kw["a_bh"], kw["m_bh"], units) if fields is not None else None
    st_o = oracle.trace(oracle.make_params(**okw), G, A, max_steps, 1.0, 0, 4)
    st_h = harness.trace(grt_amd.default_params(**kw), fields, B, max_steps, force_general)
    return A, B, st_o, st_h


def assert_same(A, B, st_o, st_h, freq_tol=1e-14):
    for k in STATE:
        assert np.array_equal(A[k], B[k], equal_nan=True), k
    assert np.array_equal(A["phase"], B["phase"])
    assert np.array_equal(A["terminate"], B["terminate"])
    f = np.abs(A["frequency"] - B["frequency"]) / np.maximum(np.abs(A["frequency"]), 1e-300)
    assert f.max() <= freq_tol
    for k in ("functor_calls", "accepted_steps", "rejected_attempts", "failed_steps",
              "fluid_samples", "fluid_hits"):
        assert st_o[k] == st_h[k], k


@pytest.fixture(scope="module")
def c1():
    return c1_fields()


def test_c1_image_rk45(oracle, c1):
    kw = c1_params()
    su, sv = pixel_screen(48)
    ph0 = oracle.init_image(oracle.make_params(**kw), su, sv, 0.1)
    A, B, st_o, st_h = run(oracle, c1, kw, ph0)
    assert st_o["accepted_steps"] > 20 * len(su)
    assert_same(A, B, st_o, st_h)
    assert np.all(A["phase"] == 3)


def test_c1_image_general_search(oracle, c1):
    kw = c1_params()
    su, sv = pixel_screen(12)
    ph0 = oracle.init_image(oracle.make_params(**kw), su, sv, 0.1)
    A, B, st_o, st_h = run(oracle, c1, kw, ph0, force_general=True)
    assert_same(A, B, st_o, st_h)


def test_torus_image_rk4_and_horizon(oracle):
    """RK4 with a fixed step through a floor-free torus: rays are captured by
    the horizon, escape, or emit; every outcome must match."""
    from grt_amd.fm_torus import fishbone_moncrief
    fields = fishbone_moncrief(nr=64, ntheta=32, nphi=32, r_min=1.5, r_max=100.0, spin=0.9,
                               r_in=6.0, r_pressure_max=9.0, rho_floor=0.0)
    kw = c1_params(theta=np.pi / 4, integrator=0, camera_width=30.0, camera_height=30.0)
    su, sv = pixel_screen(24, width=30.0, height=30.0)
    ph0 = oracle.init_image(oracle.make_params(**kw), su, sv, 0.002)
    A, B, st_o, st_h = run(oracle, fields, kw, ph0, max_steps=3000)
    assert_same(A, B, st_o, st_h)
    counts = np.bincount(A["phase"], minlength=5)
    assert counts[4] > 0 and counts[3] > 0


def test_image_mode_scattering_never_fires(oracle, c1):
    """In image mode the forward trace after the emission event retraces the
    backward path, which met no fluid before the emission point, so the
    forward phase never samples fluid: transfer and scattering after emission
    never fire (integrate_cart_ks.hpp:323-376), whatever the optical depth.
    Kernel logic and oracle agree bit for bit, and the albedo changes
    nothing."""
    out = []
    for albedo in (0.7, 0.2):
        kw = c1_params(scattering_enabled=1, albedo=albedo, seed=99,
                       length_cm_per_code=1e24)   # optical depth >> 1 per step in fluid
        su, sv = pixel_screen(24)
        ph0 = oracle.init_image(oracle.make_params(**split_units(kw)[0]), su, sv, 0.1)
        A, B, st_o, st_h = run(oracle, c1, kw, ph0, max_steps=2000)
        assert_same(A, B, st_o, st_h)
        out.append(A)
    for k in ("k1", "k2", "k3", "I", "phase"):
        assert np.array_equal(out[0][k], out[1][k]), k


def ergoregion_camera_params():
    """A camera inside the ergoregion (equatorial, r = 1.8 < 2 M for a = 0.9):
    at most arrival points g_00 = 2H - 1 > 0, so camera_observer_energy is 0
    and the arrival zeroes the Stokes vector (integrate_cart_ks.hpp:122-130,
    392-400); slightly off the plane some arrivals stay valid."""
    return c1_params(theta=np.pi / 2, camera_distance=1.8, camera_width=2.0,
                     camera_height=2.0, r_term_max=300.0)


def test_arrival_inside_the_ergoregion(oracle, c1):
    """The kernel completes a deferred photon's arrival after the trace (its
    redshift factor, or the zeroing of an invalid one, from the stored
    state); the oracle does both at the arrival.  Both agree bit for bit,
    on arrivals of both kinds."""
    kw = ergoregion_camera_params()
    su, sv = pixel_screen(16, width=2.0, height=2.0)
    ph0 = oracle.init_image(oracle.make_params(**split_units(kw)[0]), su, sv, 0.01)
    A, B, st_o, st_h = run(oracle, c1, kw, ph0, max_steps=5000)
    assert_same(A, B, st_o, st_h)
    arrived = A["phase"] == 3
    assert (arrived & (A["I"] == 0)).sum() > 100 and (arrived & (A["I"] > 0)).sum() > 20


def disk_packets(oracle, fields, kw):
    okw, units = split_units(kw)
    G = oracle.GridHolder(fields, kw["a_bh"], 1.0, units)
    return oracle.init_disk(oracle.make_params(**okw), G, 0.1)


def test_disk_mode_packets(oracle):
    fields = c1_fields(nr=16, nth=8, nph=8)
    kw = c1_params(mode=1, superphoton_count_normalization=3.0, generator=1, nu_min_hz=1e9,
                   nu_max_hz=1e12)
    ph0 = disk_packets(oracle, fields, kw)
    assert len(ph0["x0"]) > 100
    A, B, st_o, st_h = run(oracle, fields, kw, ph0, max_steps=3000)
    assert_same(A, B, st_o, st_h)


def test_disk_mode_multiple_scattering(oracle):
    """With repeated scattering, last-bit differences of the fluid tetrad (the
    kernel's factored metric) are amplified along each packet's path, so the
    comparison is statistical: same termination pattern and the same
    distribution of escape directions and intensities."""
    fields = c1_fields(nr=16, nth=8, nph=8)
    kw = c1_params(mode=1, superphoton_count_normalization=6.0, scattering_enabled=1,
                   albedo=0.9, length_cm_per_code=1e23, generator=2)
    ph0 = disk_packets(oracle, fields, kw)
    A, B, st_o, st_h = run(oracle, fields, kw, ph0, max_steps=3000)
    n = len(A["x0"])
    assert n > 300
    assert np.mean(A["phase"] == B["phase"]) > 0.98
    from scipy import stats
    for f in (lambda P: np.arctan2(P["x2"], P["x1"]),
              lambda P: P["x3"] / np.sqrt(P["x1"] ** 2 + P["x2"] ** 2 + P["x3"] ** 2),
              lambda P: P["I"]):
        assert stats.ks_2samp(f(A), f(B)).pvalue > 0.01
    assert abs(st_o["accepted_steps"] - st_h["accepted_steps"]) < 0.02 * st_o["accepted_steps"]


@pytest.mark.parametrize("generator", [1, 2])
def test_image_mode_generators(oracle, c1, generator):
    """Image mode with the power-law and the custom generator
    (photon_generation.hpp:99-156): the first reads only the emission-frame
    energy of the fluid tetrad, the second the whole sampled cell."""
    kw = c1_params(generator=generator, nu_min_hz=1e9, nu_max_hz=1e12)
    su, sv = pixel_screen(12)
    ph0 = oracle.init_image(oracle.make_params(**kw), su, sv, 0.1)
    A, B, st_o, st_h = run(oracle, c1, kw, ph0, max_steps=3000)
    assert_same(A, B, st_o, st_h)
    assert (A["phase"] == 3).sum() > 10 and np.count_nonzero(A["frequency"]) > 10


def test_twelve_failures_terminate(oracle, c1):
    """Twelve rejected attempts in one functor call terminate the photon with
    its phase unchanged (integrate_cart_ks.hpp:268-299).  A non-finite
    momentum makes every attempt non-finite, so each shrinks dlambda by
    min_step_scale until the call gives up."""
    kw = c1_params()
    su, sv = pixel_screen(4)
    ph0 = oracle.init_image(oracle.make_params(**kw), su, sv, 0.1)
    bad = [0, 5, 9]
    ph0["k2"][bad] = np.nan
    A, B, st_o, st_h = run(oracle, c1, kw, ph0, max_steps=50)
    assert_same(A, B, st_o, st_h)
    assert st_o["failed_steps"] == len(bad)
    failed = (A["terminate"] == 1) & (A["phase"] == 1)
    assert sorted(np.nonzero(failed)[0]) == bad
    assert np.allclose(A["dlambda"][bad], ph0["dlambda"][bad] * 0.01 ** 12, rtol=1e-12)


def test_zero_error_steps_grow(oracle, c1):
    """A tolerance of 1e-300 rejects every step whose embedded error is not
    exactly zero; a step whose 4th- and 5th-order solutions round to the same
    state is accepted and grows by max_step_scale (integrate_cart_ks.hpp:
    214-227).  Far from the hole the fused stage sums often round so."""
    kw = c1_params(atol=1e-300, rtol=1e-300)
    su, sv = pixel_screen(4)
    ph0 = oracle.init_image(oracle.make_params(**kw), su, sv, 0.1)
    A, B, st_o, st_h = run(oracle, c1, kw, ph0, max_steps=50)
    assert_same(A, B, st_o, st_h)
    assert st_o["rejected_attempts"] > 0 and st_o["accepted_steps"] == 16 * 50


def test_vacuum_nan_and_inside_horizon(oracle):
    kw = c1_params(absorption_enabled=0, emission_enabled=0)
    su, sv = pixel_screen(4)
    ph0 = oracle.init_image(oracle.make_params(**kw), su, sv, 0.1)
    ph0["x1"][0] = np.nan
    ph0["x1"][1], ph0["x2"][1], ph0["x3"][1] = 0.5, 0.2, 0.1      # inside r_min
    ph0["I"][1] = 5.0
    A, B, st_o, st_h = run(oracle, None, kw, ph0, max_steps=400)
    assert_same(A, B, st_o, st_h)
    assert A["phase"][1] == 4 and A["I"][1] == 0.0


def test_fast_rhs_equals_reference_rhs(oracle):
    """The tracing kernel's right-hand side equals the oracle's fused one
    bit for bit; both lie within a few ulp of the oracle's right-hand side
    in the reference's own rounding (kerr_schild_core.hpp:161-192 term by
    term), and so does ks_rhs_fast (the mixed-precision form's, the same
    equations with the contractions formed directly), from the horizon
    region out to the camera."""
    import harness
    rng = np.random.default_rng(8)
    for a in (0.0, 0.5, 0.94, -0.7):
        st = []
        for _ in range(400):
            R = np.exp(rng.uniform(np.log(1.2), np.log(2000.0)))
            th = np.arccos(rng.uniform(-1, 1))
            ph = rng.uniform(0, 2 * np.pi)
            X = np.array([rng.normal(), R * np.sin(th) * np.cos(ph), R * np.sin(th) * np.sin(ph),
                          R * np.cos(th)])
            p = rng.normal(size=4)
            st.append(np.concatenate([X, p]))
        st = np.array(st)
        ex, fa = harness.rhs(st, a, 1.0)
        ref = np.array([oracle.rhs(s[:4], s[4:], a, 1.0) for s in st])
        fused = np.array([oracle.rhs_fused(s[:4], s[4:], a, 1.0) for s in st])
        ref[:, 4] = fused[:, 4] = 0.0
        scale = np.abs(ref).max(axis=1, keepdims=True)
        assert np.array_equal(ex, fused)
        print("fused vs reference rounding: %.2e" % (np.abs(fused - ref) / scale).max())
        assert (np.abs(fused - ref) / scale).max() < 1e-14, a
        assert (np.abs(fa - ref) / scale).max() < 1e-12, a
