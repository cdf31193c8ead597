"""GPU parity of the HIP tracing path against the CPU oracle (FP64).

Tolerances (stated for the traced path): per-photon final x^mu and k_mu agree
to 1e-9 relative, per-pixel Stokes to 1e-8 relative.  A 4-vector's relative
error is max_mu |a_mu - b_mu| / max_mu |b_mu| (components cross zero).  The
device contracts a*b+c into FMA and its libm differs from glibc by an ulp, so
agreement is to rounding; rays grazing the photon sphere amplify rounding, so
the per-photon criterion is asserted for >= 99.9 % of rays and the worst ray
is reported.
"""
import numpy as np
import pytest

from conftest import c1_fields, c1_params, pixel_screen

pytestmark = pytest.mark.gpu

TOL_STATE = 1e-9
TOL_PIXEL = 1e-8
UNIT_KEYS = ("length_cm_per_code", "time_s_per_code", "density_g_cm3_per_code",
             "temperature_k_per_code", "four_velocity_cm_s_per_code", "magnetic_gauss_per_code")


def vec_rel_err(a, b, names):
    A = np.stack([np.asarray(a[k], dtype=np.float64) for k in names])
    B = np.stack([np.asarray(b[k], dtype=np.float64) for k in names])
    scale = np.maximum(np.abs(B).max(axis=0), 1e-300)
    err = (np.abs(A - B) / scale).max(axis=0)
    err[np.all(np.isnan(A) & np.isnan(B), axis=0)] = 0.0
    return err


def oracle_parts(oracle, kw, fields):
    okw = {k: v for k, v in kw.items() if k not in UNIT_KEYS}
    units = {k: v for k, v in kw.items() if k in UNIT_KEYS}
    G = oracle.GridHolder(fields, kw["a_bh"], kw["m_bh"], units) if fields is not None else None
    return oracle.make_params(**okw), G


def gpu_ctx(kw, fields):
    import grt_amd
    ctx = grt_amd.Context(grt_amd.default_params(**kw))
    if fields is not None:
        ctx.set_grid(fields)
    return ctx


def trace_both(oracle, ctx, po, G, ph0, max_steps=10000, thr=1.0, interval=0):
    import grt_amd
    g = grt_amd.to_device(ph0)
    st_g = ctx.trace(g, max_steps, thr, interval)
    g = grt_amd.to_host(g)
    o = {k: v.copy() for k, v in ph0.items()}
    st_o = oracle.trace(po, G, o, max_steps, thr, interval, 8)
    return g, o, st_g, st_o


def check_states(g, o, frac=0.999):
    assert np.array_equal(g["phase"], o["phase"])
    assert np.array_equal(g["terminate"], o["terminate"])
    ex = vec_rel_err(g, o, ["x0", "x1", "x2", "x3"])
    ek = vec_rel_err(g, o, ["k0", "k1", "k2", "k3"])
    print("x err max %.3e  frac>tol %.5f | k err max %.3e frac>tol %.5f"
          % (ex.max(), np.mean(ex > TOL_STATE), ek.max(), np.mean(ek > TOL_STATE)))
    assert np.mean(ex <= TOL_STATE) >= frac and np.mean(ek <= TOL_STATE) >= frac
    return ex, ek


def check_image(ctx, oracle, g, o, nx, width=50.0):
    img_g = ctx.image_products(g, nx, nx)
    img_o = oracle.image_products(o, nx, nx, 128, 1e8, 1e22, width, width)
    assert img_g["arrived_count"] == img_o["arrived_count"]
    I_g, I_o = img_g["image"], img_o["image"]
    rel = np.abs(I_g - I_o) / np.maximum(np.abs(I_o), 1e-300)
    rel[(I_o == 0) & (I_g == 0)] = 0.0
    print("pixel err: max %.3e median %.3e frac>tol %.5f"
          % (rel.max(), np.median(rel), np.mean(rel > TOL_PIXEL)))
    assert np.mean(rel <= TOL_PIXEL) >= 0.999
    assert np.allclose(img_g["spectrum_frequency_hz"], img_o["spectrum_frequency_hz"], rtol=1e-15)
    return rel


@pytest.fixture(scope="module")
def c1():
    return c1_fields()


def test_image_init_matches_oracle(oracle, gpu):
    import grt_amd
    kw = c1_params(camera_phi=0.3)
    su, sv = pixel_screen(32)
    ctx = gpu_ctx(kw, None)
    g = grt_amd.empty_photons(len(su))
    ctx.init_image(g, 0.1, su, sv, id_offset=5)
    po, _ = oracle_parts(oracle, kw, None)
    o = oracle.init_image(po, su, sv, 0.1, id_offset=5)
    assert np.array_equal(g["id"], o["id"]) and np.array_equal(g["phase"], o["phase"])
    assert vec_rel_err(g, o, ["x1", "x2", "x3"]).max() < 1e-15
    assert vec_rel_err(g, o, ["k0", "k1", "k2", "k3"]).max() < 1e-12
    assert np.array_equal(g["dlambda"], o["dlambda"])
    # random screen positions fill the plane uniformly
    r = grt_amd.empty_photons(20000)
    ctx.init_image(r, 0.1)
    assert np.abs(r["theta_disp"]).max() <= 25.0 and np.abs(r["phi_disp"]).max() <= 25.0
    assert abs(r["theta_disp"].mean()) < 0.5 and abs(r["theta_disp"].std() - 50 / np.sqrt(12)) < 0.5


@pytest.mark.parametrize("nx,norm", [(64, 1e8), (128, 1e8), (128, 1.0)])
def test_c1_image(oracle, gpu, c1, nx, norm):
    """The first reference configuration (64x32x32, a = 0.9, 60 deg, RK45
    rtol 1e-8) at its 128^2 rays and at 64^2, with the emitting normalisation
    and with the literal default 1.0 (int(rho * 1) = 0 packets below rho = 1:
    every ray is rejected at its emission event and the image is zero)."""
    kw = c1_params(superphoton_count_normalization=norm)
    su, sv = pixel_screen(nx)
    ctx = gpu_ctx(kw, c1)
    po, G = oracle_parts(oracle, kw, c1)
    ph0 = oracle.init_image(po, su, sv, 0.1)
    g, o, st_g, st_o = trace_both(oracle, ctx, po, G, ph0)
    print("stats gpu", st_g, "oracle", st_o)
    check_states(g, o)
    check_image(ctx, oracle, g, o, nx)
    for k in ("functor_calls", "accepted_steps", "rejected_attempts", "fluid_samples",
              "fluid_hits", "failed_steps", "global_steps", "terminated"):
        assert st_g[k] == st_o[k], k
    if norm == 1.0:
        assert np.all(g["phase"] == 4) and not g["I"].any()


def test_arrival_inside_the_ergoregion(oracle, gpu, c1):
    """A camera inside the ergoregion (equatorial, r = 1.8 < 2 M): most
    arrivals have g_00 > 0, an observer energy of 0 and a zeroed Stokes
    vector; the kernel completes these deferred arrivals after the launch
    (resolve_deferred), the oracle at the arrival, and both agree."""
    kw = c1_params(theta=np.pi / 2, camera_distance=1.8, camera_width=2.0,
                   camera_height=2.0, r_term_max=300.0)
    su, sv = pixel_screen(16, width=2.0, height=2.0)
    ctx = gpu_ctx(kw, c1)
    po, G = oracle_parts(oracle, kw, c1)
    ph0 = oracle.init_image(po, su, sv, 0.01)
    g, o, st_g, st_o = trace_both(oracle, ctx, po, G, ph0, max_steps=5000)
    check_states(g, o, frac=1.0)
    for k in ("I", "Q", "U", "V", "theta_disp", "phi_disp"):
        assert np.array_equal(g[k], o[k]), k
    # the emission-frame energy comes from the factored tetrad: 1e-14
    f = np.abs(g["frequency"] - o["frequency"]) / np.maximum(np.abs(o["frequency"]), 1e-300)
    assert f.max() <= 1e-14
    arrived = o["phase"] == 3
    assert (arrived & (o["I"] == 0)).sum() > 100 and (arrived & (o["I"] > 0)).sum() > 20


@pytest.mark.parametrize("generator", [1, 2])
def test_image_generators(oracle, gpu, c1, generator):
    """Image mode with the power-law generator (the emission event reads only
    the first tetrad vector) and the custom generator (it reads the whole
    fluid cell, magnetic field included): every photon as the oracle."""
    kw = c1_params(generator=generator, nu_min_hz=1e9, nu_max_hz=1e12)
    su, sv = pixel_screen(32)
    ctx = gpu_ctx(kw, c1)
    po, G = oracle_parts(oracle, kw, c1)
    ph0 = oracle.init_image(po, su, sv, 0.1)
    g, o, st_g, st_o = trace_both(oracle, ctx, po, G, ph0)
    check_states(g, o)
    check_image(ctx, oracle, g, o, 32)
    assert st_g["fluid_samples"] == st_o["fluid_samples"]
    assert st_g["fluid_hits"] == st_o["fluid_hits"]


def test_work_list_edges(oracle, gpu, c1):
    """The persistent loop's work list at its edges: batches smaller than,
    equal to and just above one wavefront, photons already terminated on
    entry (so the first list is not the identity), and tiny step budgets
    that split every trace over many compacting launches, under several
    launch shapes.  Every photon must end exactly as in the oracle."""
    kw = c1_params()
    su, sv = pixel_screen(16)
    ctx = gpu_ctx(kw, c1)
    po, G = oracle_parts(oracle, kw, c1)
    full = oracle.init_image(po, su, sv, 0.1)
    rng = np.random.default_rng(12)
    for n in (1, 3, 63, 64, 65, 130, 256):
        pick = np.sort(rng.choice(len(full["x0"]), n, replace=False))
        ph0 = {k: np.ascontiguousarray(v[pick]) for k, v in full.items()}
        if n > 3:
            ph0["terminate"][rng.choice(n, n // 5, replace=False)] = 1
        for tune in ({"block": 64, "blocks_per_cu": 1, "sub_budget": 3},
                     {"block": 256, "blocks_per_cu": 2, "sub_budget": 1048576},
                     {"block": 128, "blocks_per_cu": 4, "sub_budget": 7}):
            ctx.tuning = tune
            g, o, st_g, st_o = trace_both(oracle, ctx, po, G, ph0)
            check_states(g, o, frac=1.0)
            for k in ("x0", "x1", "k1", "I", "dlambda"):
                assert np.array_equal(g[k], o[k]), (n, tune, k)
            # thermal frequencies drawn at emission (log() may differ by an ulp)
            assert np.allclose(g["frequency"], o["frequency"], rtol=1e-14, atol=0.0,
                               equal_nan=True), (n, tune)
            for k in ("functor_calls", "accepted_steps", "rejected_attempts", "global_steps",
                      "terminated"):
                assert st_g[k] == st_o[k], (n, tune, k)
    ctx.tuning = {"block": 256, "blocks_per_cu": 2, "sub_budget": 256}


def test_stage_refills_over_many_photons_per_lane(oracle, gpu, c1):
    """Four photons per lane: every lane takes several photons from its
    wavefront's stage, which claims 64 slots while more than two photons per
    lane remain and then short claims (rounded up to eight slots while the
    wavefront's lanes take photons quickly, exact otherwise, as in the first
    iterations of every launch), in one launch and in launches of five calls
    that compact the survivors.  Some photons are terminated on entry, so the
    first list is not the identity.  Every photon must end exactly as in the
    oracle."""
    import grt_amd
    kw = c1_params()
    su, sv = pixel_screen(256)
    ctx = gpu_ctx(kw, c1)
    po, G = oracle_parts(oracle, kw, c1)
    ph0 = oracle.init_image(po, su, sv, 0.1)
    ph0["terminate"][::97] = 1
    o = {k: v.copy() for k, v in ph0.items()}
    st_o = oracle.trace(po, G, o, 10000, 1.0, 0, 8)
    import torch
    lanes = torch.cuda.get_device_properties(0).multi_processor_count * 64
    assert len(su) >= 4 * lanes > 0, (len(su), lanes)
    for tune in ({"block": 64, "blocks_per_cu": 1, "sub_budget": 1048576},
                 {"block": 64, "blocks_per_cu": 1, "sub_budget": 5}):
        ctx.tuning = tune
        g = grt_amd.to_device(ph0)
        st_g = ctx.trace(g, 10000)
        g = grt_amd.to_host(g)
        check_states(g, o, frac=1.0)
        for k in ("x0", "x1", "x2", "x3", "k0", "k1", "k2", "k3", "I", "Q", "dlambda"):
            assert np.array_equal(g[k], o[k]), (tune, k)
        assert np.allclose(g["frequency"], o["frequency"], rtol=1e-14, atol=0.0,
                           equal_nan=True), tune
        for k in ("functor_calls", "accepted_steps", "rejected_attempts", "global_steps",
                  "terminated", "fluid_hits"):
            assert st_g[k] == st_o[k], (tune, k)
        assert st_g["launches"] > (1 if tune["sub_budget"] == 5 else 0)
    ctx.tuning = {"block": 256, "blocks_per_cu": 2, "sub_budget": 256}


def test_empty_batches(gpu, c1):
    """Zero photons: every entry point succeeds and does nothing (the
    reference's loops run over an empty photon view)."""
    import grt_amd
    ctx = gpu_ctx(c1_params(), c1)
    for dev in (None, gpu):
        ph = grt_amd.empty_photons(0, device=dev)
        st = ctx.trace(ph, 100)
        assert st["functor_calls"] == 0 and st["accepted_steps"] == 0 and st["terminated"] == 0
        ctx.init_image(ph, 0.1)
        img = ctx.image_products(ph, 8, 8)
        assert img["arrived_count"] == 0 and not np.asarray(
            img["image"].cpu() if hasattr(img["image"], "cpu") else img["image"]).any()
    # more than 2^31 photons in one call is refused before any array is read
    import ctypes as C
    small = grt_amd.empty_photons(4, device=gpu)
    s = grt_amd.photon_struct(small)
    st = grt_amd.Stats()
    rc = ctx._L.grt_trace_device(ctx._h, C.byref(s), (1 << 31) + 1, 10, 1.0, 0, None,
                                 C.byref(st))
    assert rc == -1                                   # GRT_E_INVALID
    assert "2^31" in ctx._L.grt_last_error(ctx._h).decode()


def test_precision_workload_geometry_128(oracle, gpu):
    """The precision workload's configuration (FM torus a=0.94 without a
    density floor on 256x128x128, 17 deg, RK45 rtol 1e-7) at 128^2 rays:
    positions, momenta, step sizes and counters bit-identical to the oracle;
    Stokes I and the image to the last ulp of exp (1e-14)."""
    spin = 0.94
    fields = c1_fields(rho_floor=0.0, nr=256, nth=128, nph=128, spin=spin)
    kw = c1_params(spin=spin, theta=np.radians(17.0), atol=1e-7, rtol=1e-7,
                   superphoton_count_normalization=1e9)
    su, sv = pixel_screen(128)
    ctx = gpu_ctx(kw, fields)
    po, G = oracle_parts(oracle, kw, fields)
    ph0 = oracle.init_image(po, su, sv, 0.1)
    g, o, st_g, st_o = trace_both(oracle, ctx, po, G, ph0)
    for k in ("x0", "x1", "x2", "x3", "k0", "k1", "k2", "k3", "dlambda", "phase", "terminate",
              "theta_disp", "phi_disp"):
        assert np.array_equal(g[k], o[k], equal_nan=True), k
    for k in ("functor_calls", "accepted_steps", "rejected_attempts", "fluid_samples",
              "fluid_hits", "failed_steps", "global_steps", "terminated"):
        assert st_g[k] == st_o[k], k
    assert np.allclose(g["I"], o["I"], rtol=1e-14, atol=0.0)
    img_g = ctx.image_products(g, 128, 128)
    img_o = oracle.image_products(o, 128, 128, 128, 1e8, 1e22, 50.0, 50.0)
    assert np.allclose(img_g["image"], img_o["image"], rtol=1e-14, atol=0.0)


def test_all_terminated_on_entry(oracle, gpu, c1):
    """Photons all terminated on entry: the first launch reads a live count of
    zero on the device and leaves every photon untouched, as the reference's
    functor returns at once for a terminated photon."""
    import grt_amd
    kw = c1_params()
    su, sv = pixel_screen(16)
    ctx = gpu_ctx(kw, c1)
    po, G = oracle_parts(oracle, kw, c1)
    ph0 = oracle.init_image(po, su, sv, 0.1)
    ph0["terminate"][:] = 1
    g, o, st_g, st_o = trace_both(oracle, ctx, po, G, ph0, max_steps=50)
    assert st_g["functor_calls"] == 0 and st_g["terminated"] == len(su)
    for k in ("x0", "x1", "k1", "I", "frequency", "dlambda", "phase"):
        assert np.array_equal(g[k], ph0[k]) and np.array_equal(o[k], ph0[k]), k


def test_rk4_general_grid_and_capture(oracle, gpu):
    """RK4, the non-separable search path and horizon capture together."""
    from grt_amd.fm_torus import fishbone_moncrief
    f = fishbone_moncrief(nr=48, ntheta=24, nphi=16, r_min=1.5, r_max=100.0, spin=0.9,
                          r_in=6.0, r_pressure_max=9.0, rho_floor=0.0)
    f["r"] = f["r"] * (1.0 + 1e-6 * f["phi"])         # breaks separability
    kw = c1_params(theta=np.pi / 4, integrator=0, camera_width=30.0, camera_height=30.0)
    su, sv = pixel_screen(20, width=30.0, height=30.0)
    ctx = gpu_ctx(kw, f)
    po, G = oracle_parts(oracle, kw, f)
    ph0 = oracle.init_image(po, su, sv, 0.002)
    g, o, st_g, st_o = trace_both(oracle, ctx, po, G, ph0, max_steps=3000)
    check_states(g, o, frac=0.99)
    counts = np.bincount(o["phase"], minlength=5)
    assert counts[3] > 0 and counts[4] > 0


def test_termination_fraction_schedule(oracle, gpu, c1):
    """termination_percent < 1 stops every photon at the first output step
    where the terminated fraction reaches it (integrate_cart_ks.hpp:492-526)."""
    kw = c1_params()
    su, sv = pixel_screen(24)
    ctx = gpu_ctx(kw, c1)
    po, G = oracle_parts(oracle, kw, c1)
    ph0 = oracle.init_image(po, su, sv, 0.1)
    for thr, interval, max_steps in ((0.5, 5, 10000), (0.9, 1, 10000), (1.0, 3, 17), (0.2, 7, 9)):
        g, o, st_g, st_o = trace_both(oracle, ctx, po, G, ph0, max_steps, thr, interval)
        check_states(g, o, frac=1.0)
        # photons stopped between emission and the camera keep their
        # emitted frequency; arrived ones carry the redshift factor
        assert np.allclose(g["frequency"], o["frequency"], rtol=1e-14, atol=0.0,
                           equal_nan=True), (thr, interval)
        assert st_g["global_steps"] == st_o["global_steps"], (thr, interval)
        assert st_g["accepted_steps"] == st_o["accepted_steps"], (thr, interval)


def test_trace_resumes_after_stop(oracle, gpu, c1):
    """A trace stopped by max_steps between emission and the camera (the
    photon's frequency is drawn before the call returns) and then traced
    again to the end ends exactly as in the oracle."""
    import grt_amd
    kw = c1_params()
    su, sv = pixel_screen(24)
    ctx = gpu_ctx(kw, c1)
    po, G = oracle_parts(oracle, kw, c1)
    ph0 = oracle.init_image(po, su, sv, 0.1)
    g, o, _, _ = trace_both(oracle, ctx, po, G, ph0, max_steps=17)
    assert np.count_nonzero(o["phase"] == 2) > 100      # emitted, not yet at the camera
    assert np.allclose(g["frequency"], o["frequency"], rtol=1e-14, atol=0.0, equal_nan=True)
    gd = grt_amd.to_device(g)
    ctx.trace(gd, 10000)
    g = grt_amd.to_host(gd)
    oracle.trace(po, G, o, 10000, 1.0, 0, 8)
    check_states(g, o, frac=1.0)
    assert np.allclose(g["frequency"], o["frequency"], rtol=1e-14, atol=0.0, equal_nan=True)
    check_image(ctx, oracle, g, o, 24)


def test_scattering_image(oracle, gpu, c1):
    """Image mode with scattering enabled: the forward trace retraces the
    backward path and meets no fluid, so no scattering fires (the reference's
    algorithm); GPU and oracle agree.  Scattering itself is tested in disk
    mode (test_disk_init_and_scattering_statistics)."""
    kw = c1_params(scattering_enabled=1, albedo=0.7, seed=99, length_cm_per_code=1e24)
    su, sv = pixel_screen(24)
    ctx = gpu_ctx(kw, c1)
    po, G = oracle_parts(oracle, kw, c1)
    ph0 = oracle.init_image(po, su, sv, 0.1)
    g, o, st_g, st_o = trace_both(oracle, ctx, po, G, ph0, max_steps=2000)
    # same Philox draws: identical decisions, states to rounding
    assert np.mean(g["phase"] == o["phase"]) > 0.99
    ok = g["phase"] == o["phase"]
    ex = vec_rel_err({k: g[k][ok] for k in ("x1", "x2", "x3")},
                     {k: o[k][ok] for k in ("x1", "x2", "x3")}, ["x1", "x2", "x3"])
    assert np.median(ex) < 1e-9


def test_disk_init_and_scattering_statistics(oracle, gpu):
    """Disk-mode packets from the device initialiser agree with the oracle's
    (same Philox stream); after multiple scattering the outcome categories,
    the escape direction (azimuth, cos theta) and the escaped packets'
    log Stokes I agree (two-sample chi-squared, p > 0.01).  The paths share
    their draws until the kernel's factored tetrad and the oracle's
    16-term inner products part in the last bit, which repeated scattering
    amplifies."""
    import os
    import sys
    import grt_amd
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    from c5_scattering import chi2_two_sample
    f = c1_fields(nr=24, nth=12, nph=12)
    kw = c1_params(mode=1, superphoton_count_normalization=60.0, scattering_enabled=1,
                   albedo=0.9, length_cm_per_code=2e23, generator=0,
                   temperature_k_per_code=1e10, r_term_max=1e4)
    ctx = gpu_ctx(kw, f)
    n = ctx.disk_photon_count()
    po, G = oracle_parts(oracle, kw, f)
    assert n == oracle.disk_count(po, G) and n > 15000
    g0 = grt_amd.empty_photons(n)
    ctx.init_disk(g0, 0.1)
    o0 = oracle.init_disk(po, G, 0.1)
    assert np.array_equal(g0["id"], o0["id"]) and np.array_equal(g0["phase"], o0["phase"])
    assert vec_rel_err(g0, o0, ["x1", "x2", "x3"]).max() < 1e-13
    assert vec_rel_err(g0, o0, ["k0", "k1", "k2", "k3"]).max() < 1e-12
    assert np.allclose(g0["frequency"], o0["frequency"], rtol=1e-13)
    g, o, st_g, st_o = trace_both(oracle, ctx, po, G, o0, max_steps=4000)

    def outcomes(P):
        r = np.sqrt(P["x1"] ** 2 + P["x2"] ** 2 + P["x3"] ** 2)
        esc = (P["terminate"] != 0) & (r > 1e4) & (P["I"] > 0)
        cats = np.array([esc.sum(), ((P["terminate"] != 0) & ~esc).sum(),
                         (P["terminate"] == 0).sum()])
        return esc, cats, r

    esc_g, cat_g, r_g = outcomes(g)
    esc_o, cat_o, r_o = outcomes(o)
    print("categories gpu", cat_g, "oracle", cat_o)
    assert cat_g[0] > 5000
    logi = np.log10(np.concatenate([g["I"][esc_g], o["I"][esc_o]]))
    li_edges = np.quantile(logi, np.linspace(0, 1, 21))
    li_edges[0], li_edges[-1] = -np.inf, np.inf
    hists = {
        "categories": (cat_g, cat_o),
        "azimuth": tuple(np.histogram(np.arctan2(P["x2"], P["x1"])[e],
                                      np.linspace(-np.pi, np.pi, 17))[0]
                         for P, e in ((g, esc_g), (o, esc_o))),
        "cos_theta": tuple(np.histogram((P["x3"] / r)[e], np.linspace(-1, 1, 17))[0]
                           for P, e, r in ((g, esc_g, r_g), (o, esc_o, r_o))),
        "log_I": tuple(np.histogram(np.log10(P["I"][e]), li_edges)[0]
                       for P, e in ((g, esc_g), (o, esc_o))),
    }
    for name, (hg, ho) in hists.items():
        chi2, dof, p = chi2_two_sample(hg, ho)
        print(f"{name}: chi2 {chi2:.1f} / {dof}, p {p:.3g}")
        assert p > 0.01, name


def test_image_products_device(oracle, gpu):
    import grt_amd
    rng = np.random.default_rng(1)
    n = 50000
    ph = grt_amd.empty_photons(n)
    ph["phase"][:] = rng.choice([3, 4], n, p=[0.8, 0.2]).astype(np.uint8)
    ph["theta_disp"][:] = rng.uniform(-26, 26, n)
    ph["phi_disp"][:] = rng.uniform(-26, 26, n)
    ph["theta_disp"][:10] = 25.0                      # exactly the last bin edge
    for k in ("I", "Q", "U", "V"):
        ph[k][:] = rng.normal(size=n)
    ph["frequency"][:] = 10 ** rng.uniform(7, 23, n)
    ctx = gpu_ctx(c1_params(), None)
    a = ctx.image_products(ph, 37, 29, 64, 1e8, 1e22)
    b = oracle.image_products(ph, 37, 29, 64, 1e8, 1e22, 50.0, 50.0)
    assert a["arrived_count"] == b["arrived_count"]
    assert np.allclose(a["image"], b["image"], rtol=1e-12, atol=1e-12)
    assert np.allclose(a["spectrum_I"], b["spectrum_I"], rtol=1e-12, atol=1e-12)


def test_full_size_invariants(gpu):
    """The production image (1024^2 rays, 256x128x128 grid): every ray
    arrives, p_t is exactly conserved and L_z is conserved along each path to
    the integrator's tolerance (rtol 1e-8 per step over ~30 steps)."""
    import grt_amd
    import torch
    from grt_amd.fm_torus import fishbone_moncrief
    f = fishbone_moncrief(nr=256, ntheta=128, nphi=128, r_min=1.5, r_max=100.0, spin=0.9,
                          r_in=6.0, r_pressure_max=12.0)
    kw = c1_params()
    ctx = gpu_ctx(kw, f)
    su, sv = pixel_screen(1024)
    dev = {k: v for k, v in grt_amd.to_device(grt_amd.empty_photons(len(su))).items()}
    ctx.init_image(dev, 0.1, torch.from_numpy(su).cuda(), torch.from_numpy(sv).cuda())
    h0 = grt_amd.to_host(dev)
    st = ctx.trace(dev, 10000)
    h = grt_amd.to_host(dev)
    assert st["terminated"] == len(su) and np.all(h["phase"] == 3)
    assert np.array_equal(h["k0"], h0["k0"])
    lz0 = h0["x1"] * h0["k2"] - h0["x2"] * h0["k1"]
    lz = h["x1"] * h["k2"] - h["x2"] * h["k1"]
    assert np.median(np.abs(lz - lz0) / np.maximum(np.abs(lz0), 1.0)) < 1e-5
    img = ctx.image_products(dev, 1024, 1024)
    assert img["arrived_count"] == len(su)
    assert np.isfinite(img["image"].cpu().numpy()).all()


def test_autotune_keeps_results(oracle, gpu, c1):
    """Every launch shape the autotuner tries gives the same photons."""
    import grt_amd
    from grt_amd.autotune import autotune
    kw = c1_params()
    su, sv = pixel_screen(48)
    ctx = gpu_ctx(kw, c1)
    po, _ = oracle_parts(oracle, kw, None)
    g0 = grt_amd.to_device(oracle.init_image(po, su, sv, 0.1))
    best, table = autotune(ctx, g0, 10000, shapes=((256, 2), (64, 8)), budgets=(8, 1 << 20),
                           repeats=1)
    assert len(table) == 4 and ctx.tuning == best
    ref = None
    for tune, _ in table:
        ctx.tuning = tune
        g = {k: v.clone() for k, v in g0.items()}
        st = ctx.trace(g, 10000)
        h = grt_amd.to_host(g)
        if ref is None:
            ref = h
        for k in ("x1", "k1", "I", "phase", "theta_disp"):
            assert np.array_equal(h[k], ref[k]), (tune, k)
        if tune["sub_budget"] == 8:
            assert st["launches"] > 1


def test_correctly_rounded_sqrt_and_reciprocal(gpu):
    """sqrt_rn / rcp_rn of the FP64 kernel equal the compiler's sqrt() and
    1.0 / d bit for bit: 2^27 pseudo-random arguments over binary exponents
    [-700, 700] and edge values (powers of two, neighbours of 1, the ends
    of the range, values the right-hand side meets)."""
    import ctypes as C
    import os
    here = os.path.join(os.path.dirname(__file__), "harness", "libgrt_math_check.so")
    L = C.CDLL(here)
    L.grt_math_check.restype = C.c_int
    L.grt_math_check.argtypes = [C.c_uint64, C.c_uint64, C.c_void_p, C.c_uint64, C.c_void_p]
    edge = [2.0 ** k for k in range(-700, 701, 7)]
    edge += [np.nextafter(1.0, 0), 1.0, np.nextafter(1.0, 2), 2.0 ** -700, 2.0 ** 700 * 1.9999,
             0.81, 1.0 + 2 ** -52, 4.0 - 2 ** -50, 1e-30, 1e30, 123456.789, 2.5e25]
    v = np.ascontiguousarray(edge, dtype=np.float64)
    bad = (C.c_ulonglong * 4)()
    assert L.grt_math_check(1 << 27, 12345, v.ctypes.data, len(v), bad) == 0
    print("mismatches sqrt/rcp random", bad[0], bad[1], "edge", bad[2], bad[3])
    assert list(bad) == [0, 0, 0, 0]


def test_controller_root_fast_path_equals_the_oracle_form(gpu):
    """The FP64 kernel's step-size root quarter_root_rcp(err) (short correctly
    rounded sequences, err in (2^-500, 2^500)) equals quarter_root(1.0 / err),
    the oracle's candidate check with the device's sqrt and nextafter, bit for
    bit on 2^27 pseudo-random err over that range (the oracle's form is itself
    pinned against mpmath on the CPU)."""
    import ctypes as C
    import os
    here = os.path.join(os.path.dirname(__file__), "harness", "libgrt_math_check.so")
    L = C.CDLL(here)
    L.grt_quarter_root_check.restype = C.c_int
    L.grt_quarter_root_check.argtypes = [C.c_uint64, C.c_uint64, C.c_void_p]
    bad = C.c_ulonglong(0)
    assert L.grt_quarter_root_check(1 << 27, 4242, C.byref(bad)) == 0
    print("quarter root mismatches", bad.value)
    assert bad.value == 0
