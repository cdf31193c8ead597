"""The mixed FP32/FP64 kernel against the FP64 kernel on the same rays.

Workload: the precision-form torus (a = 0.94, r_in = 6, r_pmax = 12, no
density floor so the torus and its shadow are visible) on a 64 x 32 x 32
grid, camera at 17 deg, RK45 rtol = atol = 1e-7.  Criteria: median per-pixel
relative Stokes-I error over the lit pixels <= 1e-4 (the stated bound of the
mixed form), the unlit central region (when there is one) within two pixels,
the photons as close to null as in the FP64 form, and the mixed kernel
really ran (its states differ from the FP64 ones in the last digits).

This torus reaches the grid edge and fills 0.30 < theta < 2.84, so at 17 deg
the centre of the image sees torus surface and no shadow; the ring diameter
is then undefined (NaN) for both forms.
"""
import numpy as np
import pytest

from conftest import c1_fields, c1_params, pixel_screen

pytestmark = pytest.mark.gpu

MEDIAN_TOL = 1e-4


def precision_form(nx, spin=0.94, theta=np.radians(17.0)):
    fields = c1_fields(rho_floor=0.0, spin=spin)
    kw = c1_params(spin=spin, theta=theta, atol=1e-7, rtol=1e-7,
                   superphoton_count_normalization=1e9)
    su, sv = pixel_screen(nx)
    return fields, kw, su, sv


def render(kw, fields, su, sv, precision):
    import grt_amd
    ctx = grt_amd.Context(grt_amd.default_params(precision=precision, **kw))
    ctx.set_grid(fields)
    ph = grt_amd.empty_photons(len(su))
    ctx.init_image(ph, 0.1, su, sv)
    st = ctx.trace(ph, 20000)
    n = int(np.sqrt(len(su)))
    img = ctx.image_products(ph, n, n)
    return ph, st, img


def null_norm(ph, a, m):
    """|g^{mu nu} p_mu p_nu| / p_t^2 in Kerr-Schild form."""
    x, y, z = ph["x1"], ph["x2"], ph["x3"]
    p0, p = ph["k0"], np.stack([ph["k1"], ph["k2"], ph["k3"]])
    A = x * x + y * y + z * z - a * a
    r = np.sqrt(0.5 * (A + np.sqrt(A * A + 4 * a * a * z * z)))
    l = np.stack([(r * x + a * y) / (r * r + a * a), (r * y - a * x) / (r * r + a * a), z / r])
    H = m * r ** 3 / (r ** 4 + a * a * z * z)
    lp = -p0 + (l * p).sum(axis=0)
    return np.abs(-p0 * p0 + (p * p).sum(axis=0) - 2 * H * lp * lp) / (p0 * p0)


def test_mixed_image_accuracy(gpu):
    import grt_amd
    from grt_amd.image_metrics import pixel_errors, ring_diameter
    fields, kw, su, sv = precision_form(128)
    ph64, st64, img64 = render(kw, fields, su, sv, grt_amd.PRECISION_FP64)
    phm, stm, imgm = render(kw, fields, su, sv, grt_amd.PRECISION_MIXED)
    I64, Im = img64["image"][0], imgm["image"][0]
    e = pixel_errors(Im, I64)
    d64, dm = ring_diameter(I64, 50.0), ring_diameter(Im, 50.0)
    print("fp64", st64, "\nmixed", stm, "\nerrors", e, "ring", d64, dm)
    assert e["lit_pixels"] > 1000
    assert e["median_rel_err"] <= MEDIAN_TOL
    if np.isfinite(d64):
        assert abs(dm - d64) <= 2 * 50.0 / 128
    # both forms take comparable numbers of steps
    assert abs(stm["accepted_steps"] / st64["accepted_steps"] - 1.0) < 0.05
    # the mixed kernel ran: results differ in the last digits, and the arrival
    # points on the screen agree far below a pixel (the final position itself
    # includes the overshoot past the camera plane, which moves with the
    # last step's length)
    ok = (ph64["phase"] == grt_amd.PHASE_ARRIVED) & (phm["phase"] == grt_amd.PHASE_ARRIVED)
    assert np.abs(phm["x1"][ok] - ph64["x1"][ok]).max() > 0.0
    du = np.abs(phm["theta_disp"][ok] - ph64["theta_disp"][ok])
    dv = np.abs(phm["phi_disp"][ok] - ph64["phi_disp"][ok])
    assert np.median(du) < 1e-6 and np.median(dv) < 1e-6
    assert np.mean((du < 1e-3) & (dv < 1e-3)) > 0.99
    # photons stay on the null cone as well as in the FP64 form (the RK45
    # truncation error at rtol 1e-7 sets both)
    nm = np.nanmax(null_norm(phm, kw["a_bh"], 1.0)[ok])
    nf = np.nanmax(null_norm(ph64, kw["a_bh"], 1.0)[ok])
    print("null norm max: mixed %.3e fp64 %.3e" % (nm, nf))
    assert nm <= 10.0 * nf + 1e-12


def test_mixed_accuracy_on_the_precision_grid(gpu):
    """The precision configuration's grid (256 x 128 x 128, a = 0.94, no
    density floor, 17 deg, RK45 rtol 1e-7) at 256^2 rays: the mixed form and
    its FP64 control (the same relaxed FP64 geodesic, FP64 tetrad) against
    the FP64 render -- the median per-pixel Stokes-I error within 1e-4,
    under 2% of the lit pixels beyond 1e-4, the per-pixel observed
    frequency to 1e-6 in the median, the spectrum to 5% in L1 -- beside
    the FP64 render's own sensitivity to a one-ulp shift of the screen
    coordinates, which bounds what any form that is not bit-identical can
    reach."""
    import grt_amd
    from grt_amd.fm_torus import fishbone_moncrief
    from grt_amd.image_metrics import frequency_errors, pixel_errors, spectrum_l1
    fields = fishbone_moncrief(nr=256, ntheta=128, nphi=128, r_min=1.5, r_max=100.0, mass=1.0,
                               spin=0.94, r_in=6.0, r_pressure_max=12.0, rho_floor=0.0)
    kw = c1_params(spin=0.94, theta=np.radians(17.0), atol=1e-7, rtol=1e-7,
                   superphoton_count_normalization=1e9)
    su, sv = pixel_screen(256)
    out = {}
    for name, prec, u, v in (("fp64", 0, su, sv), ("mixed", 1, su, sv), ("fast_rhs", 2, su, sv),
                             ("ulp", 0, np.nextafter(su, np.inf), np.nextafter(sv, np.inf))):
        out[name] = render(kw, fields, u, v, prec)
    ph64, _, img64 = out["fp64"]
    I64 = img64["image"][0]
    arr64 = ph64["phase"] == grt_amd.PHASE_ARRIVED
    floor = pixel_errors(out["ulp"][2]["image"][0], I64)
    print("one-ulp screen shift of the FP64 render:", floor)
    for name in ("mixed", "fast_rhs"):
        ph, st, img = out[name]
        e = pixel_errors(img["image"][0], I64)
        fe = frequency_errors(ph["frequency"], ph64["frequency"],
                              ph["phase"] == grt_amd.PHASE_ARRIVED, arr64)
        sl = spectrum_l1(img["spectrum_I"], img64["spectrum_I"])
        print(name, st, "\n  I:", e, "\n  frequency:", fe, "\n  spectrum L1:", sl)
        assert e["lit_pixels"] > 40000
        assert e["median_rel_err"] <= MEDIAN_TOL
        assert e["frac_rel_err_gt_1e-4"] < 0.02
        assert fe["median_rel_err"] <= 1e-6 and fe["rays"] > 60000
        assert sl < 0.05


def test_mixed_rk4_and_general_grid(gpu):
    """RK4 and the non-separable search in the mixed kernel: same phases as
    FP64 for almost every ray, images close."""
    import grt_amd
    from grt_amd.image_metrics import pixel_errors
    fields, kw, su, sv = precision_form(64)
    fields = {k: v.copy() for k, v in fields.items()}
    fields["r"] = fields["r"] * (1.0 + 1e-12 * np.arange(fields["r"].shape[1])[None, :, None])
    kw = dict(kw, integrator=grt_amd.GRT_RK4)
    out = {}
    for prec in (grt_amd.PRECISION_FP64, grt_amd.PRECISION_MIXED):
        import grt_amd as G
        ctx = G.Context(G.default_params(precision=prec, **kw))
        ctx.set_grid(fields)
        ph = G.empty_photons(len(su))
        ctx.init_image(ph, 0.001, su, sv)
        st = ctx.trace(ph, 40000)
        out[prec] = (ph, ctx.image_products(ph, 64, 64))
    (p64, i64), (pm, im) = out[grt_amd.PRECISION_FP64], out[grt_amd.PRECISION_MIXED]
    assert np.mean(p64["phase"] == pm["phase"]) > 0.99
    e = pixel_errors(im["image"][0], i64["image"][0])
    print(e)
    assert e["median_rel_err"] <= MEDIAN_TOL


def chi2_same(a, b, edges):
    """Two-sample chi-squared p-value of two histograms on the same bins."""
    from scipy import stats
    ha, _ = np.histogram(a, edges)
    hb, _ = np.histogram(b, edges)
    # scale for unequal sample sizes (Press et al., Numerical Recipes 14.3.2)
    ka, kb = np.sqrt(hb.sum() / ha.sum()), np.sqrt(ha.sum() / hb.sum())
    m = (ha + hb) > 0
    chi2 = (((ka * ha - kb * hb)[m]) ** 2 / (ha + hb)[m]).sum()
    return stats.chi2.sf(chi2, m.sum() - 1)


def test_mixed_scattering_statistics(oracle, gpu):
    """Batched scattering Monte Carlo in disk mode at the settings of the
    scattering configuration (a = 0.5 torus, albedo 1, absorption and
    emission on, RK45 rtol 1e-7) on a 32 x 16 x 32 grid, density scale
    0.125 g/cm^3 (about half the packets scatter).  (1) On the same packets
    and the same Philox stream (seed 1234) the mixed kernel follows the FP64
    oracle's paths: most packets leave in the same direction.  (2) Against
    an independent oracle sample (seed 4321) the escape-direction and
    log Stokes-I histograms agree by chi-squared, p > 0.01."""
    import grt_amd
    from grt_amd.fm_torus import fishbone_moncrief
    f = fishbone_moncrief(nr=32, ntheta=16, nphi=32, r_min=2.0, r_max=100.0, mass=1.0,
                          spin=0.5, r_in=6.0, r_pressure_max=12.0)
    units = {"density_g_cm3_per_code": 0.125}

    def kwargs(seed):
        return c1_params(spin=0.5, mode=1, atol=1e-7, rtol=1e-7,
                         superphoton_count_normalization=10.0, scattering_enabled=1, albedo=1.0,
                         seed=seed, r_term_max=1000.0, absorption_coefficient=1.0,
                         emission_coefficient=1.0)

    def oracle_run(seed):
        po = oracle.make_params(**kwargs(seed))
        G = oracle.GridHolder(f, 0.5, 1.0, units)
        o0 = oracle.init_disk(po, G, 0.1)
        o = {k: v.copy() for k, v in o0.items()}
        oracle.trace(po, G, o, 20000, 1.0, 0, 8)
        return o0, o

    o0, o = oracle_run(1234)
    _, o2 = oracle_run(4321)
    n = len(o0["x0"])
    assert n > 5000
    ctx = grt_amd.Context(grt_amd.default_params(precision=grt_amd.PRECISION_MIXED,
                                                 **kwargs(1234), **units))
    ctx.set_grid(f)
    g = grt_amd.to_device(o0)
    st = ctx.trace(g, 20000)
    g = grt_amd.to_host(g)
    esc_g, esc_o, esc_2 = g["phase"] == 4, o["phase"] == 4, o2["phase"] == 4
    print("packets", n, "escaped", esc_g.sum(), esc_o.sum(), esc_2.sum())

    def direction(P, m):
        r = np.sqrt(P["x1"] ** 2 + P["x2"] ** 2 + P["x3"] ** 2)[m]
        return P["x3"][m] / r, np.arctan2(P["x2"][m], P["x1"][m])

    both = esc_g & esc_o
    cg, pg = direction(g, both)
    co, pc = direction(o, both)
    same = (np.abs(cg - co) < 1e-3) & (np.abs(np.angle(np.exp(1j * (pg - pc)))) < 1e-3)
    print("same-stream packets leaving in the same direction:", same.mean())
    assert both.mean() > 0.9 and same.mean() > 0.9

    cg, pg = direction(g, esc_g)
    c2, p2 = direction(o2, esc_2)
    ps = [chi2_same(cg, c2, np.linspace(-1, 1, 11)),
          chi2_same(pg, p2, np.linspace(-np.pi, np.pi, 9))]
    lit_g = esc_g & (g["I"] > 0)
    lit_2 = esc_2 & (o2["I"] > 0)
    lo, hi = np.percentile(np.log10(o2["I"][lit_2]), [1, 99])
    ps.append(chi2_same(np.log10(g["I"][lit_g]), np.log10(o2["I"][lit_2]),
                        np.linspace(lo, hi, 11)))
    print("chi2 p-values against an independent sample", ps)
    assert min(ps) > 0.01


@pytest.mark.parametrize("precision", [1, 2])
def test_mixed_kernels_with_several_photons_per_lane(gpu, precision):
    """The mixed kernel and its FP64 control on 65,536 rays: with 64-thread
    workgroups, one per CU, every lane takes several photons through its
    wavefront's stage (and the stage's claims shrink near the end of the
    list); in one launch each photon must end exactly as when 256-thread
    workgroups, two per CU, give every lane one photon (a photon's
    arithmetic does not depend on where or when it runs).  In launches of
    five calls the geodesics and Stokes I are again exact; the frequency of
    a photon that crosses a launch boundary after its emission agrees to
    1e-12 relative: its deferred thermal frequency is drawn after that
    launch, so its arrival redshift is applied in the kernel, with this
    form's relaxed reciprocal and root, instead of after the trace in FP64."""
    import torch
    import grt_amd
    fields, kw, su, sv = precision_form(256)
    ctx = grt_amd.Context(grt_amd.default_params(precision=precision, **kw))
    ctx.set_grid(fields)
    lanes = torch.cuda.get_device_properties(0).multi_processor_count * 64
    assert len(su) >= 4 * lanes, (len(su), lanes)
    ph0 = grt_amd.empty_photons(len(su), device=gpu)
    ctx.init_image(ph0, 0.1, torch.from_numpy(su).to(gpu), torch.from_numpy(sv).to(gpu))
    out = []
    for tune in ({"block": 256, "blocks_per_cu": 2, "sub_budget": 1048576},
                 {"block": 64, "blocks_per_cu": 1, "sub_budget": 1048576},
                 {"block": 64, "blocks_per_cu": 1, "sub_budget": 5}):
        ctx.tuning = tune
        ph = {k: v.clone() for k, v in ph0.items()}
        st = ctx.trace(ph, 20000)
        out.append((grt_amd.to_host(ph), st))
    (ref, st_ref), (one, st_one), (five, st_five) = out
    assert st_one["launches"] == 1 and st_five["launches"] > 1
    keys = ("x0", "x1", "x2", "x3", "k0", "k1", "k2", "k3", "phase", "I", "dlambda")
    for k in keys + ("frequency",):
        assert np.array_equal(one[k], ref[k], equal_nan=True), (precision, k)
    for k in keys:
        assert np.array_equal(five[k], ref[k], equal_nan=True), (precision, k)
    f, fr = five["frequency"], ref["frequency"]
    both = np.isfinite(fr) & (fr != 0)
    assert np.array_equal(np.isfinite(f), np.isfinite(fr))
    assert np.all(np.abs(f[both] - fr[both]) <= 1e-12 * np.abs(fr[both])), precision
    print("frequencies differing in the five-call launches:", int((f[both] != fr[both]).sum()))
    for st in (st_one, st_five):
        for k in ("functor_calls", "accepted_steps", "rejected_attempts", "fluid_hits",
                  "global_steps", "terminated"):
            assert st[k] == st_ref[k], (precision, k)
    ctx.close()
