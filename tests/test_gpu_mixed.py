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
