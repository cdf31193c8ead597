"""Two reference configurations at their own sizes on the GPU.

Batched scattering Monte Carlo (disk mode): the a = 0.5 torus on
256 x 128 x 256, about 1e6 packets, albedo 1, absorption and emission,
Philox seed 1234, RK45 rtol 1e-7 (tests/golden/c5_scattering.py), traced by
the FP64 and by the mixed-precision kernel.  Scattering makes the packets'
paths random, so parity is statistical: the outcome categories, the escape
direction (cos theta, phi of the final position) and the escaped packets'
log I and log frequency agree with the oracle's histograms of the same
workload traced with an independent seed (tests/golden/c5_oracle_histograms.npz)
by a two-sample chi-squared test at p > 0.01.

Differentiation form: the a = 0.9 torus on 128 x 64 x 64, RK4 at a fixed
affine step of 0.05 M.  The 512^2 image at 45 deg has an identically zero
gradient (image rays meet no fluid after their emission event).  The
gradient of L = ||I - I_target||_2 over disk packets emitted in the torus,
with the target traced through rho * exp(0.1 g) (g: seed-42 Gaussian noise
smoothed by a three-cell Gaussian kernel, sigma one cell, periodic in phi),
agrees with central finite differences at relative step 1e-6 to 1e-4 on 64
cells chosen with default_rng(42) among the cells L resolves at that step,
and on both transfer coefficients; an RK45 (rtol 1e-8) run repeats the
check on eight cells and the coefficients.
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))


@pytest.fixture(scope="module")
def c5_fields():
    from grt_amd.fm_torus import fishbone_moncrief
    from c5_scattering import GRID
    return fishbone_moncrief(**GRID)


@pytest.mark.parametrize("precision", [0, 1])
def test_scattering_monte_carlo_histograms(gpu, c5_fields, precision):
    import torch
    import grt_amd
    from c5_scattering import (DENSITY_SCALE, MAX_STEPS, chi2_two_sample, outcome_histograms,
                               params)
    ref = np.load(os.path.join(HERE, "golden", "c5_oracle_histograms.npz"))
    kw = params(1234)
    ctx = grt_amd.Context(grt_amd.default_params(precision=precision,
                                                 density_g_cm3_per_code=DENSITY_SCALE, **kw))
    ctx.set_grid(c5_fields)
    n = ctx.disk_photon_count()
    assert n == int(ref["packets"][0])
    ph = grt_amd.empty_photons(n, device=gpu)
    ctx.init_disk(ph, 0.1)
    st = ctx.trace(ph, MAX_STEPS)
    torch.cuda.synchronize()
    h = outcome_histograms(grt_amd.to_host(ph))
    print("precision", precision, st, "categories", h["categories"], ref["categories"])
    assert h["categories"][2] == 0 and ref["categories"][2] == 0    # all terminated
    for name in ("categories", "cos_theta", "phi", "log_I", "log_nu"):
        chi2, dof, p = chi2_two_sample(h[name], ref[name])
        print(f"  {name}: chi2 {chi2:.1f} / {dof} dof, p = {p:.3g}")
        assert p > 0.01, name
    # the accepted-step totals of the two seeds agree to the sampling noise
    assert st["accepted_steps"] == pytest.approx(float(ref["accepted_steps"][0]), rel=0.02)


def c4_fields():
    from grt_amd.fm_torus import fishbone_moncrief
    return fishbone_moncrief(nr=128, ntheta=64, nphi=64, r_min=1.5, r_max=100.0, mass=1.0,
                             spin=0.9, r_in=6.0, r_pressure_max=12.0)


def c4_params(**kw):
    d = dict(a_bh=0.9, m_bh=1.0, r_term_min=1.0 + np.sqrt(1.0 - 0.81), atol=1e-8, rtol=1e-8,
             integrator=0, absorption_enabled=1, emission_enabled=1, absorption_coefficient=0.05,
             emission_coefficient=1.0, camera_distance=1000.0, camera_theta=np.pi / 4,
             camera_phi=0.0, camera_width=50.0, camera_height=50.0)
    d.update(kw)
    return d


def test_differentiation_image_gradient_is_zero(gpu):
    """The 512^2 image at 45 deg with RK4 at 0.05 M: every ray's forward
    pass retraces vacuum, so dL/drho, dL/dT and dL/dkappa vanish."""
    import torch
    import grt_amd
    fields = c4_fields()
    ctx = grt_amd.Context(grt_amd.default_params(**c4_params(
        mode=0, r_term_max=1500.0, superphoton_count_normalization=1e8)))
    ctx.set_grid(fields)
    m = 512
    u = -25.0 + (np.arange(m) + 0.5) * 50.0 / m
    U, V = np.meshgrid(u, u, indexing="xy")
    g = grt_amd.empty_photons(m * m, device=gpu)
    ctx.init_image(g, 5e-5, torch.from_numpy(U.ravel().copy()).to(gpu),
                   torch.from_numpy(V.ravel().copy()).to(gpu))
    st = ctx.trace_taped(g, 50000, slots_per_photon=8)
    arrived = (g["phase"] == grt_amd.PHASE_ARRIVED)
    print(st, int(arrived.sum()))
    assert st["terminated"] == m * m and int(arrived.sum()) > m * m // 2
    gr, gt, gc = ctx.adjoint(torch.ones(m * m, dtype=torch.float64, device=gpu))
    assert not gr.any() and not gt.any() and not gc.any()


def smoothed_noise(shape, seed=42):
    """Seed-42 standard normal noise smoothed by a Gaussian of one cell
    truncated at one cell (a three-cell kernel), periodic in phi, nearest
    at the r and theta edges."""
    from scipy.ndimage import gaussian_filter
    g = np.random.default_rng(seed).standard_normal(shape)
    return gaussian_filter(g, sigma=1.0, truncate=1.0, mode=("nearest", "nearest", "wrap"))


@pytest.mark.parametrize("integrator,cells", [(0, 64), (1, 8)])
def test_differentiation_protocol(gpu, integrator, cells):
    import torch
    import grt_amd
    from grt_amd.autograd import transfer_intensity
    fields = c4_fields()
    extra = dict(integrator=integrator, mode=1, r_term_max=150.0,
                 superphoton_count_normalization=1.5, seed=42)
    ctx = grt_amd.Context(grt_amd.default_params(**c4_params(**extra)))
    ctx.set_grid(fields)
    n = ctx.disk_photon_count()
    assert 2000 < n < 20000
    ph0 = grt_amd.empty_photons(n, device=gpu)
    ctx.init_disk(ph0, 0.05)          # RK4: the fixed step; RK45: the first step
    max_steps = 40000
    rho0 = torch.from_numpy(fields["rho"]).to(gpu)
    t0 = torch.from_numpy(fields["Tgas"]).to(gpu)
    ka0, ke0 = 0.05, 1.0

    def forward(rho, tg, ka, ke):
        p = ctx.params
        p.absorption_coefficient, p.emission_coefficient = ka, ke
        ctx.set_params(p)
        ctx.set_fluid_scalars(rho, tg)
        g = {k: v.clone() for k, v in ph0.items()}
        st = ctx.trace(g, max_steps)
        assert st["terminated"] == n
        return g["I"]

    target = forward(rho0 * torch.from_numpy(np.exp(0.1 * smoothed_noise(rho0.shape))).to(gpu),
                     t0, ka0, ke0).clone()

    def loss(rho, tg, ka, ke):
        return float(torch.linalg.vector_norm(forward(rho, tg, ka, ke) - target))

    rho = rho0.clone().requires_grad_()
    tg = t0.clone().requires_grad_()
    ka = torch.tensor(ka0, dtype=torch.float64, requires_grad=True)
    ke = torch.tensor(ke0, dtype=torch.float64, requires_grad=True)
    I = transfer_intensity(ctx, ph0, rho, tg, ka, ke, max_steps=max_steps, slots_per_photon=256)
    L = torch.linalg.vector_norm(I - target)
    L.backward()
    L0 = float(L.detach())
    assert L0 > 0
    gr = rho.grad.cpu().numpy().ravel()
    h = 1e-6
    # cells whose central difference at relative step h resolves dL/drho to
    # ~1e-5: |dL/drho rho| >= 1e-4 L (the difference then stands 2e-10 L
    # above L's rounding)
    sens = np.abs(gr * fields["rho"].ravel())
    ok = np.flatnonzero(sens >= 1e-4 * L0)
    print("photons", n, "L", L0, "cells with gradient", np.count_nonzero(gr), "resolvable",
          len(ok))
    assert len(ok) >= cells
    pick = np.random.default_rng(42).choice(ok, cells, replace=False)
    worst = 0.0
    for c in pick:
        v0 = float(fields["rho"].ravel()[c])
        rp = rho0.clone().view(-1)
        rm = rho0.clone().view(-1)
        rp[c] = v0 * (1 + h)
        rm[c] = v0 * (1 - h)
        fd = (loss(rp.view(rho0.shape), t0, ka0, ke0) -
              loss(rm.view(rho0.shape), t0, ka0, ke0)) / (2 * h * v0)
        rel = abs(gr[c] - fd) / abs(fd)
        worst = max(worst, rel)
        assert rel <= 1e-4, (c, gr[c], fd)
    for name, grad, base in (("kappa_abs", ka.grad.item(), ka0), ("kappa_emit", ke.grad.item(), ke0)):
        args_p = (ka0 * (1 + h), ke0) if name == "kappa_abs" else (ka0, ke0 * (1 + h))
        args_m = (ka0 * (1 - h), ke0) if name == "kappa_abs" else (ka0, ke0 * (1 - h))
        fd = (loss(rho0, t0, *args_p) - loss(rho0, t0, *args_m)) / (2 * h * base)
        print(name, grad, fd)
        assert grad == pytest.approx(fd, rel=1e-4), name
    print("worst cell rel err", worst)
    ctx.set_fluid_scalars(rho0, t0)
