"""The differentiable form on the GPU: the checkpointed FP64 adjoint
(grt_trace_taped_device + grt_adjoint_device) against the oracle's reference
gradient (full record, no checkpoints) and against central finite
differences of the GPU forward pass; the autograd binding.

Workload: disk-mode photons emitted inside a small torus (24 x 12 x 12 grid,
a = 0.9), absorption and emission on, scattering off: every photon carries
radiative transfer along its path through the grid, so the gradient is
dense.  Loss L = sum_i w_i I_i with fixed random weights.  Bars: the final
photon geodesics equal the oracle's bit for bit (the taped trace is the FP64
kernel) and the Stokes parameters agree to 1e-13 relative (the device exp()
behind exp(-dtau) and glibc's may differ in the last bit); gradients agree
with the oracle's to 1e-10 relative (summation order of the atomics aside);
central differences agree to 1e-5 relative (relative step 1e-4, on the cells
whose difference L resolves).
"""
import numpy as np
import pytest

from conftest import c1_params
from test_oracle_adjoint import H_CELL, disk_setup, fd_cells, small_torus

pytestmark = pytest.mark.gpu


def gpu_setup(kw, fields):
    import grt_amd
    ctx = grt_amd.Context(grt_amd.default_params(**kw))
    ctx.set_grid(fields)
    return ctx


def rel_close(a, b, rel, floor):
    a = np.asarray(a)
    b = np.asarray(b)
    return np.all(np.abs(a - b) <= rel * np.abs(b) + floor)


@pytest.mark.parametrize("integrator", [1, 0])
def test_adjoint_matches_reference_gradient(oracle, gpu, integrator):
    import torch
    import grt_amd
    fields = small_torus()
    kw, po, ph0 = disk_setup(oracle, fields, integrator=integrator)
    n = len(ph0["x0"])
    w = np.random.default_rng(3).uniform(0.5, 1.5, n)
    steps = 4000 if integrator == 1 else 20000
    G = oracle.GridHolder(fields, 0.9, 1.0)
    pho = {k: v.copy() for k, v in ph0.items()}
    st_o, gr_o, gt_o, gc_o = oracle.trace_adjoint(po, G, pho, steps, w, nthreads=8)
    ctx = gpu_setup(kw, fields)
    g = grt_amd.to_device(ph0)
    st_g = ctx.trace_taped(g, steps, slots_per_photon=256)
    gr, gt, gc = ctx.adjoint(torch.from_numpy(w).to(gpu))
    gh = grt_amd.to_host(g)
    for k in ("x0", "x1", "x2", "x3", "k0", "k1", "k2", "k3", "phase"):
        assert np.array_equal(gh[k], pho[k]), k
    for k in ("I", "Q", "U", "V"):
        assert rel_close(gh[k], pho[k], 1e-13, 1e-300), k
    gr, gt, gc = gr.cpu().numpy(), gt.cpu().numpy(), gc.cpu().numpy()
    print("nonzero cells", np.count_nonzero(gr_o), "coeff", gc, gc_o, st_g, st_o)
    assert np.count_nonzero(gr_o) > 10
    assert np.array_equal(gr != 0, gr_o != 0)
    assert rel_close(gr, gr_o, 1e-10, 1e-14 * np.abs(gr_o).max())
    assert rel_close(gt, gt_o, 1e-10, 1e-14 * np.abs(gt_o).max())
    assert rel_close(gc, gc_o, 1e-10, 0.0)


def test_geometry_recording_build_gives_the_same_gradient(oracle, gpu):
    """The library built with GRT_GRADIENT_READS_GEOMETRY=1 (its adjoint also
    records each transfer step's position, time and magnetic field, for
    gradient companions that read them) returns the default build's gradient
    with the default companions, which read neither (summation order of the
    atomics aside)."""
    import torch
    import grt_amd
    fields = small_torus()
    kw, po, ph0 = disk_setup(oracle, fields)
    n = len(ph0["x0"])
    w = torch.from_numpy(np.random.default_rng(7).uniform(0.5, 1.5, n)).to(gpu)
    out = []
    for lib in (None, grt_amd.load(grt_amd.GEOMETRY_LIB_PATH)):
        ctx = grt_amd.Context(grt_amd.default_params(**kw), lib=lib)
        ctx.set_grid(fields)
        g = grt_amd.to_device(ph0)
        ctx.trace_taped(g, 4000, slots_per_photon=256)
        out.append([t.cpu().numpy() for t in ctx.adjoint(w)])
        ctx.close()
    (gr0, gt0, gc0), (gr1, gt1, gc1) = out
    print("nonzero cells", np.count_nonzero(gr0), "coeff", gc0, gc1)
    assert np.count_nonzero(gr0) > 10
    assert np.array_equal(gr0 != 0, gr1 != 0)
    assert rel_close(gr1, gr0, 1e-12, 1e-14 * np.abs(gr0).max())
    assert rel_close(gt1, gt0, 1e-12, 1e-14 * np.abs(gt0).max())
    assert rel_close(gc1, gc0, 1e-12, 0.0)


def test_adjoint_matches_gpu_finite_differences(oracle, gpu):
    import torch
    import grt_amd
    fields = small_torus()
    kw, po, ph0 = disk_setup(oracle, fields)
    n = len(ph0["x0"])
    w = torch.from_numpy(np.random.default_rng(5).uniform(0.5, 1.5, n)).to(gpu)
    ctx = gpu_setup(kw, fields)
    g0 = grt_amd.to_device(ph0)

    def L(rho, tg):
        ctx.set_fluid_scalars(rho, tg)
        g = {k: v.clone() for k, v in g0.items()}
        ctx.trace(g, 4000)
        return float((w * g["I"]).sum())

    g = {k: v.clone() for k, v in g0.items()}
    ctx.trace_taped(g, 4000)
    gr, gt, gc = ctx.adjoint(w)
    gr, gt = gr.cpu().numpy(), gt.cpu().numpy()
    rho0, t0 = fields["rho"], fields["Tgas"]
    base = {"rho": rho0, "Tgas": t0}
    L0 = L(rho0, t0)
    rng = np.random.default_rng(6)
    for name, grad in (("rho", gr), ("Tgas", gt)):
        for c in fd_cells(grad, base[name], L0, rng):
            v0 = base[name].ravel()[c]
            plus = {k: a.copy() for k, a in base.items()}
            minus = {k: a.copy() for k, a in base.items()}
            plus[name].ravel()[c] = v0 * (1 + H_CELL)
            minus[name].ravel()[c] = v0 * (1 - H_CELL)
            fd = (L(plus["rho"], plus["Tgas"]) - L(minus["rho"], minus["Tgas"])) / \
                (2 * H_CELL * v0)
            assert grad.ravel()[c] == pytest.approx(fd, rel=1e-5), (name, c)
    ctx.set_fluid_scalars(rho0, t0)


def test_adjoint_slot_overflow_and_sequence_errors(oracle, gpu):
    import torch
    import grt_amd
    fields = small_torus()
    kw, po, ph0 = disk_setup(oracle, fields)
    ctx = gpu_setup(kw, fields)
    n = len(ph0["x0"])
    with pytest.raises(grt_amd.GrtError, match="no taped trace"):
        ctx.adjoint(torch.ones(n, dtype=torch.float64, device=gpu))
    g = grt_amd.to_device(ph0)
    with pytest.raises(grt_amd.GrtError, match="slots_per_photon"):
        ctx.trace_taped(g, 4000, slots_per_photon=1)
    sc = gpu_setup(dict(kw, scattering_enabled=1), fields)
    with pytest.raises(grt_amd.GrtError, match="scattering"):
        sc.trace_taped(grt_amd.to_device(ph0), 4000)


def test_autograd_intensity_loss(oracle, gpu):
    """torch.autograd through transfer_intensity equals the direct adjoint of
    the same per-photon seeds."""
    import torch
    import grt_amd
    from grt_amd.autograd import transfer_intensity
    fields = small_torus()
    kw, po, ph0 = disk_setup(oracle, fields)
    ctx = gpu_setup(kw, fields)
    g0 = grt_amd.to_device(ph0)
    rho = torch.from_numpy(fields["rho"]).to(gpu).requires_grad_()
    tg = torch.from_numpy(fields["Tgas"]).to(gpu).requires_grad_()
    ka = torch.tensor(kw["absorption_coefficient"], dtype=torch.float64, requires_grad=True)
    ke = torch.tensor(kw["emission_coefficient"], dtype=torch.float64, requires_grad=True)
    out = {}
    I = transfer_intensity(ctx, g0, rho, tg, ka, ke, max_steps=4000, out=out)
    target = I.detach() * 0.9
    loss = torch.linalg.vector_norm(I - target)
    loss.backward()
    seed = ((I - target) / torch.linalg.vector_norm(I - target)).detach()
    g = {k: v.clone() for k, v in g0.items()}
    ctx.trace_taped(g, 4000)
    gr, gt, gc = ctx.adjoint(seed)
    assert torch.allclose(rho.grad, gr, rtol=1e-10, atol=1e-14 * gr.abs().max().item())
    assert torch.allclose(tg.grad, gt, rtol=1e-10, atol=1e-14 * gt.abs().max().item())
    assert ka.grad.item() == pytest.approx(gc[0].item(), rel=1e-10)
    assert ke.grad.item() == pytest.approx(gc[1].item(), rel=1e-10)
    # the photons themselves were not modified
    assert torch.equal(g0["x1"], grt_amd.to_device(ph0)["x1"])


def test_image_mode_gradient_is_zero(oracle, gpu):
    """Image rays never meet fluid after their emission event (their forward
    pass retraces the vacuum part of the backward path), so their gradient
    is zero, as the oracle's."""
    import torch
    import grt_amd
    from conftest import c1_fields, pixel_screen
    fields = c1_fields()
    kw = c1_params(theta=np.pi / 4)
    ctx = gpu_setup(kw, fields)
    su, sv = pixel_screen(32)
    g = grt_amd.empty_photons(len(su), device=gpu)
    ctx.init_image(g, 0.1, torch.from_numpy(su).to(gpu), torch.from_numpy(sv).to(gpu))
    ctx.trace_taped(g, 10000)
    gr, gt, gc = ctx.adjoint(torch.ones(len(su), dtype=torch.float64, device=gpu))
    assert not gr.any() and not gt.any() and not gc.any()


def test_tape_and_grid_invalidation(oracle, gpu):
    """A taped trace is replayed only while its grid, fluid and geometry
    stand: a new fluid or grid, or a parameter change other than the two
    transfer coefficients, drops it.  A spin or unit-scale change after
    set_grid leaves the cell records stale until the grid is set again."""
    import torch
    import grt_amd
    fields = small_torus()
    kw, po, ph0 = disk_setup(oracle, fields)
    ctx = gpu_setup(kw, fields)
    n = len(ph0["x0"])
    seed = torch.ones(n, dtype=torch.float64, device=gpu)

    def taped():
        ctx.trace_taped(grt_amd.to_device(ph0), 4000)

    taped()
    p = grt_amd.default_params(**dict(kw, absorption_coefficient=0.5))
    ctx.set_params(p)                       # coefficients only: the tape stands
    ctx.adjoint(seed)
    ctx.set_fluid_scalars(fields["rho"], fields["Tgas"])
    with pytest.raises(grt_amd.GrtError, match="no taped trace"):
        ctx.adjoint(seed)
    taped()
    ctx.set_params(grt_amd.default_params(**dict(kw, rtol=1e-7)))
    with pytest.raises(grt_amd.GrtError, match="no taped trace"):
        ctx.adjoint(seed)
    ctx.set_params(grt_amd.default_params(**kw))
    taped()
    ctx.set_grid(fields)
    with pytest.raises(grt_amd.GrtError, match="no taped trace"):
        ctx.adjoint(seed)
    # stale cell records
    ctx.set_params(grt_amd.default_params(**dict(kw, a_bh=0.5)))
    with pytest.raises(grt_amd.GrtError, match="call grt_set_grid again"):
        ctx.trace(grt_amd.to_device(ph0), 10)
    with pytest.raises(grt_amd.GrtError, match="call grt_set_grid again"):
        ctx.disk_photon_count()
    ctx.set_params(grt_amd.default_params(**dict(kw, density_g_cm3_per_code=2.0)))
    ctx.set_grid(fields)
    ctx.trace(grt_amd.to_device(ph0), 10)


def test_autograd_two_forwards_one_backward(oracle, gpu):
    """loss = f(x1) + f(x2) through one context: the first backward finds its
    tape replaced and raises instead of returning the second's gradient."""
    import torch
    import grt_amd
    from grt_amd.autograd import transfer_intensity
    fields = small_torus()
    kw, po, ph0 = disk_setup(oracle, fields)
    ctx = gpu_setup(kw, fields)
    g0 = grt_amd.to_device(ph0)
    rho = torch.from_numpy(fields["rho"]).to(gpu).requires_grad_()
    tg = torch.from_numpy(fields["Tgas"]).to(gpu)
    I1 = transfer_intensity(ctx, g0, rho, tg, 1.0, 1.0, max_steps=4000)
    I2 = transfer_intensity(ctx, g0, rho * 1.01, tg, 1.0, 1.0, max_steps=4000)
    with pytest.raises(grt_amd.GrtError, match="another taped trace"):
        (I1.sum() + I2.sum()).backward()


def test_taped_trace_with_several_photons_per_lane(oracle, gpu):
    """About 43,000 disk photons on 64-thread workgroups, one per CU (the
    launch never has more lanes than photons, so only more photons than
    lanes make a lane take several through its wavefront's stage): the taped
    kernel's photons equal the plain FP64 kernel's bit for bit, and the
    gradient equals that of 256-thread workgroups, two per CU, where every
    lane traces one photon (summation order of the atomics aside)."""
    import torch
    import grt_amd
    fields = small_torus()
    kw = c1_params(mode=1, superphoton_count_normalization=150.0, absorption_coefficient=0.05,
                   emission_coefficient=0.3, integrator=1, seed=7)
    ctx = gpu_setup(kw, fields)
    n = ctx.disk_photon_count()
    lanes = torch.cuda.get_device_properties(0).multi_processor_count * 64
    assert n > 2 * lanes, (n, lanes)
    ph0 = grt_amd.empty_photons(n, device=gpu)
    ctx.init_disk(ph0, 0.1)
    w = torch.from_numpy(np.random.default_rng(5).uniform(0.5, 1.5, n)).to(gpu)
    small = {"block": 64, "blocks_per_cu": 1, "sub_budget": 1048576}
    wide = {"block": 256, "blocks_per_cu": 2, "sub_budget": 1048576}
    ctx.tuning = small
    ref = {k: v.clone() for k, v in ph0.items()}
    ctx.trace(ref, 4000)
    grads = []
    for tune in (small, wide):
        ctx.tuning = tune
        g = {k: v.clone() for k, v in ph0.items()}
        ctx.trace_taped(g, 4000, slots_per_photon=256)
        for k in ("x0", "x1", "x2", "x3", "k0", "k1", "k2", "k3", "phase", "I", "Q"):
            assert torch.equal(g[k], ref[k]), (tune, k)
        grads.append([t.cpu().numpy() for t in ctx.adjoint(w)])
        del g
    (gr0, gt0, gc0), (gr1, gt1, gc1) = grads
    print("photons", n, "nonzero cells", np.count_nonzero(gr0), "coeff", gc0, gc1)
    assert np.count_nonzero(gr0) > 100
    assert np.array_equal(gr0 != 0, gr1 != 0)
    assert rel_close(gr0, gr1, 1e-12, 1e-14 * np.abs(gr1).max())
    assert rel_close(gt0, gt1, 1e-12, 1e-14 * np.abs(gt1).max())
    assert rel_close(gc0, gc1, 1e-12, 0.0)
    ctx.tuning = {"block": 256, "blocks_per_cu": 2, "sub_budget": 256}
