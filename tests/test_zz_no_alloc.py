"""grt_trace_device_noalloc (the trace with the fixed device scratch, no
allocation) against grt_trace_device on the same photons: every photon field
and every statistic bit for bit, on two calls in a row, the second smaller
than the first (so stale scratch contents beyond its photons cannot leak
in), in image mode (deferred thermal emission, the arrival redshift factor)
and in disk mode with scattering (the per-photon RNG draw counters)."""
import ctypes as C

import numpy as np
import pytest

from conftest import c1_fields, c1_params, pixel_screen

pytestmark = pytest.mark.gpu

STATS = ("functor_calls", "accepted_steps", "rejected_attempts", "failed_steps",
         "fluid_samples", "fluid_hits", "global_steps", "terminated")


def _same(a, b, st_a, st_b):
    import grt_amd
    for name, _ in grt_amd.PHOTON_FIELDS:
        assert np.array_equal(a[name], b[name], equal_nan=True), name
    for k in STATS:
        assert st_a[k] == st_b[k], k


def _image_batch(ctx, n_side, gpu):
    import torch
    import grt_amd
    su, sv = pixel_screen(n_side)
    ph = grt_amd.empty_photons(len(su), device=gpu)
    ctx.init_image(ph, 0.1, torch.from_numpy(su).to(gpu), torch.from_numpy(sv).to(gpu))
    return ph


def test_noalloc_equals_trace_device_image(gpu):
    import torch
    import grt_amd
    fields = c1_fields()
    ctx = grt_amd.Context(grt_amd.default_params(**c1_params()))
    ctx.set_grid(fields)
    for n_side in (64, 24):           # the second call is smaller
        ph = _image_batch(ctx, n_side, gpu)
        a = {k: v.clone() for k, v in ph.items()}
        b = {k: v.clone() for k, v in ph.items()}
        st_a = ctx.trace(a, 10000)
        st_b = ctx.trace_noalloc(b, 10000)
        torch.cuda.synchronize()
        print(n_side, st_a)
        assert st_b["terminated"] == n_side * n_side
        _same(grt_amd.to_host(a), grt_amd.to_host(b), st_a, st_b)


def test_noalloc_equals_trace_device_scattering(gpu):
    import torch
    import grt_amd
    from test_oracle_adjoint import small_torus
    fields = small_torus()
    kw = c1_params(mode=1, r_term_max=200.0, superphoton_count_normalization=6.0,
                   absorption_coefficient=0.05, emission_coefficient=0.3, scattering_enabled=1,
                   albedo=0.8, seed=77, density_g_cm3_per_code=4.0)
    ctx = grt_amd.Context(grt_amd.default_params(**kw))
    ctx.set_grid(fields)
    total = ctx.disk_photon_count()
    assert total > 100
    for n in (total, total // 3):
        ph = grt_amd.empty_photons(n, device=gpu)
        ctx.init_disk(ph, 0.1)
        a = {k: v.clone() for k, v in ph.items()}
        b = {k: v.clone() for k, v in ph.items()}
        st_a = ctx.trace(a, 4000)
        st_b = ctx.trace_noalloc(b, 4000)
        torch.cuda.synchronize()
        print(n, st_a)
        _same(grt_amd.to_host(a), grt_amd.to_host(b), st_a, st_b)


def test_noalloc_refuses_more_than_its_scratch(gpu):
    """More photons than the fixed scratch holds is refused before any
    photon array is read."""
    import grt_amd
    ctx = grt_amd.Context(grt_amd.default_params(**c1_params()))
    ph = grt_amd.empty_photons(4, device=gpu)
    s = grt_amd.photon_struct(ph)
    st = grt_amd.Stats()
    rc = grt_amd.load().grt_trace_device_noalloc(ctx._h, C.byref(s),
                                                 grt_amd.NOALLOC_MAX_PHOTONS + 1, 10, 1.0, 0,
                                                 None, C.byref(st))
    assert rc == -6   # GRT_E_RANGE
    assert "at most 1048576 photons" in grt_amd.load().grt_last_error(ctx._h).decode()
