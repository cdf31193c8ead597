"""torch.autograd binding of the differentiable form.

`transfer_intensity(context, photons0, rho, tgas, kappa_abs, kappa_emit)`
traces a fixed set of initial photons (dict of GPU tensors, left untouched)
through the context's grid with the given density, temperature and
coefficients, and returns the photons' final Stokes I as a tensor that
autograd differentiates with respect to rho, tgas (shaped like the grid)
and both coefficients (0-d tensors).  The forward pass is the FP64 tracing
kernel with geodesic checkpoints (grt_trace_taped_device); the backward pass
is the adjoint kernel (grt_adjoint_device), which replays the checkpointed
geodesics and sweeps the radiative-transfer recurrence backwards.

`image_intensity(I, photons, nx, ny, width, height)` bins arrived photons'
I into the nx x ny Stokes-I image with the observation rule of
build_observation_products (observation.hpp:22-84), differentiably in I.
"""
import torch

from . import PHASE_ARRIVED, GrtError


class _TransferIntensity(torch.autograd.Function):
    @staticmethod
    def forward(fctx, rho, tgas, kappa_abs, kappa_emit, context, photons0, max_steps, slots, out):
        context.set_fluid_scalars(rho.detach(), tgas.detach())
        p = context.params
        p.absorption_coefficient = float(kappa_abs)
        p.emission_coefficient = float(kappa_emit)
        context.set_params(p)
        ph = {k: v.clone() for k, v in photons0.items()}
        stats = context.trace_taped(ph, max_steps, slots)
        fctx.context = context
        fctx.generation = context.tape_generation
        if out is not None:
            out["photons"] = ph
            out["stats"] = stats
        return ph["I"].clone()

    @staticmethod
    def backward(fctx, grad_I):
        # The context holds one tape: the last taped trace's, with the fluid
        # and parameters it set.  A later forward on the same context replaced
        # it, and replaying that one would return another problem's gradient.
        if fctx.context.tape_generation != fctx.generation:
            raise GrtError("transfer_intensity: another taped trace ran on this context after "
                           "this forward pass; run backward before the next forward on the "
                           "same context, or use one context per forward pass")
        gr, gt, gc = fctx.context.adjoint(grad_I.contiguous())
        return gr, gt, gc[0], gc[1], None, None, None, None, None


def transfer_intensity(context, photons0, rho, tgas, kappa_abs, kappa_emit, max_steps=10000,
                       slots_per_photon=64, out=None):
    """Final Stokes I of `photons0` traced through (rho, tgas, kappa_abs,
    kappa_emit); differentiable in those four.  `out`, if a dict, receives the
    final photons ("photons") and the trace statistics ("stats")."""
    kappa_abs = torch.as_tensor(kappa_abs, dtype=torch.float64)
    kappa_emit = torch.as_tensor(kappa_emit, dtype=torch.float64)
    return _TransferIntensity.apply(rho, tgas, kappa_abs, kappa_emit, context, photons0,
                                    max_steps, slots_per_photon, out)


def _bins(v, lo, hi, n):
    if n <= 1 or not hi > lo:
        return torch.zeros_like(v, dtype=torch.int64)
    fr = ((v - lo) / (hi - lo)).clamp(0.0, 1.0)
    return (fr * float(n - 1)).to(torch.int64)


def image_intensity(I, photons, nx, ny, width, height):
    """Stokes-I image [ny, nx] of the arrived photons, differentiable in I."""
    arrived = photons["phase"] == PHASE_ARRIVED
    px = _bins(photons["theta_disp"], -0.5 * width, 0.5 * width, nx)
    py = _bins(photons["phi_disp"], -0.5 * height, 0.5 * height, ny)
    pix = (py * nx + px)[arrived]
    img = torch.zeros(nx * ny, dtype=I.dtype, device=I.device)
    img = img.index_add(0, pix, I[arrived])
    return img.view(ny, nx)
