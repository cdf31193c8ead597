"""Host-side checks of the Python binding that need no GPU: photon arrays are
validated before any pointer reaches the library (a short, mistyped or
host-resident array would otherwise be read out of bounds by a kernel), and
the autograd binding refuses to replay a tape that a later forward pass on
the same context replaced."""
import numpy as np
import pytest
import torch

import grt_amd
from grt_amd import GrtError, photon_struct


def host_photons(n=8):
    return grt_amd.empty_photons(n)


def torch_photons(n=8):
    tdt = {np.dtype(np.float64): torch.float64, np.dtype(np.uint64): torch.int64,
           np.dtype(np.uint8): torch.uint8}
    return {name: torch.zeros(n, dtype=tdt[np.dtype(dt)]) for name, dt in grt_amd.PHOTON_FIELDS}


def test_valid_host_photons_pass():
    ph = host_photons()
    s = photon_struct(ph)
    assert s.x0 == ph["x0"].ctypes.data and s.phase == ph["phase"].ctypes.data


@pytest.mark.parametrize("field", ["x1", "id", "terminate"])
def test_short_host_field_is_refused(field):
    ph = host_photons()
    ph[field] = ph[field][:5].copy()
    with pytest.raises(GrtError, match=f"{field} must be one-dimensional of length 8"):
        photon_struct(ph)


def test_wrong_host_dtype_is_refused():
    ph = host_photons()
    ph["I"] = ph["I"].astype(np.float32)
    with pytest.raises(GrtError, match="I must be a numpy float64"):
        photon_struct(ph)
    ph = host_photons()
    ph["phase"] = ph["phase"].astype(np.int32)
    with pytest.raises(GrtError, match="phase must be a numpy uint8"):
        photon_struct(ph)


def test_two_dimensional_host_field_is_refused():
    ph = host_photons()
    ph["k2"] = np.zeros((8, 1))
    with pytest.raises(GrtError, match="k2 must be one-dimensional"):
        photon_struct(ph)


def test_cpu_torch_tensors_are_refused():
    """A CPU tensor has a data_ptr, so it would select the device entry points
    and hand a host address to a kernel."""
    with pytest.raises(GrtError, match="CPU tensor"):
        photon_struct(torch_photons())


def test_mixed_numpy_and_torch_is_refused():
    ph = host_photons()
    ph["x2"] = torch.zeros(8, dtype=torch.float64)
    with pytest.raises(GrtError, match="all numpy arrays or all torch tensors"):
        photon_struct(ph)


class _FakeContext:
    """Stands in for grt_amd.Context in the autograd binding: a taped trace
    sets I to the current emission coefficient and counts generations, as the
    real context does."""

    def __init__(self):
        self.params = grt_amd.Params()
        self._tape_gen = 0
        self.adjoint_calls = 0

    def set_fluid_scalars(self, rho, tgas):
        self.rho = rho

    def set_params(self, p):
        self.params = p

    def trace_taped(self, ph, max_steps, slots):
        ph["I"].fill_(self.params.emission_coefficient)
        self._tape_gen += 1
        return {}

    @property
    def tape_generation(self):
        return self._tape_gen

    def adjoint(self, seed):
        self.adjoint_calls += 1
        g = torch.zeros_like(self.rho)
        return g, g.clone(), torch.tensor([0.0, float(seed.sum())], dtype=torch.float64)


def _forward(ctx, ke):
    from grt_amd.autograd import transfer_intensity
    ph = torch_photons(4)
    rho = torch.ones(2, 2, 2, dtype=torch.float64, requires_grad=True)
    tg = torch.ones(2, 2, 2, dtype=torch.float64, requires_grad=True)
    ka = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    ke = torch.tensor(ke, dtype=torch.float64, requires_grad=True)
    return transfer_intensity(ctx, ph, rho, tg, ka, ke), ke


def test_autograd_backward_of_the_last_forward():
    ctx = _FakeContext()
    I, ke = _forward(ctx, 2.0)
    I.sum().backward()
    assert ctx.adjoint_calls == 1 and ke.grad.item() == 4.0


def test_autograd_refuses_a_replaced_tape():
    """loss = f(x1) + f(x2) on one context: the first forward's tape is gone
    when its backward runs, so backward raises instead of replaying the
    second forward's tape."""
    ctx = _FakeContext()
    I1, _ = _forward(ctx, 2.0)
    I2, _ = _forward(ctx, 3.0)
    with pytest.raises(GrtError, match="another taped trace ran on this context"):
        (I1.sum() + I2.sum()).backward()
