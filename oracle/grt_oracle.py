"""ctypes binding of the CPU oracle (oracle/libgrt_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
benchmark's CPU-baseline leg.  The product library never imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libgrt_oracle.so")


def build(force=False):
    """Compile the oracle with the Makefile next to this file."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE],
                       check=True)
    return _LIB_PATH


class Params(C.Structure):
    _fields_ = [
        ("a_bh", C.c_double), ("m_bh", C.c_double),
        ("r_term_min", C.c_double), ("r_term_max", C.c_double),
        ("atol", C.c_double), ("rtol", C.c_double),
        ("min_step_scale", C.c_double), ("max_step_scale", C.c_double),
        ("safety", C.c_double),
        ("integrator", C.c_int32), ("mode", C.c_int32),
        ("absorption_enabled", C.c_int32), ("emission_enabled", C.c_int32),
        ("absorption_coefficient", C.c_double), ("emission_coefficient", C.c_double),
        ("scattering_enabled", C.c_int32), ("albedo", C.c_double),
        ("seed", C.c_uint64), ("rng_stream", C.c_uint32),
        ("generator", C.c_int32),
        ("superphoton_count_normalization", C.c_double),
        ("energy_per_cell_erg", C.c_double), ("power_law_slope", C.c_double),
        ("nu_min_hz", C.c_double), ("nu_max_hz", C.c_double),
        ("custom_frequency_hz", C.c_double),
        ("camera_distance", C.c_double), ("camera_theta", C.c_double),
        ("camera_phi", C.c_double), ("camera_width", C.c_double),
        ("camera_height", C.c_double),
    ]


class Grid(C.Structure):
    _fields_ = [
        ("nr", C.c_int32), ("nth", C.c_int32), ("nph", C.c_int32),
        ("r", C.c_void_p), ("theta", C.c_void_p), ("phi", C.c_void_p),
        ("rho", C.c_void_p), ("tgas", C.c_void_p), ("vel", C.c_void_p),
        ("mag", C.c_void_p),
        ("length_scale", C.c_double), ("time_scale", C.c_double),
        ("density_scale", C.c_double), ("temperature_scale", C.c_double),
        ("velocity_scale", C.c_double), ("magnetic_scale", C.c_double),
        ("spin", C.c_double), ("mass", C.c_double),
        ("available", C.c_int32),
    ]


PHOTON_FIELDS = [
    ("id", np.uint64), ("frequency", np.float64), ("emission_frame_energy", np.float64),
    ("x0", np.float64), ("x1", np.float64), ("x2", np.float64), ("x3", np.float64),
    ("k0", np.float64), ("k1", np.float64), ("k2", np.float64), ("k3", np.float64),
    ("I", np.float64), ("Q", np.float64), ("U", np.float64), ("V", np.float64),
    ("dlambda", np.float64), ("terminate", np.uint8), ("phase", np.uint8),
    ("theta_disp", np.float64), ("phi_disp", np.float64),
]


class Photons(C.Structure):
    _fields_ = [(name, C.c_void_p) for name, _ in PHOTON_FIELDS]


class Fluid(C.Structure):
    _fields_ = [
        ("position_cm", C.c_double * 3), ("coordinate_time_s", C.c_double),
        ("density_g_cm3", C.c_double), ("temperature_k", C.c_double),
        ("magnetic_gauss", C.c_double * 4),
        ("coordinate_four_velocity_cm_s", C.c_double * 4),
    ]


class Stats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "functor_calls", "accepted_steps", "rejected_attempts", "failed_steps",
        "fluid_samples", "global_steps", "terminated", "fluid_hits")]


class Rng(C.Structure):
    _fields_ = [("id", C.c_uint64), ("draw", C.c_uint32), ("stream", C.c_uint32),
                ("seed", C.c_uint64)]


_D = C.POINTER(C.c_double)
_RHS_FN = C.CFUNCTYPE(None, _D, _D, C.c_void_p)

_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(_LIB_PATH)
        P, G, PH, F, S = (C.POINTER(Params), C.POINTER(Grid), C.POINTER(Photons),
                          C.POINTER(Fluid), C.POINTER(Stats))
        sig = {
            "orc_compute_r": (C.c_double, [C.c_double] * 4),
            "orc_compute_r_and_deriv": (None, [C.c_double] * 4 + [_D, _D]),
            "orc_compute_metric": (None, [_D, C.c_double, C.c_double, _D]),
            "orc_compute_inverse_metric": (None, [_D, C.c_double, C.c_double, _D]),
            "orc_compute_inverse_metric_deriv": (None, [_D, C.c_double, C.c_double, _D]),
            "orc_hamiltonian_rhs": (None, [_D, _D, C.c_double, C.c_double, _D]),
            "orc_hamiltonian_rhs_fused": (None, [_D, C.c_double, C.c_double, _D]),
            "orc_set_unfused_rounding": (None, [C.c_int]),
            "orc_null_covariant_momentum": (C.c_int, [_D, _D, C.c_double, C.c_double, _D]),
            "orc_inverse_norm": (C.c_double, [_D, _D, C.c_double, C.c_double]),
            "orc_geodesic_observables": (None, [_D, C.c_double, C.c_double, _D]),
            "orc_rk4_generic": (None, [C.c_int, _D, C.c_double, _RHS_FN, C.c_void_p]),
            "orc_rk45_generic": (C.c_int, [C.c_int, _D, _D, _D, _RHS_FN, C.c_void_p]
                                 + [C.c_double] * 5),
            "orc_rk4_step_ks": (None, [_D, C.c_double, C.c_double, C.c_double]),
            "orc_rk45_step_ks": (C.c_int, [P, _D, _D, _D]),
            "orc_quarter_root": (C.c_double, [C.c_double]),
            "orc_locate_cell": (C.c_int, [G, C.c_double, C.c_double, C.c_double,
                                          C.POINTER(C.c_int32)]),
            "orc_spherical_to_cartesian_vector": (None, [C.c_double] * 4 + [_D, _D]),
            "orc_build_fluid_tetrad": (C.c_int, [_D, _D, C.c_double, C.c_double, _D]),
            "orc_sample_fluid": (C.c_int, [_D, G, F, _D]),
            "orc_apply_radiative_transfer": (None, [_D, F, C.c_double, C.c_double, C.c_int,
                                                    C.c_int, C.c_double, C.c_double]),
            "orc_philox4x32_10": (None, [C.POINTER(C.c_uint32)] * 3),
            "orc_rng_drand": (C.c_double, [C.POINTER(Rng)]),
            "orc_maybe_scatter": (C.c_int, [_D, _D, C.c_int, C.c_double, C.c_double,
                                            C.c_double, F, _D, C.c_double, C.c_double,
                                            C.POINTER(Rng)]),
            "orc_superphotons_per_cell": (C.c_int, [C.c_double, C.c_double]),
            "orc_generate_photon": (C.c_int, [P, F, C.POINTER(Rng), C.c_int, _D, _D]),
            "orc_sample_generator": (None, [P, F, C.c_int, C.c_int64, _D, _D]),
            "orc_camera_plane_crossing": (C.c_int, [P, _D, _D, _D]),
            "orc_camera_observer_energy": (C.c_double, [P, _D]),
            "orc_step_photon": (None, [P, G, PH, C.c_int64, C.POINTER(C.c_uint32), S]),
            "orc_trace": (None, [P, G, PH, C.c_int64, C.c_int32, C.c_double, C.c_int64,
                                 C.c_int32, S]),
            "orc_trace_adjoint": (None, [P, G, PH, C.c_int64, C.c_int32, C.c_double, C.c_int64,
                                         C.c_int32, _D, _D, _D, _D, S]),
            "orc_optical_depth_increment_grad": (C.c_double, [F, C.c_double, C.c_double,
                                                              C.c_double, _D]),
            "orc_emissivity_increment_grad": (C.c_double, [F, C.c_double, C.c_double,
                                                           C.c_double, _D]),
            "orc_init_image_photon": (None, [P, C.c_double, C.c_double, C.c_double, _D, _D]),
            "orc_init_image": (None, [P, _D, _D, C.c_int64, C.c_uint64, C.c_double, PH]),
            "orc_init_disk": (None, [P, G, C.c_int64, C.c_uint64, C.c_double, PH,
                                     C.POINTER(C.c_uint64)]),
            "orc_observation_bin": (C.c_int64, [C.c_double, C.c_double, C.c_double, C.c_int64]),
            "orc_image_products": (None, [PH, C.c_int64, C.c_int64, C.c_int64, C.c_int64,
                                          C.c_double, C.c_double, C.c_double, C.c_double,
                                          _D, _D, _D, C.POINTER(C.c_uint64)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def dptr(a):
    """Pointer to the data of a contiguous float64 numpy array."""
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


# Defaults of the reference's runtime globals (utils.cpp:4-37, utils.hpp:178-180,
# simulation_options.hpp:13-32, photon_generation.hpp:40-51).
DEFAULTS = dict(
    a_bh=1.0, m_bh=1.0, r_term_min=1.0, r_term_max=15000.0,
    atol=1e-8, rtol=1e-8, min_step_scale=0.01, max_step_scale=10.0, safety=0.9,
    integrator=1, mode=0, absorption_enabled=0, emission_enabled=0,
    absorption_coefficient=0.0, emission_coefficient=0.0,
    scattering_enabled=0, albedo=1.0, seed=12345, rng_stream=0,
    generator=0, superphoton_count_normalization=1.0, energy_per_cell_erg=1.0,
    power_law_slope=2.0, nu_min_hz=1e9, nu_max_hz=1e20, custom_frequency_hz=2.30e11,
    camera_distance=10000.0, camera_theta=np.pi / 4, camera_phi=np.pi / 4,
    camera_width=20.0, camera_height=20.0,
)


def make_params(**kw):
    d = dict(DEFAULTS)
    for k, v in kw.items():
        if k not in d:
            raise KeyError(k)
        d[k] = v
    p = Params()
    for k, v in d.items():
        setattr(p, k, v)
    return p


class GridHolder:
    """Keeps the numpy arrays alive behind an orc_grid."""

    def __init__(self, fields, spin, mass, units=None, available=True):
        u = dict(length_cm_per_code=1.0, time_s_per_code=1.0, density_g_cm3_per_code=1.0,
                 temperature_k_per_code=1.0, four_velocity_cm_s_per_code=1.0,
                 magnetic_gauss_per_code=1.0)
        if units:
            u.update(units)
        self.arrays = {k: np.ascontiguousarray(fields[k], dtype=np.float64)
                       for k in ("r", "theta", "phi", "rho", "Tgas", "vel", "mag")}
        nr, nth, nph = self.arrays["r"].shape
        g = Grid()
        g.nr, g.nth, g.nph = nr, nth, nph
        g.r = self.arrays["r"].ctypes.data
        g.theta = self.arrays["theta"].ctypes.data
        g.phi = self.arrays["phi"].ctypes.data
        g.rho = self.arrays["rho"].ctypes.data
        g.tgas = self.arrays["Tgas"].ctypes.data
        g.vel = self.arrays["vel"].ctypes.data
        g.mag = self.arrays["mag"].ctypes.data
        g.length_scale = u["length_cm_per_code"]
        g.time_scale = u["time_s_per_code"]
        g.density_scale = u["density_g_cm3_per_code"]
        g.temperature_scale = u["temperature_k_per_code"]
        g.velocity_scale = u["four_velocity_cm_s_per_code"]
        g.magnetic_scale = u["magnetic_gauss_per_code"]
        g.spin, g.mass = spin, mass
        g.available = 1 if available else 0
        self.grid = g


def empty_photons(n):
    return {name: np.zeros(n, dtype=dt) for name, dt in PHOTON_FIELDS}


def photon_struct(ph):
    s = Photons()
    for name, dt in PHOTON_FIELDS:
        a = ph[name]
        assert a.dtype == dt and a.flags.c_contiguous, name
        setattr(s, name, a.ctypes.data)
    return s


def trace(params, grid_holder, ph, max_steps, termination_percent=1.0, output_interval=0,
          nthreads=0):
    n = len(ph["x0"])
    st = Stats()
    g = grid_holder.grid if grid_holder is not None else Grid()
    lib().orc_trace(C.byref(params), C.byref(g), C.byref(photon_struct(ph)), n,
                    max_steps, termination_percent, output_interval, nthreads, C.byref(st))
    return {name: getattr(st, name) for name, _ in Stats._fields_}


def trace_adjoint(params, grid_holder, ph, max_steps, dL_dI, termination_percent=1.0,
                  output_interval=0, nthreads=0):
    """Trace like trace() and return (stats, dL/drho, dL/dTgas, [dL/dkappa_abs,
    dL/dkappa_emit]) for the loss whose gradient with respect to each photon's
    final Stokes I is dL_dI (the reference gradient, no checkpointing)."""
    n = len(ph["x0"])
    st = Stats()
    g = grid_holder.grid
    shape = (g.nr, g.nth, g.nph)
    gr = np.zeros(shape)
    gt = np.zeros(shape)
    gc = np.zeros(2)
    seeds = np.ascontiguousarray(dL_dI, dtype=np.float64)
    assert seeds.shape == (n,)
    lib().orc_trace_adjoint(C.byref(params), C.byref(g), C.byref(photon_struct(ph)), n,
                            max_steps, termination_percent, output_interval, nthreads,
                            dptr(seeds), dptr(gr), dptr(gt), dptr(gc), C.byref(st))
    return {name: getattr(st, name) for name, _ in Stats._fields_}, gr, gt, gc


def increment_grads(fluid, nu, dl, kabs, kemit):
    """(dtau, [d/drho, d/dT, d/dkappa]), (dI, [...]) of the default increments."""
    f = Fluid(**fluid) if isinstance(fluid, dict) else fluid
    da = np.zeros(3)
    de = np.zeros(3)
    t = lib().orc_optical_depth_increment_grad(C.byref(f), nu, dl, kabs, dptr(da))
    e = lib().orc_emissivity_increment_grad(C.byref(f), nu, dl, kemit, dptr(de))
    return (t, da), (e, de)


def sample_generator(params, fluid, packets, n):
    """n draws (frequency, packet energy) of the configured generator for one
    fluid cell (dict of Fluid fields); draw i from photon id i's stream."""
    f = Fluid(**fluid) if isinstance(fluid, dict) else fluid
    nu = np.zeros(n)
    en = np.zeros(n)
    lib().orc_sample_generator(C.byref(params), C.byref(f), packets, n, dptr(nu), dptr(en))
    return nu, en


def init_image(params, su, sv, initial_step, id_offset=0):
    su = np.ascontiguousarray(su, dtype=np.float64)
    sv = np.ascontiguousarray(sv, dtype=np.float64)
    ph = empty_photons(len(su))
    lib().orc_init_image(C.byref(params), dptr(su), dptr(sv), len(su), id_offset, initial_step,
                         C.byref(photon_struct(ph)))
    return ph


def disk_count(params, grid_holder):
    tot = C.c_uint64(0)
    lib().orc_init_disk(C.byref(params), C.byref(grid_holder.grid), 0, 0, 0.0, None,
                        C.byref(tot))
    return tot.value


def init_disk(params, grid_holder, initial_step, id_offset=0):
    n = disk_count(params, grid_holder)
    ph = empty_photons(n)
    tot = C.c_uint64(0)
    lib().orc_init_disk(C.byref(params), C.byref(grid_holder.grid), n, id_offset, initial_step,
                        C.byref(photon_struct(ph)), C.byref(tot))
    return ph


def image_products(ph, nx, ny, bins, smin, smax, width, height):
    n = len(ph["x0"])
    img = np.zeros(4 * nx * ny)
    snu = np.zeros(bins)
    sI = np.zeros(bins)
    arr = C.c_uint64(0)
    lib().orc_image_products(C.byref(photon_struct(ph)), n, nx, ny, bins, smin, smax, width,
                             height, dptr(img), dptr(snu), dptr(sI), C.byref(arr))
    return dict(image=img.reshape(4, ny, nx), spectrum_frequency_hz=snu, spectrum_I=sI,
                arrived_count=arr.value)


def rhs(X, p, a, m):
    X = np.ascontiguousarray(X, dtype=np.float64)
    p = np.ascontiguousarray(p, dtype=np.float64)
    out = np.zeros(8)
    lib().orc_hamiltonian_rhs(dptr(X), dptr(p), a, m, dptr(out))
    return out


def quarter_root(x):
    """The step controller's x^(1/4) (the candidate check of the oracle)."""
    return lib().orc_quarter_root(float(x))


def set_unfused_rounding(unfused):
    """Switch the integrators between the fused rounding (default, what the
    tracing kernel reproduces) and the reference's expressions unfused."""
    lib().orc_set_unfused_rounding(1 if unfused else 0)


def rhs_fused(X, p, a, m):
    """The right-hand side as the integrators evaluate it (fused rounding)."""
    s = np.ascontiguousarray(np.concatenate([X, p]), dtype=np.float64)
    out = np.zeros(8)
    lib().orc_hamiltonian_rhs_fused(dptr(s), a, m, dptr(out))
    return out


def metric(X, a, m):
    X = np.ascontiguousarray(X, dtype=np.float64)
    g = np.zeros(16)
    lib().orc_compute_metric(dptr(X), a, m, dptr(g))
    return g.reshape(4, 4)


def inverse_metric(X, a, m):
    X = np.ascontiguousarray(X, dtype=np.float64)
    g = np.zeros(16)
    lib().orc_compute_inverse_metric(dptr(X), a, m, dptr(g))
    return g.reshape(4, 4)


def inverse_metric_deriv(X, a, m):
    X = np.ascontiguousarray(X, dtype=np.float64)
    g = np.zeros(64)
    lib().orc_compute_inverse_metric_deriv(dptr(X), a, m, dptr(g))
    return g.reshape(4, 4, 4)


def null_momentum(X, K, a, m):
    X = np.ascontiguousarray(X, dtype=np.float64)
    K = np.ascontiguousarray(K, dtype=np.float64)
    p = np.zeros(4)
    ok = lib().orc_null_covariant_momentum(dptr(X), dptr(K), a, m, dptr(p))
    return bool(ok), p


def observables(state, a, m):
    s = np.ascontiguousarray(state, dtype=np.float64)
    out = np.zeros(4)
    lib().orc_geodesic_observables(dptr(s), a, m, dptr(out))
    return dict(norm=out[0], energy=out[1], angular_momentum_z=out[2], carter=out[3])


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)
