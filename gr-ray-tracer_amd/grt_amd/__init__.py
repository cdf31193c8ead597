"""Python host binding of libgrt_amd.so (the HIP tracing path, C ABI in include/grt.h).

This mirrors the reference's host interface for the traced path:
`Context(params)` stands for the Geodesic_cartesian_kerr_schild functor's
captured constants (integrate_cart_ks.hpp:63-94), `set_grid` for the
.npy field bundle (load_python_arrays.hpp:195-223), `trace` for
integrate_geodesics (integrate_cart_ks.hpp:423-535) and `image_products`
for build_observation_products (observation.hpp:31-84).  Photons are the
20-field structure of arrays of utils.hpp:40-54, given as a dict of numpy
arrays (host) or of torch tensors on the GPU (device, no copies).

Every call goes to the HIP library; there is no CPU fallback.  Loading fails
loudly if libgrt_amd.so has not been built.
"""
import ctypes as C
import os

import numpy as np

_PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("GRT_AMD_LIB") or os.path.join(_PKG_ROOT, "libgrt_amd.so")
# The build for ...Gradient companions that read the sample's position, time or
# magnetic field (GRT_GRADIENT_READS_GEOMETRY=1), built beside the default one.
GEOMETRY_LIB_PATH = os.path.join(_PKG_ROOT, "libgrt_amd_geometry.so")

GRT_RK4, GRT_RK45 = 0, 1
MODE_IMAGE, MODE_DISK = 0, 1
GEN_BLACKBODY, GEN_POWER_LAW, GEN_CUSTOM = 0, 1, 2
PHASE_DISK, PHASE_BACKWARD, PHASE_FORWARD, PHASE_ARRIVED, PHASE_REJECTED = range(5)
PRECISION_FP64, PRECISION_MIXED, PRECISION_FP64_FAST_RHS = 0, 1, 2
NOALLOC_MAX_PHOTONS = 1 << 20     # GRT_NOALLOC_MAX_PHOTONS of grt.h
ABI_VERSION = 2                   # GRT_ABI_VERSION of grt.h (the Stats layout below)


class GrtError(RuntimeError):
    pass


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
        ("length_cm_per_code", C.c_double), ("time_s_per_code", C.c_double),
        ("density_g_cm3_per_code", C.c_double), ("temperature_k_per_code", C.c_double),
        ("four_velocity_cm_s_per_code", C.c_double), ("magnetic_gauss_per_code", C.c_double),
        ("precision", C.c_int32),
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


class Stats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "functor_calls", "accepted_steps", "rejected_attempts", "failed_steps",
        "fluid_samples", "fluid_hits", "global_steps", "terminated", "launches")] + \
        [("kernel_ms", C.c_double)]


class Timers(C.Structure):
    """grt_timers: the reference's per-phase Timers of integrate_geodesics."""
    _fields_ = [("geodesic_integration_s", C.c_double), ("termination_count_s", C.c_double),
                ("output_s", C.c_double), ("archives", C.c_uint64),
                ("last_archive_step", C.c_int64)]


class Tuning(C.Structure):
    _fields_ = [("block", C.c_int32), ("blocks_per_cu", C.c_int32), ("sub_budget", C.c_int32)]


EXPORTED = [
    "grt_params_default", "grt_create", "grt_destroy", "grt_set_params", "grt_set_grid",
    "grt_trace", "grt_trace_device", "grt_trace_device_noalloc", "grt_image_products",
    "grt_image_products_device",
    "grt_init_image", "grt_init_image_device", "grt_disk_photon_count", "grt_init_disk",
    "grt_init_disk_device", "grt_trace_taped_device", "grt_adjoint_device",
    "grt_set_fluid_scalars", "grt_set_fluid_scalars_device",
    "grt_load_config", "grt_read_npy", "grt_load_numpy_fields", "grt_write_image_products_npz",
    "grt_write_photons_npz", "grt_integrate_device",
    "grt_get_tuning", "grt_set_tuning", "grt_last_error", "grt_abi_version",
]

_lib = None


def load(path=None):
    """Load libgrt_amd.so (raises GrtError if it has not been built).  path:
    another build of the library, such as GEOMETRY_LIB_PATH, loaded beside
    it (not cached; pass the handle to Context)."""
    global _lib
    if path is not None:
        return _bind(path)
    if _lib is None:
        _lib = _bind(LIB_PATH)
    return _lib


def _bind(lib_path):
    """A handle on the library at lib_path with every entry point's signature."""
    if not os.path.exists(lib_path):
        raise GrtError(f"{lib_path} is missing: build it with "
                       "python gr-ray-tracer_amd/grt_amd/build.py")
    L = C.CDLL(lib_path)
    P, PH, ST, TU = C.POINTER(Params), C.POINTER(Photons), C.POINTER(Stats), C.POINTER(Tuning)
    vp, dp = C.c_void_p, C.POINTER(C.c_double)
    sig = {
        "grt_params_default": (None, [P]),
        "grt_create": (C.c_int, [P, C.POINTER(vp)]),
        "grt_destroy": (C.c_int, [vp]),
        "grt_set_params": (C.c_int, [vp, P]),
        "grt_set_grid": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32] + [vp] * 7),
        "grt_trace": (C.c_int, [vp, PH, C.c_uint64, C.c_int32, C.c_double, C.c_int64, ST]),
        "grt_trace_device": (C.c_int, [vp, PH, C.c_uint64, C.c_int32, C.c_double, C.c_int64,
                                       vp, ST]),
        "grt_trace_device_noalloc": (C.c_int, [vp, PH, C.c_uint64, C.c_int32, C.c_double,
                                               C.c_int64, vp, ST]),
        # the run config pointer is a grt_amd.io.RunConfig (passed by address)
        "grt_integrate_device": (C.c_int, [vp, PH, C.c_uint64, vp, vp, ST,
                                           C.POINTER(Timers)]),
        "grt_image_products": (C.c_int, [vp, PH, C.c_uint64, C.c_int64, C.c_int64, C.c_int64,
                                         C.c_double, C.c_double, vp, vp, vp, vp]),
        "grt_image_products_device": (C.c_int, [vp, PH, C.c_uint64, C.c_int64, C.c_int64,
                                                C.c_int64, C.c_double, C.c_double, vp, vp, vp,
                                                vp, vp]),
        "grt_init_image": (C.c_int, [vp, PH, C.c_uint64, C.c_uint64, C.c_double, vp, vp]),
        "grt_init_image_device": (C.c_int, [vp, PH, C.c_uint64, C.c_uint64, C.c_double, vp, vp,
                                            vp]),
        "grt_disk_photon_count": (C.c_int, [vp, C.POINTER(C.c_uint64)]),
        "grt_init_disk": (C.c_int, [vp, PH, C.c_uint64, C.c_uint64, C.c_double]),
        "grt_init_disk_device": (C.c_int, [vp, PH, C.c_uint64, C.c_uint64, C.c_double, vp]),
        "grt_trace_taped_device": (C.c_int, [vp, PH, C.c_uint64, C.c_int32, C.c_double,
                                             C.c_int64, C.c_int32, vp, ST]),
        "grt_adjoint_device": (C.c_int, [vp, vp, C.c_uint64, vp, vp, vp, vp]),
        "grt_set_fluid_scalars": (C.c_int, [vp, vp, vp]),
        "grt_set_fluid_scalars_device": (C.c_int, [vp, vp, vp, vp]),
        "grt_get_tuning": (C.c_int, [vp, TU]),
        "grt_set_tuning": (C.c_int, [vp, TU]),
        "grt_last_error": (C.c_char_p, [vp]),
        "grt_abi_version": (C.c_int, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    if L.grt_abi_version() != ABI_VERSION:
        raise GrtError(f"{lib_path} has ABI version {L.grt_abi_version()}, this binding "
                       f"version {ABI_VERSION}: rebuild the library")
    return L


def default_params(**overrides):
    """grt_params filled with the reference defaults, then `overrides`."""
    p = Params()
    load().grt_params_default(C.byref(p))
    for k, v in overrides.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def empty_photons(n, device=None):
    """Zeroed photon arrays: numpy on the host, or torch tensors on `device`."""
    if device is None:
        return {name: np.zeros(n, dtype=dt) for name, dt in PHOTON_FIELDS}
    import torch
    tdt = {np.dtype(np.float64): torch.float64, np.dtype(np.uint64): torch.int64,
           np.dtype(np.uint8): torch.uint8}
    return {name: torch.zeros(n, dtype=tdt[np.dtype(dt)], device=device)
            for name, dt in PHOTON_FIELDS}


def _ptr(a):
    if hasattr(a, "data_ptr"):               # torch tensor
        if not a.is_contiguous():
            raise GrtError("photon tensors must be contiguous")
        return a.data_ptr()
    if not a.flags.c_contiguous:
        raise GrtError("photon arrays must be C-contiguous")
    return a.ctypes.data


def _torch_dtypes(dt):
    """torch dtypes accepted for a photon field of numpy dtype dt (same width;
    ids may be signed, as torch has no full unsigned 64-bit arithmetic)."""
    import torch
    if dt == np.uint64:
        return tuple(t for t in (torch.int64, getattr(torch, "uint64", None)) if t is not None)
    return {np.float64: (torch.float64,), np.uint8: (torch.uint8,)}[dt]


def photon_struct(ph, device_index=None):
    """grt_photons over the 20 arrays of `ph`: all numpy arrays (host entry
    points) or all torch tensors on the GPU `device_index` (device entry
    points), one-dimensional, each of its field's dtype and all of the length
    of x0.  Anything else raises GrtError before a pointer reaches the library
    (a short, mistyped or host array would be read out of bounds by a kernel)."""
    s = Photons()
    n = len(ph["x0"])
    if hasattr(ph["x0"], "data_ptr"):
        # torch tensors: one pass of cheap accessors per field (this runs on
        # every trace call); get_device() is -1 for a CPU tensor
        shape = (n,)
        for name, dt, accepted in _torch_fields():
            a = ph[name]
            if not hasattr(a, "data_ptr"):
                raise GrtError("photon fields must be all numpy arrays or all torch tensors "
                               f"(field {name} is not)")
            dev = a.get_device()
            if dev < 0:
                raise GrtError(f"photon field {name} is a CPU tensor; the device entry "
                               "points need GPU tensors (or pass numpy arrays)")
            if device_index is not None and dev != device_index:
                raise GrtError(f"photon field {name} is on cuda:{dev}, the "
                               f"context on cuda:{device_index}")
            if a.dtype not in accepted:
                raise GrtError(f"photon field {name} must be {np.dtype(dt).name}, not {a.dtype}")
            if a.shape != shape:
                raise GrtError(f"photon field {name} must be one-dimensional of length {n}")
            if not a.is_contiguous():
                raise GrtError("photon tensors must be contiguous")
            setattr(s, name, a.data_ptr())
        return s
    for name, dt in PHOTON_FIELDS:
        a = ph[name]
        if hasattr(a, "data_ptr"):
            raise GrtError("photon fields must be all numpy arrays or all torch tensors "
                           f"(field {name} is not)")
        if not isinstance(a, np.ndarray) or a.dtype != dt:
            raise GrtError(f"photon field {name} must be a numpy {np.dtype(dt).name} array")
        if (a.shape[0] if a.ndim == 1 else -1) != n:
            raise GrtError(f"photon field {name} must be one-dimensional of length {n}")
        setattr(s, name, _ptr(a))
    return s


_TORCH_FIELDS = None


def _torch_fields():
    """(name, numpy dtype, accepted torch dtypes) of every photon field, built
    once: photon_struct runs on every trace call."""
    global _TORCH_FIELDS
    if _TORCH_FIELDS is None:
        _TORCH_FIELDS = [(name, dt, _torch_dtypes(dt)) for name, dt in PHOTON_FIELDS]
    return _TORCH_FIELDS


def _device_vector(t, what, n=None):
    """Pointer of a float64 GPU tensor (contiguous; n elements when given)."""
    import torch
    if not hasattr(t, "data_ptr") or not t.is_cuda:
        raise GrtError(f"{what} must be a GPU tensor")
    if t.dtype != torch.float64:
        raise GrtError(f"{what} must be float64")
    if n is not None and t.numel() != n:
        raise GrtError(f"{what} must have {n} elements")
    return _ptr(t)


def _stats_dict(st):
    return {name: getattr(st, name) for name, _ in Stats._fields_}


class Context:
    """A tracing context on the current HIP device (grt_create)."""

    def __init__(self, params, lib=None):
        """lib: a handle from load(path) to run another build of the library."""
        self._L = lib if lib is not None else load()
        h = C.c_void_p()
        rc = self._L.grt_create(C.byref(params), C.byref(h))
        if rc != 0:
            raise GrtError(self._L.grt_last_error(None).decode())
        self._h = h
        self.params = params
        self._keep = None
        self._grid_shape = None
        # the HIP device grt_create took (the current one)
        self._device_index = _current_device_index()
        # generation of the last taped trace: the adjoint replays the tape a
        # context holds, which a later taped trace replaces
        self._tape_gen = 0

    def _check(self, rc):
        if rc != 0:
            raise GrtError(self._L.grt_last_error(self._h).decode())

    def close(self):
        if self._h:
            self._L.grt_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params):
        self._check(self._L.grt_set_params(self._h, C.byref(params)))
        self.params = params

    def set_grid(self, fields):
        """fields: dict with r, theta, phi, rho, Tgas (nr, nth, nph) and vel, mag (.., 4)."""
        arr = {k: np.ascontiguousarray(fields[k], dtype=np.float64)
               for k in ("r", "theta", "phi", "rho", "Tgas", "vel", "mag")}
        nr, nth, nph = arr["r"].shape
        for k in ("theta", "phi", "rho", "Tgas"):
            if arr[k].shape != (nr, nth, nph):
                raise GrtError(f"Grid field has incompatible dimensions: {k}")
        for k in ("vel", "mag"):
            if arr[k].shape != (nr, nth, nph, 4):
                raise GrtError(f"Grid field has incompatible dimensions: {k}")
        self._check(self._L.grt_set_grid(
            self._h, nr, nth, nph, *[arr[k].ctypes.data for k in
                                     ("r", "theta", "phi", "rho", "Tgas", "vel", "mag")]))
        self._grid_shape = (nr, nth, nph)

    def trace(self, photons, max_steps, termination_percent=1.0, output_interval=0,
              stream=None):
        """Advance photons in place.  numpy dict -> host entry point; torch
        tensors on the GPU -> device entry point on `stream` (torch stream or
        raw handle, default: torch's current stream)."""
        n = len(photons["x0"])
        st = Stats()
        s = photon_struct(photons, self._device_index)
        if hasattr(photons["x0"], "data_ptr"):
            handle = _stream_handle(stream)
            self._check(self._L.grt_trace_device(self._h, C.byref(s), n, max_steps,
                                                 termination_percent, output_interval,
                                                 handle, C.byref(st)))
        else:
            self._check(self._L.grt_trace(self._h, C.byref(s), n, max_steps,
                                          termination_percent, output_interval, C.byref(st)))
        return _stats_dict(st)

    def integrate(self, photons, run, stream=None):
        """integrate_geodesics with its output (grt_integrate_device): trace
        the GPU photons with the run's step limit, termination fraction and
        output interval, writing <output_dir>/<name>_step_<s>_rank0.npz on
        every recount step and, in image mode, <name>_image_products.npz.
        `run` is a grt_amd.io.RunConfig or the dict grt_amd.io.load_config
        returns.  Returns (stats, timers) as dicts."""
        from .io import run_config_struct
        n = len(photons["x0"])
        if not hasattr(photons["x0"], "data_ptr"):
            raise GrtError("integrate takes GPU tensors")
        s = photon_struct(photons, self._device_index)
        rc = run_config_struct(run)
        st = Stats()
        tm = Timers()
        self._check(self._L.grt_integrate_device(self._h, C.byref(s), n, C.addressof(rc),
                                                 _stream_handle(stream), C.byref(st),
                                                 C.byref(tm)))
        return _stats_dict(st), {name: getattr(tm, name) for name, _ in Timers._fields_}

    def trace_noalloc(self, photons, max_steps, termination_percent=1.0, output_interval=0,
                      stream=None):
        """trace() of GPU tensors through grt_trace_device_noalloc: no device
        allocation, at most NOALLOC_MAX_PHOTONS photons, and calls on one GPU
        share one fixed scratch array (serialise them)."""
        n = len(photons["x0"])
        st = Stats()
        s = photon_struct(photons, self._device_index)
        if not hasattr(photons["x0"], "data_ptr"):
            raise GrtError("trace_noalloc takes GPU tensors")
        self._check(self._L.grt_trace_device_noalloc(self._h, C.byref(s), n, max_steps,
                                                     termination_percent, output_interval,
                                                     _stream_handle(stream), C.byref(st)))
        return _stats_dict(st)

    def trace_taped(self, photons, max_steps, slots_per_photon=64, termination_percent=1.0,
                    output_interval=0, stream=None):
        """trace() with the FP64 kernel, keeping the geodesic checkpoints the
        adjoint replays (photons: dict of GPU tensors)."""
        n = len(photons["x0"])
        st = Stats()
        s = photon_struct(photons, self._device_index)
        self._check(self._L.grt_trace_taped_device(self._h, C.byref(s), n, max_steps,
                                                   termination_percent, output_interval,
                                                   slots_per_photon, _stream_handle(stream),
                                                   C.byref(st)))
        self._taped_n = n
        self._tape_gen += 1
        return _stats_dict(st)

    @property
    def tape_generation(self):
        """Counts the taped traces of this context; the adjoint replays the
        one with the current generation."""
        return self._tape_gen

    def adjoint(self, dL_dI, stream=None):
        """Reverse sweep of the last taped trace.  dL_dI: float64 GPU tensor,
        one entry per photon.  Returns (dL/drho, dL/dTgas) shaped like the
        grid and dL/d[absorption_coefficient, emission_coefficient]."""
        import torch
        if not hasattr(dL_dI, "data_ptr") or not dL_dI.is_cuda:
            raise GrtError("dL_dI must be a GPU tensor")
        dL_dI = dL_dI.to(torch.float64).contiguous()
        dev = dL_dI.device
        gr = torch.empty(self._grid_shape, dtype=torch.float64, device=dev)
        gt = torch.empty(self._grid_shape, dtype=torch.float64, device=dev)
        gc = torch.empty(2, dtype=torch.float64, device=dev)
        self._check(self._L.grt_adjoint_device(self._h, dL_dI.data_ptr(), dL_dI.numel(),
                                               gr.data_ptr(), gt.data_ptr(), gc.data_ptr(),
                                               _stream_handle(stream)))
        return gr, gt, gc

    def set_fluid_scalars(self, rho, tgas, stream=None):
        """Replace every cell's density and temperature (numpy arrays or GPU
        tensors shaped like the grid), keeping the rest of the grid."""
        if hasattr(rho, "data_ptr"):
            import torch
            r = rho.to(torch.float64).contiguous()
            t = tgas.to(torch.float64).contiguous()
            if tuple(r.shape) != self._grid_shape or tuple(t.shape) != self._grid_shape:
                raise GrtError("rho and Tgas must have the grid's shape")
            self._check(self._L.grt_set_fluid_scalars_device(
                self._h, _device_vector(r, "rho"), _device_vector(t, "Tgas"),
                _stream_handle(stream)))
        else:
            r = np.ascontiguousarray(rho, dtype=np.float64)
            t = np.ascontiguousarray(tgas, dtype=np.float64)
            if r.shape != self._grid_shape or t.shape != self._grid_shape:
                raise GrtError("rho and Tgas must have the grid's shape")
            self._check(self._L.grt_set_fluid_scalars(self._h, r.ctypes.data, t.ctypes.data))

    def init_image(self, photons, initial_step, screen_u=None, screen_v=None, id_offset=0,
                   stream=None):
        """Fill `photons` (numpy dict or GPU tensor dict) with image-mode rays."""
        n = len(photons["x0"])
        s = photon_struct(photons, self._device_index)
        if hasattr(photons["x0"], "data_ptr"):
            su = _device_vector(screen_u, "screen_u", n) if screen_u is not None else None
            sv = _device_vector(screen_v, "screen_v", n) if screen_v is not None else None
            self._check(self._L.grt_init_image_device(self._h, C.byref(s), n, id_offset,
                                                      initial_step, su, sv,
                                                      _stream_handle(stream)))
        else:
            su = sv = None
            if screen_u is not None:
                screen_u = np.ascontiguousarray(screen_u, dtype=np.float64)
                screen_v = np.ascontiguousarray(screen_v, dtype=np.float64)
                if screen_u.shape != (n,) or screen_v.shape != (n,):
                    raise GrtError(f"screen_u and screen_v must have {n} elements")
                su, sv = screen_u.ctypes.data, screen_v.ctypes.data
            self._check(self._L.grt_init_image(self._h, C.byref(s), n, id_offset, initial_step,
                                               su, sv))
        return photons

    def disk_photon_count(self):
        t = C.c_uint64(0)
        self._check(self._L.grt_disk_photon_count(self._h, C.byref(t)))
        return t.value

    def init_disk(self, photons, initial_step, id_offset=0, stream=None):
        """Fill `photons` with the first len(photons) disk-mode packets."""
        n = len(photons["x0"])
        s = photon_struct(photons, self._device_index)
        if hasattr(photons["x0"], "data_ptr"):
            self._check(self._L.grt_init_disk_device(self._h, C.byref(s), n, id_offset,
                                                     initial_step, _stream_handle(stream)))
        else:
            self._check(self._L.grt_init_disk(self._h, C.byref(s), n, id_offset, initial_step))
        return photons

    def image_products(self, photons, nx, ny, spectrum_bins=128, spectrum_min_hz=1e8,
                       spectrum_max_hz=1e22):
        n = len(photons["x0"])
        s = photon_struct(photons, self._device_index)
        if hasattr(photons["x0"], "data_ptr"):
            import torch
            dev = photons["x0"].device
            img = torch.empty(4 * nx * ny, dtype=torch.float64, device=dev)
            nu = torch.empty(spectrum_bins, dtype=torch.float64, device=dev)
            sp = torch.empty(spectrum_bins, dtype=torch.float64, device=dev)
            arr = torch.zeros(1, dtype=torch.int64, device=dev)
            self._check(self._L.grt_image_products_device(
                self._h, C.byref(s), n, nx, ny, spectrum_bins, spectrum_min_hz,
                spectrum_max_hz, img.data_ptr(), nu.data_ptr(), sp.data_ptr(), arr.data_ptr(),
                _stream_handle(None)))
            return dict(image=img.view(4, ny, nx), spectrum_frequency_hz=nu, spectrum_I=sp,
                        arrived_count=int(arr.item()))
        img = np.zeros(4 * nx * ny)
        nu = np.zeros(spectrum_bins)
        sp = np.zeros(spectrum_bins)
        arr = np.zeros(1, dtype=np.uint64)
        self._check(self._L.grt_image_products(
            self._h, C.byref(s), n, nx, ny, spectrum_bins, spectrum_min_hz, spectrum_max_hz,
            img.ctypes.data, nu.ctypes.data, sp.ctypes.data, arr.ctypes.data))
        return dict(image=img.reshape(4, ny, nx), spectrum_frequency_hz=nu, spectrum_I=sp,
                    arrived_count=int(arr[0]))

    @property
    def tuning(self):
        t = Tuning()
        self._check(self._L.grt_get_tuning(self._h, C.byref(t)))
        return dict(block=t.block, blocks_per_cu=t.blocks_per_cu, sub_budget=t.sub_budget)

    @tuning.setter
    def tuning(self, d):
        t = Tuning()
        self._check(self._L.grt_get_tuning(self._h, C.byref(t)))
        for k, v in d.items():
            setattr(t, k, v)
        self._check(self._L.grt_set_tuning(self._h, C.byref(t)))


def _current_device_index():
    """torch's current GPU index (the device grt_create runs on), or None when
    torch or a GPU is absent."""
    try:
        import torch
        return torch.cuda.current_device() if torch.cuda.is_available() else None
    except ImportError:
        return None


def _stream_handle(stream):
    if stream is None:
        import torch
        return C.c_void_p(torch.cuda.current_stream().cuda_stream)
    if hasattr(stream, "cuda_stream"):
        return C.c_void_p(stream.cuda_stream)
    return C.c_void_p(stream)


def to_device(photons, device="cuda:0"):
    """Copy a numpy photon dict to torch tensors on the GPU."""
    import torch
    out = {}
    for name, dt in PHOTON_FIELDS:
        a = np.ascontiguousarray(photons[name])
        if dt == np.uint64:
            a = a.view(np.int64)
        out[name] = torch.from_numpy(a.copy()).to(device)
    return out


def to_host(photons):
    """Copy torch photon tensors back to a numpy dict."""
    out = {}
    for name, dt in PHOTON_FIELDS:
        a = photons[name].detach().cpu().numpy()
        out[name] = a.view(dt) if dt == np.uint64 else a.astype(dt, copy=False)
    return out
