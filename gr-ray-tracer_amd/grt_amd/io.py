"""The file formats either side of the traced path, and a run driver.

load_config(path)            the reference's TOML (toml_config.hpp:129-277):
                             (Params, run settings dict)
read_npy(path)               an f8 C-order .npy array (cnpy.h:302-349)
load_numpy_fields(ctx, dir)  the seven-file .npy bundle into the context
                             (load_python_arrays.hpp:15-223)
write_image_products_npz     <name>_image_products.npz (write_output.hpp:107-142)
write_photons_npz            <name>_step_<s>_rank0.npz (write_output.hpp:90-103)
run(config_path)             main.cpp:246-266 on one GPU: config -> grid ->
                             photons (image camera or disk emission) ->
                             integrate_geodesics -> archives

All parsing and writing is done by libgrt_amd.so (grt_io.cpp).
"""
import ctypes as C
import os

import numpy as np

from . import (GrtError, Params, Photons, default_params, load, photon_struct, to_host,
               MODE_IMAGE)


class RunConfig(C.Structure):
    _fields_ = [
        ("name", C.c_char * 256), ("numpy_dir", C.c_char * 1024),
        ("output_dir", C.c_char * 1024), ("output_variables", C.c_char * 512),
        ("decomposition", C.c_char * 128),
        ("vacuum", C.c_int32), ("max_steps", C.c_int32), ("photon_count", C.c_uint64),
        ("termination_percent", C.c_double), ("initial_step", C.c_double),
        ("output_interval", C.c_uint64), ("output_stride", C.c_uint64),
        ("image_nx", C.c_uint64), ("image_ny", C.c_uint64), ("spectrum_bins", C.c_uint64),
        ("spectrum_min_hz", C.c_double), ("spectrum_max_hz", C.c_double),
    ]


def _lib():
    L = load()
    if not hasattr(L, "_grt_io_ready"):
        vp = C.c_void_p
        L.grt_load_config.restype = C.c_int
        L.grt_load_config.argtypes = [C.c_char_p, C.POINTER(Params), C.POINTER(RunConfig)]
        L.grt_read_npy.restype = C.c_int
        L.grt_read_npy.argtypes = [C.c_char_p, vp, C.c_uint64, C.POINTER(C.c_uint64),
                                   C.POINTER(C.c_int32)]
        L.grt_load_numpy_fields.restype = C.c_int
        L.grt_load_numpy_fields.argtypes = [vp, C.c_char_p]
        L.grt_write_image_products_npz.restype = C.c_int
        L.grt_write_image_products_npz.argtypes = [C.c_char_p, C.c_int64, C.c_int64, C.c_int64,
                                                   vp, vp, vp, C.c_uint64]
        L.grt_write_photons_npz.restype = C.c_int
        L.grt_write_photons_npz.argtypes = [C.c_char_p, C.POINTER(Photons), C.c_uint64,
                                            C.c_char_p, C.c_uint64]
        L._grt_io_ready = True
    return L


def _err():
    return GrtError(_lib().grt_last_error(None).decode())


def run_config_struct(run):
    """A RunConfig from a RunConfig or from the dict load_config returns
    (missing keys keep the reference defaults of a fresh parse)."""
    if isinstance(run, RunConfig):
        return run
    r = RunConfig()
    for name, ctype in RunConfig._fields_:
        if name not in run:
            continue
        v = run[name]
        if isinstance(v, str):
            v = v.encode()
        setattr(r, name, v)
    return r


def load_config(path):
    """Parse a reference TOML config: (Params, dict of the run settings)."""
    p = Params()
    r = RunConfig()
    if _lib().grt_load_config(os.fsencode(path), C.byref(p), C.byref(r)) != 0:
        raise _err()
    run = {}
    for name, _ in RunConfig._fields_:
        v = getattr(r, name)
        run[name] = v.decode() if isinstance(v, bytes) else v
    run["vacuum"] = bool(run["vacuum"])
    return p, run


def read_npy(path):
    shape = (C.c_uint64 * 4)()
    nd = C.c_int32(0)
    L = _lib()
    if L.grt_read_npy(os.fsencode(path), None, 0, shape, C.byref(nd)) != 0:
        raise _err()
    dims = tuple(int(shape[i]) for i in range(nd.value))
    out = np.empty(dims, dtype=np.float64)
    if L.grt_read_npy(os.fsencode(path), out.ctypes.data, out.size, shape, C.byref(nd)) != 0:
        raise _err()
    return out


def load_numpy_fields(ctx, directory):
    """Load r, theta, phi, rho, Tgas, vel, mag .npy files of `directory` into
    the context (ctx None: only check the bundle)."""
    L = _lib()
    h = ctx._h if ctx is not None else None
    if L.grt_load_numpy_fields(h, os.fsencode(directory)) != 0:
        raise GrtError(L.grt_last_error(h).decode())
    if ctx is not None:
        r = read_npy(os.path.join(directory, "r.npy"))
        ctx._grid_shape = r.shape


def _host(a):
    return a.detach().cpu().numpy() if hasattr(a, "detach") else np.asarray(a)


def write_image_products_npz(path, products):
    """products: the dict of Context.image_products (numpy or torch)."""
    img = np.ascontiguousarray(_host(products["image"]), dtype=np.float64)
    nu = np.ascontiguousarray(_host(products["spectrum_frequency_hz"]), dtype=np.float64)
    sI = np.ascontiguousarray(_host(products["spectrum_I"]), dtype=np.float64)
    _, ny, nx = img.shape
    if _lib().grt_write_image_products_npz(os.fsencode(path), nx, ny, len(nu), img.ctypes.data,
                                           nu.ctypes.data, sI.ctypes.data,
                                           int(products["arrived_count"])) != 0:
        raise _err()


def write_photons_npz(path, photons, variables=None, stride=1):
    """photons: host (numpy) photon dict."""
    s = photon_struct(photons)
    v = variables.encode() if variables else None
    if _lib().grt_write_photons_npz(os.fsencode(path), C.byref(s), len(photons["x0"]), v,
                                    stride) != 0:
        raise _err()


def print_timers(timers, steps):
    """The reference's timer table (display.hpp:95-120) for `steps` completed
    iterations."""
    rows = (("Geodesic Integration", timers["geodesic_integration_s"]),
            ("Termination Count", timers["termination_count_s"]),
            ("Output", timers["output_s"]))
    total = sum(t for _, t in rows)
    print(f"Iteration: [{steps}]")
    for label, t in rows:
        pct = 100.0 * t / total if total > 0 else 0.0
        print(f"{label}[% total]: {pct}% \t[time]: {t} seconds")
    print(f"Time elapsed: {total} seconds")


def run(config_path, device="cuda:0", verbose=False):
    """One run of the reference program (main.cpp:246-266 with
    integrate_geodesics, integrate_cart_ks.hpp:423-535) on one GPU: config ->
    grid -> photons -> the global loop with its step archives on every
    recount step and, in image mode, the image products.  Returns the trace
    statistics, the per-phase timers and the archive paths written."""
    import torch
    import grt_amd
    params, rc = load_config(config_path)
    if rc["vacuum"]:
        params.scattering_enabled = 0
    ctx = grt_amd.Context(params)
    if not rc["vacuum"]:
        load_numpy_fields(ctx, rc["numpy_dir"])
    if params.mode == MODE_IMAGE:
        n = rc["photon_count"]
        ph = grt_amd.empty_photons(n, device=device)
        ctx.init_image(ph, rc["initial_step"])
    else:
        n = ctx.disk_photon_count()
        ph = grt_amd.empty_photons(n, device=device)
        ctx.init_disk(ph, rc["initial_step"])
    stats, timers = ctx.integrate(ph, rc)
    torch.cuda.synchronize()
    if verbose:
        print_timers(timers, stats["global_steps"])
    name = os.path.join(rc["output_dir"], rc["name"])
    steps = [s for s in range(stats["global_steps"])
             if s % rc["output_interval"] == 0 or s == rc["max_steps"] - 1]
    out = {"stats": stats, "timers": timers, "photons": n, "photons_final": ph,
           "step_archives": [f"{name}_step_{s}_rank0.npz" for s in steps]}
    out["step_archive"] = out["step_archives"][-1] if steps else None
    if params.mode == MODE_IMAGE:
        out["image_products"] = f"{name}_image_products.npz"
        with np.load(out["image_products"], allow_pickle=False) as z:
            out["arrived_count"] = int(z["arrived_count"][0])
    return out
