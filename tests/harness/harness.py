"""ctypes wrapper of libgrt_host_harness.so (TEST INFRASTRUCTURE): the
tracing kernel's per-lane logic compiled for the CPU."""
import ctypes as C
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "gr-ray-tracer_amd"))
import grt_amd  # noqa: E402  (only for the Params / Photons layouts)

LIB = os.path.join(HERE, "libgrt_host_harness.so")
_lib = None


def _sources():
    csrc = os.path.join(REPO, "gr-ray-tracer_amd", "csrc")
    return [os.path.join(HERE, f) for f in ("grt_host_harness.cpp", "grt_math_check.hip",
                                             "build.sh")] + \
        [os.path.join(csrc, f) for f in os.listdir(csrc) if f.endswith((".hpp", ".h"))] + \
        [os.path.join(REPO, "include", f) for f in os.listdir(os.path.join(REPO, "include"))]


def build(force=False):
    """Builds the harness when it is missing or older than its sources."""
    stale = not os.path.exists(LIB) or \
        os.path.getmtime(LIB) < max(os.path.getmtime(f) for f in _sources())
    if force or stale:
        subprocess.run(["sh", os.path.join(HERE, "build.sh")], check=True)


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(LIB)
        vp = C.c_void_p
        L.harness_trace.restype = C.c_int
        L.harness_trace.argtypes = [C.POINTER(grt_amd.Params), C.c_int, C.c_int, C.c_int] + \
            [vp] * 7 + [C.c_int, C.POINTER(grt_amd.Photons), C.c_uint64, C.c_int,
                        C.POINTER(C.c_uint64)]
        L.harness_locate.restype = C.c_int
        L.harness_locate.argtypes = [C.POINTER(grt_amd.Params), C.c_int, C.c_int, C.c_int] + \
            [vp] * 7 + [C.c_int, C.c_int64, vp, vp, vp, vp]
        L.harness_axis_counts.restype = C.c_int
        L.harness_axis_counts.argtypes = [C.POINTER(grt_amd.Params), C.c_int, C.c_int, C.c_int] + \
            [vp] * 7 + [C.c_int, C.c_int64, vp, vp]
        L.harness_locate_raw.restype = C.c_int
        L.harness_locate_raw.argtypes = [C.POINTER(grt_amd.Params), C.c_int, C.c_int, C.c_int] + \
            [vp] * 7 + [C.c_int, C.c_int64, vp, vp, vp, vp, vp, vp]
        _lib = L
    return _lib


def rhs(states, a, m):
    """(ks_rhs, ks_rhs_fast) of the device code, compiled for the CPU."""
    L = lib()
    st = np.ascontiguousarray(states, dtype=np.float64).reshape(-1, 8)
    ex = np.zeros_like(st)
    fa = np.zeros_like(st)
    L.harness_rhs.restype = None
    L.harness_rhs.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_void_p, C.c_void_p]
    L.harness_rhs(st.ctypes.data, len(st), a, m, ex.ctypes.data, fa.ctypes.data)
    return ex, fa


def _grid_args(fields):
    if fields is None:
        return [0, 0, 0] + [None] * 7, []
    arr = [np.ascontiguousarray(fields[k], dtype=np.float64)
           for k in ("r", "theta", "phi", "rho", "Tgas", "vel", "mag")]
    nr, nth, nph = arr[0].shape
    return [nr, nth, nph] + [a.ctypes.data for a in arr], arr


def trace(params, fields, photons, max_steps, force_general=False):
    args, keep = _grid_args(fields)
    st = (C.c_uint64 * 6)()
    s = grt_amd.photon_struct(photons)
    lib().harness_trace(C.byref(params), *args, int(force_general), C.byref(s),
                        len(photons["x0"]), max_steps, st)
    return dict(functor_calls=st[0], accepted_steps=st[1], rejected_attempts=st[2],
                failed_steps=st[3], fluid_samples=st[4], fluid_hits=st[5])


def locate(params, fields, radius, polar, az, force_general=False):
    args, keep = _grid_args(fields)
    radius = np.ascontiguousarray(radius, dtype=np.float64)
    polar = np.ascontiguousarray(polar, dtype=np.float64)
    az = np.ascontiguousarray(az, dtype=np.float64)
    out = np.zeros((len(radius), 4), dtype=np.int32)
    sep = lib().harness_locate(C.byref(params), *args, int(force_general), len(radius),
                               radius.ctypes.data, polar.ctypes.data, az.ctypes.data,
                               out.ctypes.data)
    return out, bool(sep)


def locate_raw(params, fields, radius, ct, X, Y, no_buckets=False):
    """The separable search on the kernel's own inputs; returns (found, i, j, k)
    rows and the bucket count of each axis."""
    args, keep = _grid_args(fields)
    q = [np.ascontiguousarray(v, dtype=np.float64) for v in (radius, ct, X, Y)]
    out = np.zeros((len(q[0]), 4), dtype=np.int32)
    nb = np.zeros(3, dtype=np.int32)
    lib().harness_locate_raw(C.byref(params), *args, int(no_buckets), len(q[0]),
                             *[v.ctypes.data for v in q], out.ctypes.data, nb.ctypes.data)
    return out, nb


def axis_counts(params, fields, axis, values):
    """(plain, bucketed) counts of the separable search's axis tables for the
    compared values (r, cos theta or pseudo-angle); None without buckets."""
    args, keep = _grid_args(fields)
    v = np.ascontiguousarray(values, dtype=np.float64)
    out = np.zeros((len(v), 2), dtype=np.int32)
    ok = lib().harness_axis_counts(C.byref(params), *args, axis, len(v), v.ctypes.data,
                                   out.ctypes.data)
    return (out[:, 0], out[:, 1]) if ok else None
