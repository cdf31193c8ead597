"""Compile libgrt_amd.so (the HIP tracing kernels + the C ABI) for gfx950.

The library is built in-tree next to this package so it travels with the
repository to the GPU host.  Packed FP32 instructions are disabled for the
device code (`-packed-fp32-ops`); the host pass warns that it does not know
that feature, which is harmless.
"""
import os
import shutil
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG_DIR)                 # gr-ray-tracer_amd/
REPO = os.path.dirname(ROOT)
CSRC = os.path.join(ROOT, "csrc")
INCLUDE = os.path.join(REPO, "include")
LIB_PATH = os.path.join(ROOT, "libgrt_amd.so")
# The same library for gradient companions that read the sample's position,
# time or magnetic field (GRT_GRADIENT_READS_GEOMETRY=1: the adjoint records
# those fields too); built beside the default one so its record path is tested.
GEOMETRY_LIB_PATH = os.path.join(ROOT, "libgrt_amd_geometry.so")
GEOMETRY_SOURCES = ("grt_adjoint.hip", "grt_api.cpp")   # the sources the flag changes

SOURCES = ["grt_kernels.hip", "grt_mixed.hip", "grt_adjoint.hip", "grt_disk.hip",
           "grt_noalloc.hip", "grt_api.cpp", "grt_io.cpp"]
HEADERS = ["grt_device.hpp", "grt_internal.hpp", "grt_trace.hpp"]
# The mixed-precision kernel does not reproduce the reference's rounding
# operation for operation: it may contract a * b + c into one FMA, divide by
# multiplying with a reciprocal and use the hardware's approximate
# sqrt / exp / reciprocal instructions.
RELAXED_FLAGS = {"grt_mixed.hip": ["-freciprocal-math", "-fapprox-func"]}


def hipcc():
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libgrt_amd.so")


def compile_flags(resource_usage=False, contract="off"):
    # -ffp-contract=off: the FP64 path fuses only where the source calls fma
    # explicitly and otherwise rounds every +, -, * as the reference's
    # expressions do, see grt_device.hpp.
    flags = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=" + contract,
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function",
             "-I" + INCLUDE, "-I" + CSRC]
    if resource_usage:
        flags.append("-Rpass-analysis=kernel-resource-usage")
    return flags


def build(verbose=False, resource_usage=False, extra_flags=None, out=None):
    """Compile every source into libgrt_amd.so (always recompiles).
    extra_flags: {source: [flags]} appended for that source; out: another
    output path for the library."""
    extra_flags = extra_flags or {}
    lib_path = out or LIB_PATH
    cc = hipcc()
    objs = []
    build_dir = os.path.join(ROOT, "_build") if out is None else out + ".objs"
    os.makedirs(build_dir, exist_ok=True)
    for src in SOURCES:
        obj = os.path.join(build_dir, src + ".o")
        contract = "fast" if src in RELAXED_FLAGS else "off"
        cmd = [cc] + compile_flags(resource_usage and src.endswith(".hip"), contract) + \
            RELAXED_FLAGS.get(src, []) + extra_flags.get(src, []) + \
            ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"hipcc failed on {src}")
        if verbose or resource_usage:
            sys.stderr.write(r.stderr)
        objs.append(obj)
    tmp = lib_path + ".tmp"
    cmd = [cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc link of libgrt_amd.so failed")
    os.replace(tmp, lib_path)
    if out is None and not extra_flags:
        # the geometry-recording variant: the sources the flag changes are
        # compiled again, the rest shared
        gobjs = []
        for src, obj in zip(SOURCES, objs):
            if src in GEOMETRY_SOURCES:
                gobj = os.path.join(build_dir, src + ".geometry.o")
                contract = "fast" if src in RELAXED_FLAGS else "off"
                cmd = [cc] + compile_flags(False, contract) + RELAXED_FLAGS.get(src, []) + \
                    ["-DGRT_GRADIENT_READS_GEOMETRY=1", "-c", os.path.join(CSRC, src), "-o", gobj]
                r = subprocess.run(cmd, capture_output=True, text=True)
                if r.returncode != 0:
                    sys.stderr.write(r.stdout + r.stderr)
                    raise RuntimeError(f"hipcc failed on {src} (geometry records)")
                gobjs.append(gobj)
            else:
                gobjs.append(obj)
        gtmp = GEOMETRY_LIB_PATH + ".tmp"
        r = subprocess.run([cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", gtmp] + gobjs,
                           capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("hipcc link of libgrt_amd_geometry.so failed")
        os.replace(gtmp, GEOMETRY_LIB_PATH)
    return lib_path


if __name__ == "__main__":
    build(verbose=True, resource_usage="--usage" in sys.argv)
