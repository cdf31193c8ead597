#!/bin/sh
# Builds the CPU harness of the kernel logic (test infrastructure).
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
REPO=$(cd "$HERE/../.." && pwd)
hipcc -x hip --offload-host-only -O2 -std=c++17 -fPIC -shared \
  -I"$REPO/include" -I"$REPO/gr-ray-tracer_amd/csrc" \
  "$HERE/grt_host_harness.cpp" -o "$HERE/libgrt_host_harness.so"
# GPU check of the tracing kernel's correctly rounded sqrt / reciprocal
# (same flags as the library's FP64 kernels).
hipcc --offload-arch=gfx950 -O3 -std=c++17 -fPIC -shared -ffp-contract=off \
  -Xclang -target-feature -Xclang -packed-fp32-ops \
  -I"$REPO/include" -I"$REPO/gr-ray-tracer_amd/csrc" \
  "$HERE/grt_math_check.hip" -o "$HERE/libgrt_math_check.so"
