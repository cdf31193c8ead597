#!/bin/sh
# Builds the CPU harness of the kernel logic (test infrastructure).
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
REPO=$(cd "$HERE/../.." && pwd)
hipcc -x hip --offload-host-only -O2 -std=c++17 -fPIC -shared \
  -I"$REPO/include" -I"$REPO/gr-ray-tracer_amd/csrc" \
  "$HERE/grt_host_harness.cpp" -o "$HERE/libgrt_host_harness.so"
