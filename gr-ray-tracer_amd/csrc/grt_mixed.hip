// grt_mixed.hip -- the mixed-precision form of the tracing kernel.
//
// The geodesic stays FP64: the photon state, every right-hand-side
// evaluation, the Runge-Kutta stage sums and the RK45 error norm.  The RK45
// step-size proposal is driven by |x5 - x4|, a difference of two nearly
// equal stage sums, so any FP32 rounding inside the stages (6e-8 relative)
// would move the accept/reject decisions and step sizes by ~1e-3 relative at
// rtol = 1e-7.  The image-mode emission point and the camera arrival point
// are RK step endpoints, so such a step sequence changes most pixels.
// FP64 is therefore kept for the geodesic, with relaxed rounding: the
// right-hand side is ks_rhs_fast (the same equations in about 100
// operations instead of 248, with four reciprocals instead of 21 divisions,
// and its reciprocals and roots one Newton step from the hardware's seeds),
// and this file is compiled with FMA contraction, reciprocal-based division
// (-freciprocal-math) and the hardware's approximate sqrt / rsq / rcp
// sequences (-fapprox-func).  These change the result at the 1e-14 level,
// which moves step sizes by ~1e-8 relative.
//
// FP32 is used where nothing feeds back into the step sequence: the fluid
// tetrad (Gram-Schmidt under the Kerr-Schild metric) and the attenuation
// factor exp(-dtau) of the radiative-transfer update.  The controller's
// err^(-1/4) feeds every step size and stays FP64.  A photon scattered with
// the FP32 tetrad is put back on the null cone in FP64.
//
// On this chip FP32 arithmetic issues at the FP64 rate unless it is packed
// two per lane, which this build does not use; FP32 pays only where it
// replaces FP64 division, square roots and transcendentals.
#include <hip/hip_runtime.h>

// The library's sqrt and 1/sqrt, not the exact short sequences (grt_device.hpp).
#define GRT_RELAXED_MATH 1
#include "grt_internal.hpp"
#include "grt_trace.hpp"

namespace grt {

namespace {

template <int kIntegrator, int kSep>
__global__ __launch_bounds__(256, GRT_TRACE_WAVES_PER_EU) void trace_mixed(TraceLaunch a)
{
    trace_body<kIntegrator, kSep, RhsFast, 1>(a);
}

// The control of the mixed form: the same relaxed FP64 geodesic, with the
// tetrad and the attenuation in FP64 as well (GRT_PRECISION_FP64_FAST_RHS).
template <int kIntegrator, int kSep>
__global__ __launch_bounds__(256, GRT_TRACE_WAVES_PER_EU) void trace_fp64_fast_rhs(TraceLaunch a)
{
    trace_body<kIntegrator, kSep, RhsFast, 0>(a);
}

template <int kIntegrator, int kSep>
void launch_one(const TraceLaunch& a, int blocks, int block, size_t lds_bytes, hipStream_t stream)
{
    if (a.mixed == GRT_PRECISION_FP64_FAST_RHS)
        hipLaunchKernelGGL((trace_fp64_fast_rhs<kIntegrator, kSep>), dim3(blocks), dim3(block),
                           lds_bytes, stream, a);
    else
        hipLaunchKernelGGL((trace_mixed<kIntegrator, kSep>), dim3(blocks), dim3(block), lds_bytes,
                           stream, a);
}

}  // namespace

hipError_t launch_trace_mixed(const TraceLaunch& a, int blocks, int block, size_t lds_bytes,
                              hipStream_t stream)
{
    const bool sep = a.grid.separable || !a.grid.available;
    if (a.integrator == GRT_RK4) {
        if (sep) launch_one<GRT_RK4, 1>(a, blocks, block, lds_bytes, stream);
        else launch_one<GRT_RK4, 0>(a, blocks, block, lds_bytes, stream);
    } else {
        if (sep) launch_one<GRT_RK45, 1>(a, blocks, block, lds_bytes, stream);
        else launch_one<GRT_RK45, 0>(a, blocks, block, lds_bytes, stream);
    }
    return hipGetLastError();
}

}  // namespace grt
