// grt_adjoint.hip -- kernels of the differentiable form (FP64).
//
// trace_taped: the FP64 tracing loop that also writes geodesic checkpoints
//   (one per 32-call segment that touches Stokes I, see grt_trace.hpp).
// adjoint_fp64: the reverse sweep.  For each photon with a non-zero seed
//   dL/dI_final it walks the kept segments from the last to the first,
//   replays the segment's calls from its checkpoint with the same FP64 code
//   (so the replay reproduces the forward pass bit for bit), records every
//   radiative-transfer update and every event that sets I outright, and then
//   runs the recurrence I_after = a I_before + b backwards,
//   a = exp(-dtau), b = dI (radiative_transfer.hpp:81-96):
//     dL/dtheta  += dL/dI_after * (-a I_before ddtau/dtheta + d(dI)/dtheta)
//     dL/dI_before = a dL/dI_after
//   for theta = the cell's density and temperature (scatter-added into the
//   grid by atomics, chained to code units) and the two coefficients.  An
//   event that sets I outright ends the photon's sweep.  The partial
//   derivatives come from the ...Gradient companions of the extension points.
//   The emission energy E / int(rho norm) of an image ray is piecewise
//   constant in rho and is held fixed.
//
// Compiled without FMA contraction, like the FP64 tracing kernel.
#include <hip/hip_runtime.h>

#include "grt_internal.hpp"
#include "grt_trace.hpp"

namespace grt {

namespace {

template <int kIntegrator, int kSep>
__global__ __launch_bounds__(256, GRT_TRACE_WAVES_PER_EU) void trace_taped(TraceLaunch a)
{
    trace_body<kIntegrator, kSep, RhsExact, 0, 1>(a);
}

// Commits the segment a photon was in when the trace ended (the loop's stop
// rule can leave photons live).
__global__ __launch_bounds__(256) void tape_finalize(uint64_t* hdr, uint64_t n, int slots,
                                                     uint32_t* overflow)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t h = hdr[i];
    if (!(h & kTapeActive)) return;
    uint32_t committed = (uint32_t)(h >> 32) & 0x7FFFFFFFu;
    if (committed < (uint32_t)slots) ++committed;
    else atomicOr(overflow, 1u);
    hdr[i] = ((uint64_t)committed << 32) | (uint32_t)h;
}

// Replay records of one lane, interleaved with the other lanes' records.
struct RecordHook {
    double* base;
    uint32_t stride;
    int count;

    __device__ __forceinline__ void put(int q, double v)
    {
        base[((size_t)count * kRecDoubles + q) * stride] = v;
    }
    __device__ __forceinline__ double get(int e, int q) const
    {
        return base[((size_t)e * kRecDoubles + q) * stride];
    }
    __device__ void rt(const FluidCellCGS& fl, uint32_t cell, double nu, double taken, double Ib)
    {
        if (count >= kRecEntries) return;
        put(0, (double)cell);
        put(1, Ib);
        put(2, taken);
        put(3, nu);
#if GRT_GRADIENT_READS_GEOMETRY
        put(4, fl.position_cm[0]);
        put(5, fl.position_cm[1]);
        put(6, fl.position_cm[2]);
        put(7, fl.coordinate_time_s);
#pragma unroll
        for (int q = 0; q < 4; ++q) put(8 + q, fl.magnetic_gauss[q]);
#else
        (void)fl;
#endif
        ++count;
    }
    __device__ void reset()
    {
        if (count >= kRecEntries) return;
        put(0, -1.0);
        ++count;
    }
    // The fluid of event e as the forward pass sampled it: density,
    // temperature and four-velocity are the cell's own (scaled) record, which
    // the tape's validity keeps unchanged; position, time and the magnetic
    // field are recorded only where the gradient companions read them, and
    // are NaN otherwise.
    __device__ FluidCellCGS fluid(int e, const double* cells, uint32_t cell) const
    {
        FluidCellCGS fl;
        const double* c = cells + (size_t)cell * kCellDoubles;
        fl.density_g_cm3 = c[0];
        fl.temperature_k = c[1];
#pragma unroll
        for (int q = 0; q < 4; ++q) fl.coordinate_four_velocity_cm_s[q] = c[2 + q];
#if GRT_GRADIENT_READS_GEOMETRY
        fl.position_cm[0] = get(e, 4);
        fl.position_cm[1] = get(e, 5);
        fl.position_cm[2] = get(e, 6);
        fl.coordinate_time_s = get(e, 7);
#pragma unroll
        for (int q = 0; q < 4; ++q) fl.magnetic_gauss[q] = get(e, 8 + q);
#else
        (void)e;
        const double nan = __builtin_nan("");
        fl.position_cm[0] = fl.position_cm[1] = fl.position_cm[2] = nan;
        fl.coordinate_time_s = nan;
#pragma unroll
        for (int q = 0; q < 4; ++q) fl.magnetic_gauss[q] = nan;
#endif
        return fl;
    }
};

__device__ __forceinline__ double wave_sum_d(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <int kIntegrator, int kSep>
__global__ __launch_bounds__(256, GRT_TRACE_WAVES_PER_EU) void adjoint_fp64(AdjointLaunch a)
{
    extern __shared__ double lds[];
    GridView G = a.grid;
    if (kSep) {
        for (int t = threadIdx.x; t < a.lds_doubles; t += blockDim.x) lds[t] = G.tab[t];
        __syncthreads();
        G.tab = lds;
    }
    const StepParams& P = a.consts->P;
    const uint32_t lane_id = blockIdx.x * blockDim.x + threadIdx.x;
    RecordHook h;
    h.base = a.rec + lane_id;
    h.stride = a.rec_lanes;
    h.count = 0;
    double g_abs = 0.0, g_emit = 0.0;

    uint32_t p = wave_claim(a.queue);
    while (p < a.n) {
        double lam = a.seed[p];
        // A photon's consecutive transfer steps mostly stay in one cell: the
        // sweep sums a run of events in the same cell and scatter-adds it
        // once, when the cell changes or the photon ends.
        uint32_t run_cell = 0xFFFFFFFFu;
        double run_rho = 0.0, run_T = 0.0;
        const uint64_t hdr = a.tape_hdr[p];
        const uint32_t calls_total = (uint32_t)hdr;
        for (int k = (int)((hdr >> 32) & 0x7FFFFFFFu) - 1; k >= 0 && lam != 0.0; --k) {
            Lane L;
            PhotonRng rng;
            rng.seed = P.seed;
            rng.stream = P.stream << 1;
            uint32_t call0;
            read_checkpoint(a.tape + ((size_t)p * a.tape_slots + k) * kCkptDoubles, L.ph, rng,
                            call0);
            lane_reset(L, P);
            const uint32_t ncalls = min(kSegmentCalls, calls_total - call0);
            h.count = 0;
            LaneCounters cnt{0u, 0u, 0u, 0u, 0u, 0u};
            uint32_t done = 0;
            while (done < ncalls) {
                double au = 0.0, av = 0.0;
                if (lane_iterate<kIntegrator, kSep, RhsExact, 0, RecordHook>(L, G, P, rng, cnt,
                                                                             au, av, h)) {
                    ++done;
                    if (L.ph.term) break;
                }
            }
            for (int e = h.count - 1; e >= 0; --e) {
                const double cd = h.get(e, 0);
                if (cd < 0.0) {
                    lam = 0.0;
                    break;
                }
                const uint32_t cell = (uint32_t)cd;
                const FluidCellCGS fl = h.fluid(e, G.cells, cell);
                const double Ib = h.get(e, 1), taken = h.get(e, 2), nu = h.get(e, 3);
                double att = 1.0;
                double dt[3] = {0.0, 0.0, 0.0}, di[3] = {0.0, 0.0, 0.0};
                if (P.absorb)
                    att = exp(-ComputeOpticalDepthIncrementGradient(fl, nu, taken, P.kabs, &dt[0],
                                                                    &dt[1], &dt[2]));
                if (P.emit)
                    ComputeEmissivityIncrementGradient(fl, nu, taken, P.kemit, &di[0], &di[1],
                                                       &di[2]);
                const double aI = att * Ib;
                if (cell != run_cell) {
                    if (run_cell != 0xFFFFFFFFu) {
                        unsafeAtomicAdd(&a.grad_rho[run_cell], run_rho);
                        unsafeAtomicAdd(&a.grad_T[run_cell], run_T);
                    }
                    run_cell = cell;
                    run_rho = run_T = 0.0;
                }
                run_rho += lam * (di[0] - aI * dt[0]) * a.density_scale;
                run_T += lam * (di[1] - aI * dt[1]) * a.temperature_scale;
                g_abs += lam * (-aI * dt[2]);
                g_emit += lam * di[2];
                lam *= att;
            }
        }
        if (run_cell != 0xFFFFFFFFu) {
            unsafeAtomicAdd(&a.grad_rho[run_cell], run_rho);
            unsafeAtomicAdd(&a.grad_T[run_cell], run_T);
        }
        p = wave_claim(a.queue);
    }
    g_abs = wave_sum_d(g_abs);
    g_emit = wave_sum_d(g_emit);
    if ((threadIdx.x & 63) == 0) {
        unsafeAtomicAdd(&a.grad_coeff[0], g_abs);
        unsafeAtomicAdd(&a.grad_coeff[1], g_emit);
    }
}

__global__ __launch_bounds__(256) void set_fluid_scalars(double* cells, double* rho_code,
                                                         const double* rho, const double* tgas,
                                                         uint64_t ncell, double ds, double ts)
{
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const double r = rho[c];
    cells[c * kCellDoubles + 0] = r * ds;
    cells[c * kCellDoubles + 1] = tgas[c] * ts;
    rho_code[c] = r;
}

}  // namespace

hipError_t launch_trace_taped(const TraceLaunch& a, int blocks, int block, size_t lds_bytes,
                              hipStream_t stream)
{
    const bool sep = a.grid.separable || !a.grid.available;
    if (a.integrator == GRT_RK4) {
        if (sep)
            hipLaunchKernelGGL((trace_taped<GRT_RK4, 1>), dim3(blocks), dim3(block), lds_bytes, stream, a);
        else
            hipLaunchKernelGGL((trace_taped<GRT_RK4, 0>), dim3(blocks), dim3(block), lds_bytes, stream, a);
    } else {
        if (sep)
            hipLaunchKernelGGL((trace_taped<GRT_RK45, 1>), dim3(blocks), dim3(block), lds_bytes, stream, a);
        else
            hipLaunchKernelGGL((trace_taped<GRT_RK45, 0>), dim3(blocks), dim3(block), lds_bytes, stream, a);
    }
    return hipGetLastError();
}

hipError_t launch_tape_finalize(uint64_t* hdr, uint64_t n, int slots, uint32_t* overflow,
                                hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    const uint64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(tape_finalize, dim3((unsigned)blocks), dim3(256), 0, stream, hdr, n, slots,
                       overflow);
    return hipGetLastError();
}

hipError_t launch_adjoint(const AdjointLaunch& a, int blocks, int block, size_t lds_bytes,
                          hipStream_t stream)
{
    const bool sep = a.grid.separable || !a.grid.available;
    if (a.integrator == GRT_RK4) {
        if (sep)
            hipLaunchKernelGGL((adjoint_fp64<GRT_RK4, 1>), dim3(blocks), dim3(block), lds_bytes, stream, a);
        else
            hipLaunchKernelGGL((adjoint_fp64<GRT_RK4, 0>), dim3(blocks), dim3(block), lds_bytes, stream, a);
    } else {
        if (sep)
            hipLaunchKernelGGL((adjoint_fp64<GRT_RK45, 1>), dim3(blocks), dim3(block), lds_bytes, stream, a);
        else
            hipLaunchKernelGGL((adjoint_fp64<GRT_RK45, 0>), dim3(blocks), dim3(block), lds_bytes, stream, a);
    }
    return hipGetLastError();
}

hipError_t launch_set_fluid_scalars(double* cells, double* rho_code, const double* rho,
                                    const double* tgas, uint64_t ncell, double density_scale,
                                    double temperature_scale, hipStream_t stream)
{
    if (ncell == 0) return hipSuccess;
    const uint64_t blocks = (ncell + 255) / 256;
    hipLaunchKernelGGL(set_fluid_scalars, dim3((unsigned)blocks), dim3(256), 0, stream, cells,
                       rho_code, rho, tgas, ncell, density_scale, temperature_scale);
    return hipGetLastError();
}

}  // namespace grt
