// grt_kernels.hip -- CDNA4 (gfx950) kernels of the traced path.
//
// trace_fp64: the persistent tracing loop (grt_trace.hpp) with the FP64
// arithmetic of the reference, rounded operation for operation as the
// oracle rounds it (this file is compiled without automatic FMA contraction;
// the only fused operations are the explicit fma calls of grt_device.hpp).
#include <hip/hip_runtime.h>

#include "grt_internal.hpp"
#include "grt_trace.hpp"

namespace grt {

namespace {

template <int kIntegrator, int kSep>
__global__ __launch_bounds__(256, GRT_TRACE_WAVES_PER_EU) void trace_fp64(TraceLaunch a)
{
    trace_body<kIntegrator, kSep, RhsExact, 0>(a);
}


// Live list of the photons not yet terminated.  Each thread tests
// kLiveItems consecutive entries; the block scans its counts and takes its
// slice of the list with one atomic, so entries stay ascending within a block.
// The same pass starts every photon's draw counter and deferred-frequency
// flag of the call at zero (one launch instead of two fills before it).
constexpr int kLiveItems = 8;

__global__ __launch_bounds__(256) void build_live(const uint8_t* __restrict__ term, uint64_t n,
                                                  uint32_t* __restrict__ list,
                                                  uint32_t* __restrict__ count,
                                                  uint32_t* __restrict__ draw,
                                                  uint8_t* __restrict__ fpend)
{
    __shared__ uint32_t wave_total[4];
    __shared__ uint32_t block_base;
    const uint64_t first = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * kLiveItems;
    // the block's slice of the counters and flags, cleared with consecutive
    // lanes on consecutive entries
    const uint64_t slice = (uint64_t)blockIdx.x * blockDim.x * kLiveItems;
#pragma unroll
    for (int q = 0; q < kLiveItems; ++q) {
        const uint64_t i = slice + (uint64_t)q * blockDim.x + threadIdx.x;
        if (i < n) {
            draw[i] = 0u;
            fpend[i] = 0;
        }
    }
    uint32_t live = 0;
    if (first + kLiveItems <= n && (reinterpret_cast<uintptr_t>(term) & 7u) == 0) {
        // the thread's eight flags in one load
        const uint64_t w = *reinterpret_cast<const uint64_t*>(term + first);
#pragma unroll
        for (int q = 0; q < kLiveItems; ++q)
            if (((w >> (8 * q)) & 0xFFu) == 0) live |= 1u << q;
    } else {
#pragma unroll
        for (int q = 0; q < kLiveItems; ++q)
            if (first + q < n && !term[first + q]) live |= 1u << q;
    }
    const uint32_t c = (uint32_t)__popc(live);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t incl = c;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(incl, off, 64);
        if (lane >= off) incl += t;
    }
    if (lane == 63) wave_total[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) tot += wave_total[w];
        block_base = tot ? atomicAdd(count, tot) : 0u;
    }
    __syncthreads();
    uint32_t o = block_base + incl - c;
    for (int w = 0; w < wave; ++w) o += wave_total[w];
#pragma unroll
    for (int q = 0; q < kLiveItems; ++q)
        if (live & (1u << q)) list[o++] = (uint32_t)(first + q);
}

// observation_bin, observation.hpp:22-29
__device__ __forceinline__ int64_t obs_bin(double v, double lo, double hi, int64_t bins)
{
    if (bins <= 1 || !(hi > lo)) return 0;
    const double fr = (v - lo) / (hi - lo);
    const double sc = fmax(0.0, fmin(1.0, fr));
    return (int64_t)(sc * (double)(bins - 1));
}

// Spectra of up to kLdsBins bins are summed per block in LDS first: every
// photon of a one-frequency run falls in the same bin, and direct atomics on
// that one address would serialise the whole launch.
constexpr int kLdsBins = 4096;

// build_observation_products, observation.hpp:31-84 (sums by atomics).
// Draws the thermal frequencies whose draw a tracing launch deferred: the
// five draws before the photon's saved draw counter, at T = -freq, then, for
// a photon that reached the camera, the arrival's redshift factor from its
// stored position and p_t (deferred_arrival_factor; an invalid one zeroes
// the Stokes vector).  A photon still in flight continues with the drawn
// frequency (fpend cleared).
__global__ __launch_bounds__(256) void resolve_deferred(PhotonPtrs ph, uint64_t n, uint64_t seed,
                                                        uint32_t stream, double a, double m)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !ph.fpend[i]) return;
    Photon p;
    p.freq = ph.freq[i];
    p.phase = ph.phase[i];
    p.pend = 1;
    const PhotonRng rng{ph.id[i], seed, stream, ph.rng_draw[i]};
    double fr = 1.0;
    if (p.phase == kArrived) {
        p.s[1] = ph.x[1][i];
        p.s[2] = ph.x[2][i];
        p.s[3] = ph.x[3][i];
        p.s[4] = ph.k[0][i];
        p.eem = ph.eem[i];
        if (!deferred_arrival_factor(a, m, p, fr))
            for (int q = 0; q < 4; ++q) ph.st[q][i] = 0.0;
    }
    finish_deferred(p, rng, fr);
    ph.freq[i] = p.freq;
    ph.fpend[i] = 0;
}

__global__ __launch_bounds__(256) void image_products(PhotonPtrs ph, uint64_t n, int64_t nx,
                                                      int64_t ny, int64_t bins, double smin,
                                                      double smax, double lmin, double lmax,
                                                      double w, double h, double* image,
                                                      double* spec, unsigned long long* arrived)
{
    __shared__ double hist[kLdsBins];
    const bool lds = bins <= kLdsBins;
    if (lds) {
        for (int b = threadIdx.x; b < bins; b += blockDim.x) hist[b] = 0.0;
        __syncthreads();
    }
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool hit = false;
    if (i < n && ph.phase[i] == kArrived) {
        hit = true;
        const int64_t px = obs_bin(ph.udisp[i], -0.5 * w, 0.5 * w, nx);
        const int64_t py = obs_bin(ph.vdisp[i], -0.5 * h, 0.5 * h, ny);
        const int64_t pix = py * nx + px;
        const int64_t np = nx * ny;
        const double I = ph.st[0][i];
        unsafeAtomicAdd(&image[pix], I);
        unsafeAtomicAdd(&image[np + pix], ph.st[1][i]);
        unsafeAtomicAdd(&image[2 * np + pix], ph.st[2][i]);
        unsafeAtomicAdd(&image[3 * np + pix], ph.st[3][i]);
        const double nu = ph.freq[i];
        if (nu >= smin && nu <= smax) {
            const int64_t b = obs_bin(log10(nu), lmin, lmax, bins);
            if (lds) atomicAdd(&hist[b], I);
            else unsafeAtomicAdd(&spec[b], I);
        }
    }
    const uint64_t mask = __ballot(hit);
    if (mask && (threadIdx.x & 63) == (uint32_t)(__ffsll((long long)mask) - 1))
        atomicAdd(arrived, (unsigned long long)__popcll(mask));
    if (lds) {
        __syncthreads();
        for (int b = threadIdx.x; b < bins; b += blockDim.x)
            if (hist[b] != 0.0) unsafeAtomicAdd(&spec[b], hist[b]);
    }
}

// initialize_photons_image, initialize_photons.hpp:82-170.  Screen
// positions come from su/sv, or (when NULL) from two uniform draws of the
// photon's initialisation stream, drand(-w/2, w/2) and drand(-h/2, h/2).
__global__ __launch_bounds__(256) void init_image(PhotonPtrs ph, uint64_t n, uint64_t id0,
                                                  CameraInit C, double a, double m, double w,
                                                  double h, uint64_t seed, uint32_t stream,
                                                  const double* su, const double* sv)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t id = id0 + i;
    double u, v;
    if (su) {
        u = su[i];
        v = sv[i];
    } else {
        PhotonRng rng{id, seed, stream, 0u};
        u = -w / 2 + (w / 2 - (-w / 2)) * rng.drand();
        v = -h / 2 + (h / 2 - (-h / 2)) * rng.drand();
    }
    double s[8];
    init_image_ray(C, u, v, a, m, s);
    ph.id[i] = id;
    ph.freq[i] = 0.0;
    ph.eem[i] = 1.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        ph.x[q][i] = s[q];
        ph.k[q][i] = s[4 + q];
        ph.st[q][i] = 0.0;
    }
    ph.dl[i] = C.dlambda0;
    ph.term[i] = 0;
    ph.phase[i] = kBackward;
    ph.udisp[i] = u;
    ph.vdisp[i] = v;
}

}  // namespace

hipError_t launch_init_image(const PhotonPtrs& ph, uint64_t n, uint64_t id0, const CameraInit& C,
                             double a, double m, double w, double h, uint64_t seed,
                             uint32_t stream, const double* su, const double* sv,
                             hipStream_t hs)
{
    if (n == 0) return hipSuccess;
    const int block = 256;
    const uint64_t blocks = (n + block - 1) / block;
    hipLaunchKernelGGL(init_image, dim3((unsigned)blocks), dim3(block), 0, hs, ph, n, id0, C, a, m,
                       w, h, seed, stream, su, sv);
    return hipGetLastError();
}

hipError_t launch_trace(const TraceLaunch& a, int blocks, int block, size_t lds_bytes,
                        hipStream_t stream)
{
    if (a.tape) return launch_trace_taped(a, blocks, block, lds_bytes, stream);
    if (a.mixed) return launch_trace_mixed(a, blocks, block, lds_bytes, stream);
    const bool sep = a.grid.separable || !a.grid.available;
    if (a.integrator == GRT_RK4) {
        if (sep)
            hipLaunchKernelGGL((trace_fp64<GRT_RK4, 1>), dim3(blocks), dim3(block), lds_bytes, stream, a);
        else
            hipLaunchKernelGGL((trace_fp64<GRT_RK4, 0>), dim3(blocks), dim3(block), lds_bytes, stream, a);
    } else {
        if (sep)
            hipLaunchKernelGGL((trace_fp64<GRT_RK45, 1>), dim3(blocks), dim3(block), lds_bytes, stream, a);
        else
            hipLaunchKernelGGL((trace_fp64<GRT_RK45, 0>), dim3(blocks), dim3(block), lds_bytes, stream, a);
    }
    return hipGetLastError();
}

hipError_t launch_build_live(const uint8_t* term, uint64_t n, uint32_t* list, uint32_t* count,
                             uint32_t* draw, uint8_t* fpend, hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    const int block = 256;
    const uint64_t per_block = (uint64_t)block * kLiveItems;
    const uint64_t blocks = (n + per_block - 1) / per_block;
    hipLaunchKernelGGL(build_live, dim3((unsigned)blocks), dim3(block), 0, stream, term, n, list,
                       count, draw, fpend);
    return hipGetLastError();
}

hipError_t launch_resolve_deferred(const PhotonPtrs& ph, uint64_t n, uint64_t seed,
                                   uint32_t stream_id, double a, double m, hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    const int block = 256;
    const uint64_t blocks = (n + block - 1) / block;
    hipLaunchKernelGGL(resolve_deferred, dim3((unsigned)blocks), dim3(block), 0, stream, ph, n,
                       seed, stream_id, a, m);
    return hipGetLastError();
}

hipError_t launch_image_products(const PhotonPtrs& ph, uint64_t n, int64_t nx, int64_t ny,
                                 int64_t bins, double smin, double smax, double lmin, double lmax,
                                 double width, double height, double* image, double* spec,
                                 unsigned long long* arrived, hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    const int block = 256;
    const uint64_t blocks = (n + block - 1) / block;
    hipLaunchKernelGGL(image_products, dim3((unsigned)blocks), dim3(block), 0, stream, ph, n, nx,
                       ny, bins, smin, smax, lmin, lmax, width, height, image, spec, arrived);
    return hipGetLastError();
}

hipError_t trace_kernel_attributes(int* vgprs, int* sgprs, int* lds_static, int* max_threads)
{
    hipFuncAttributes attr;
    const hipError_t e = hipFuncGetAttributes(&attr, reinterpret_cast<const void*>(
                                                         &trace_fp64<GRT_RK45, 1>));
    if (e != hipSuccess) return e;
    *vgprs = 0;
    *sgprs = 0;
    *lds_static = (int)attr.sharedSizeBytes;
    *max_threads = attr.maxThreadsPerBlock;
    return hipSuccess;
}

}  // namespace grt
