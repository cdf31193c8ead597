// grt_kernels.hip -- CDNA4 (gfx950) kernels of the traced path.
//
// trace_fp64: persistent, refilling tracing kernel.  One photon per lane.
// Each lane keeps its photon entirely in registers and loops over RK45
// attempts; when its photon terminates or uses up this launch's step budget
// the lane writes the photon back and takes the next index from the launch's
// work list with one atomic per wavefront (ballot + mbcnt), so short paths
// never idle a lane while long paths in the same wavefront continue.  Photons
// still live after the budget are appended to the next launch's list, which
// bounds the tail of every launch (compaction between launches).
#include <hip/hip_runtime.h>

#include "grt_internal.hpp"

namespace grt {

namespace {

__device__ __forceinline__ uint32_t lane_rank(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Wave-aggregated fetch-add: every active lane gets base + its rank among the
// active lanes; one atomic per wavefront.
__device__ __forceinline__ uint32_t wave_claim(uint32_t* counter)
{
    const uint64_t mask = __ballot(1);
    const uint32_t rank = lane_rank(mask);
    const uint32_t total = (uint32_t)__popcll(mask);
    uint32_t base = 0;
    if (rank == 0) base = atomicAdd(counter, total);
    base = __builtin_amdgcn_readfirstlane(base);
    return base + rank;
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ void load_photon(const PhotonPtrs& p, uint32_t idx, Photon& ph,
                                            PhotonRng& rng)
{
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        ph.s[q] = p.x[q][idx];
        ph.s[4 + q] = p.k[q][idx];
        ph.stokes[q] = p.st[q][idx];
    }
    ph.freq = p.freq[idx];
    ph.eem = p.eem[idx];
    ph.dl = p.dl[idx];
    ph.phase = p.phase[idx];
    ph.term = 0;
    rng.id = p.id[idx];
    rng.draw = p.rng_draw[idx];
}

__device__ __forceinline__ void store_photon(const PhotonPtrs& p, uint32_t idx, const Photon& ph,
                                             const PhotonRng& rng, double au, double av)
{
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        p.x[q][idx] = ph.s[q];
        p.k[q][idx] = ph.s[4 + q];
        p.st[q][idx] = ph.stokes[q];
    }
    p.freq[idx] = ph.freq;
    p.eem[idx] = ph.eem;
    p.dl[idx] = ph.dl;
    p.phase[idx] = ph.phase;
    p.term[idx] = ph.term;
    if (ph.phase == kArrived) {
        p.udisp[idx] = au;
        p.vdisp[idx] = av;
    }
    p.rng_draw[idx] = rng.draw;
}

#ifndef GRT_TRACE_WAVES_PER_EU
#define GRT_TRACE_WAVES_PER_EU 2
#endif

template <int kIntegrator, int kSep>
__global__ __launch_bounds__(256, GRT_TRACE_WAVES_PER_EU) void trace_fp64(TraceLaunch a)
{
    extern __shared__ double lds[];
    GridView G = a.grid;
    if (kSep && a.use_lds) {
        // Stage the separable search tables in LDS.
        for (int t = threadIdx.x; t < a.lds_doubles; t += blockDim.x) lds[t] = G.tab[t];
        __syncthreads();
        G.tab = lds;
    }
    const StepParams& P = a.consts->P;
    const PhotonPtrs& PH = a.consts->ph;

    LaneCounters n{0u, 0u, 0u, 0u, 0u};
    uint32_t last_term = 0;

    Lane L;
    PhotonRng rng;
    rng.seed = P.seed;
    rng.stream = P.stream << 1;
    uint32_t slot = wave_claim(a.queue);
    bool have = slot < a.n_in;
    uint32_t idx = have ? a.list_in[slot] : 0u;
    if (have) load_photon(PH, idx, L.ph, rng);
    lane_reset(L);
    int calls_left = a.budget;
    int calls_done = 0;

    while (have) {
        double au = 0.0, av = 0.0;
        const bool call_done = lane_iterate<kIntegrator, kSep>(L, G, P, rng, n, au, av);
        if (call_done) {
            ++n.calls;
            ++calls_done;
            --calls_left;
            if (L.ph.term || calls_left == 0) {
                store_photon(PH, idx, L.ph, rng, au, av);
                if (L.ph.term) {
                    const uint32_t step = (uint32_t)(a.step0 + calls_done - 1);
                    last_term = step > last_term ? step : last_term;
                } else {
                    const uint32_t o = wave_claim(a.n_out);
                    a.list_out[o] = idx;
                }
                slot = wave_claim(a.queue);
                have = slot < a.n_in;
                if (have) {
                    idx = a.list_in[slot];
                    load_photon(PH, idx, L.ph, rng);
                    lane_reset(L);
                    calls_left = a.budget;
                    calls_done = 0;
                }
            }
        }
    }

    const unsigned long long s_calls = wave_sum(n.calls);
    const unsigned long long s_acc = wave_sum(n.accepted);
    const unsigned long long s_rej = wave_sum(n.rejected);
    const unsigned long long s_fail = wave_sum(n.failed);
    const unsigned long long s_samp = wave_sum(n.samples);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t o = (uint32_t)__shfl_xor((int)last_term, off, 64);
        last_term = o > last_term ? o : last_term;
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&a.stats[kStatCalls], s_calls);
        atomicAdd(&a.stats[kStatAccepted], s_acc);
        atomicAdd(&a.stats[kStatRejected], s_rej);
        atomicAdd(&a.stats[kStatFailed], s_fail);
        atomicAdd(&a.stats[kStatSamples], s_samp);
        atomicMax(&a.stats[kStatLastTermStep], (unsigned long long)last_term);
    }
}

// Live list of the photons not yet terminated.  Each thread tests
// kLiveItems consecutive entries; the block scans its counts and takes its
// slice of the list with one atomic, so entries stay ascending within a block.
constexpr int kLiveItems = 8;

__global__ __launch_bounds__(256) void build_live(const uint8_t* __restrict__ term, uint64_t n,
                                                  uint32_t* __restrict__ list,
                                                  uint32_t* __restrict__ count)
{
    __shared__ uint32_t wave_total[4];
    __shared__ uint32_t block_base;
    const uint64_t first = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * kLiveItems;
    uint32_t live = 0;
#pragma unroll
    for (int q = 0; q < kLiveItems; ++q)
        if (first + q < n && !term[first + q]) live |= 1u << q;
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
                             hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    const int block = 256;
    const uint64_t per_block = (uint64_t)block * kLiveItems;
    const uint64_t blocks = (n + per_block - 1) / per_block;
    hipLaunchKernelGGL(build_live, dim3((unsigned)blocks), dim3(block), 0, stream, term, n, list,
                       count);
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
