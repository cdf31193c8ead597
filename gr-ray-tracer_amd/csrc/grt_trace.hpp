// grt_trace.hpp -- the persistent, refilling tracing loop shared by the
// FP64 kernel (grt_kernels.hip) and the mixed-precision kernel
// (grt_mixed.hip).  Device code only.
//
// One photon per lane.  Each lane keeps its photon entirely in registers and
// loops over RK45 attempts; when its photon terminates or uses up this
// launch's step budget the lane writes the photon back and takes the next
// index from the launch's work list with one atomic per wavefront (ballot +
// mbcnt), so short paths never idle a lane while long paths in the same
// wavefront continue.  Photons still live after the budget are appended to
// the next launch's list, which bounds the tail of every launch (compaction
// between launches).
#pragma once

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

// The persistent loop shared by the tracing kernels; Rhs and kMixed select
// the arithmetic of the step (see lane_iterate).
template <int kIntegrator, int kSep, class Rhs, int kMixed>
__device__ __forceinline__ void trace_body(const TraceLaunch& a)
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
        const bool call_done = lane_iterate<kIntegrator, kSep, Rhs, kMixed>(L, G, P, rng, n, au, av);
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

}  // namespace
}  // namespace grt
