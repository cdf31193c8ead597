// grt_trace.hpp -- the persistent, refilling tracing loop shared by the
// FP64 kernel (grt_kernels.hip) and the mixed-precision kernel
// (grt_mixed.hip).  Device code only.
//
// One photon per lane.  Each lane keeps its photon entirely in registers and
// loops over RK45 attempts; when its photon terminates or uses up this
// launch's step budget the lane writes the photon back and takes the next
// photon of the launch's work list from its wavefront's stage in LDS (the
// next slots of the list, claimed with one atomic and loaded with one
// coalesced pass), so short paths never idle a lane while long paths in the
// same wavefront continue, and a refill costs an LDS read instead of an HBM
// round trip.  Photons still live after the budget are appended to the next
// launch's list, which bounds the tail of every launch (compaction between
// launches).
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

// The photon arrays live in global memory; saying so keeps their accesses
// global (vector memory only) whatever the pointer was loaded from.
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gmem(T* p)
{
    return (__attribute__((address_space(1))) T*)p;
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Fills stage positions [0, count) with work-list slots base + position,
// lane l loading position l; slots at or past n_in stay empty.  Called with
// all 64 lanes of the wavefront active, so each photon array is read in one
// coalesced pass, and the photons' Kerr-Schild radii (spin a) are computed
// 64 at a time.
// The work list of a launch: list_in (nullptr: the identity list) of n_in
// slots, claimed through the counter queue by grid_lanes lanes.
struct WorkList {
    const uint32_t* list_in;
    uint32_t n_in;
    uint32_t* queue;
    uint32_t grid_lanes;
};

__device__ __forceinline__ void stage_fill(WaveStage* S, const WorkList& a,
                                           const PhotonPtrs& p, double spin, uint32_t base,
                                           uint32_t count)
{
    const uint32_t l = threadIdx.x & 63;
    const uint32_t slot = base + l;
    if (l >= count || slot >= a.n_in) return;
    const uint32_t i = a.list_in ? a.list_in[slot] : slot;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        S->v[q][l] = gmem(p.x[q])[i];
        S->v[4 + q][l] = gmem(p.k[q])[i];
        S->v[8 + q][l] = gmem(p.st[q])[i];
    }
    S->v[12][l] = gmem(p.freq)[i];
    S->v[13][l] = gmem(p.eem)[i];
    S->v[14][l] = gmem(p.dl)[i];
    S->id[l] = gmem(p.id)[i];
    S->draw[l] = gmem(p.rng_draw)[i];
    S->phase[l] = (uint32_t)gmem(p.phase)[i] | ((uint32_t)gmem(p.fpend)[i] << 8);
    S->idx[l] = i;
    S->v[15][l] = ks_radius(S->v[1][l], S->v[2][l], S->v[3][l], spin);
}

// Reads the photon at stage position `pos` into the lane (its radius into
// r); returns its index.
__device__ __forceinline__ uint32_t stage_read(const WaveStage* S, uint32_t pos, Photon& ph,
                                               PhotonRng& rng, double& r)
{
#pragma unroll
    for (int q = 0; q < 8; ++q) ph.s[q] = S->v[q][pos];
#pragma unroll
    for (int q = 0; q < 4; ++q) ph.stokes[q] = S->v[8 + q][pos];
    ph.freq = S->v[12][pos];
    ph.eem = S->v[13][pos];
    ph.dl = S->v[14][pos];
    r = S->v[15][pos];
    ph.phase = (uint8_t)S->phase[pos];
    ph.pend = (uint8_t)(S->phase[pos] >> 8);
    ph.term = 0;
    rng.id = S->id[pos];
    rng.draw = S->draw[pos];
    return S->idx[pos];
}

// Claims near the end of a work list are rounded up to a multiple of this
// many slots (the rest stay in the stage for the wavefront's next takers)
// while the wavefront's lanes take photons quickly: at least one per two
// loop iterations so far, so a photon left in the stage waits a few
// iterations at most.  Claiming exactly what the lanes need meant about
// 49,000 claims of a few photons each in the last quarter of a 1024^2 image
// render (paths of about 40 attempts), each an atomic on the one work counter
// and a load pass: rounding to 8 ran it 2.6-3.3% faster.  Where paths are
// long (RK4 disk photons, up to a launch's 256 calls) the photons a
// wavefront holds back wait for its own lanes while others idle, and
// rounding there ran the launch 18% slower; the rate test keeps those claims
// exact.  Results are the same either way.
constexpr uint32_t kStageClaimRound = 8;

// Hands the next work-list slots, in order, to the lanes with `need` set.
// When the stage runs out it claims the next slots (the first claim of a
// launch statically, then with one atomic on the work counter) and loads
// them: 64 at a time while more than two photons per
// lane of the grid remain, then as many as the wavefront needs (rounded up to
// kStageClaimRound, see there), so the end of the list is spread over the
// wavefronts nearly as evenly as single claims would.  Must be called with
// all 64 lanes active (st stays wave-uniform).
// Returns whether this lane received a photon (then ph, rng, idx and the
// photon's radius r hold it).
__device__ __forceinline__ bool stage_take(WaveStage* S, StageState& st, const WorkList& a,
                                           const PhotonPtrs& p, double spin, bool need, Photon& ph,
                                           PhotonRng& rng, uint32_t& idx, double& r)
{
    const uint64_t mask = __ballot(need);
    const uint32_t total = (uint32_t)__popcll(mask);
    const uint32_t rank = lane_rank(mask);
    const uint32_t avail = st.end - st.next;
    bool got = false;
    if (need && rank < avail && st.next + rank < a.n_in) {
        idx = stage_read(S, st.next + rank - st.base, ph, rng, r);
        got = true;
    }
    if (total > avail && !st.drained) {
        // the launch's lane count from the launch arguments: gridDim and
        // blockDim would be reloaded from the dispatch packet here (a global
        // load and a wait inside the loop)
        const uint32_t lanes = a.grid_lanes;
        const bool quick = 2ull * st.taken >= (uint64_t)st.iters;
        const uint32_t short_claim =
            quick ? (total - avail + kStageClaimRound - 1) & ~(kStageClaimRound - 1) : total - avail;
        const uint32_t want =
            (st.iters == 0 || (uint64_t)st.end + 2ull * lanes < (uint64_t)a.n_in) ? 64u
                                                                                  : short_claim;
        uint32_t nbase;
        if (st.iters == 0) {
            // The first claim of a launch is static: wavefront w of the grid
            // takes slots [64 w, 64 w + 64), which cover [0, grid_lanes) once,
            // and the work counter hands out the slots from grid_lanes on.
            // (Every wavefront's first claim on the one counter at once held
            // the start of a 1024^2 image render back by 1%.)
            nbase = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 64u;
        } else {
            uint32_t c = 0;
            if ((threadIdx.x & 63) == 0) c = atomicAdd(a.queue, want);
            nbase = __builtin_amdgcn_readfirstlane(c) + lanes;
        }
        stage_fill(S, a, p, spin, nbase, want);
        const uint32_t slot = nbase + (rank - avail);
        if (need && rank >= avail && slot < a.n_in) {
            idx = stage_read(S, slot - nbase, ph, rng, r);
            got = true;
        }
        st.base = nbase;
        st.next = nbase + (total - avail);
        st.end = nbase + want;
        st.drained = (uint64_t)st.end >= (uint64_t)a.n_in ? 1u : 0u;
    } else {
        st.next += total < avail ? total : avail;
    }
    return got;
}

__device__ __forceinline__ void store_photon(const PhotonPtrs& p, uint32_t idx, const Photon& ph,
                                             const PhotonRng& rng, double au, double av)
{
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        gmem(p.x[q])[idx] = ph.s[q];
        gmem(p.k[q])[idx] = ph.s[4 + q];
        gmem(p.st[q])[idx] = ph.stokes[q];
    }
    gmem(p.freq)[idx] = ph.freq;
    gmem(p.eem)[idx] = ph.eem;
    gmem(p.dl)[idx] = ph.dl;
    gmem(p.phase)[idx] = ph.phase;
    gmem(p.term)[idx] = ph.term;
    if (ph.phase == kArrived) {
        gmem(p.udisp)[idx] = au;
        gmem(p.vdisp)[idx] = av;
    }
    gmem(p.rng_draw)[idx] = rng.draw;
    gmem(p.fpend)[idx] = ph.pend;
}

#ifndef GRT_TRACE_WAVES_PER_EU
#define GRT_TRACE_WAVES_PER_EU 2
#endif

// ---------------------------------------------------------------------------
// Geodesic checkpoints of the differentiable form.
//
// A photon's functor calls are cut into segments of kSegmentCalls.  At the
// start of every segment the photon's state is written to checkpoint slot
// `committed`; the slot is kept (committed + 1) only if something in the
// segment touched Stokes I (a radiative-transfer update or an event that sets
// I outright).  Per photon a 64-bit header holds the calls made so far (low
// 32 bits), the committed count (bits 32-62) and the "segment touched I" flag
// (bit 63).  The adjoint replays each kept segment from its checkpoint.
// ---------------------------------------------------------------------------

constexpr uint64_t kTapeActive = 1ull << 63;

struct TapeFlag {
    bool touched = false;
    GRT_HD void rt(const FluidCellCGS&, uint32_t, double, double, double) { touched = true; }
    GRT_HD void reset() { touched = true; }
};

__device__ __forceinline__ void write_checkpoint(double* c, const Photon& ph, const PhotonRng& rng,
                                                 uint32_t call0)
{
#pragma unroll
    for (int q = 0; q < 8; ++q) c[q] = ph.s[q];
#pragma unroll
    for (int q = 0; q < 4; ++q) c[8 + q] = ph.stokes[q];
    c[12] = ph.freq;
    c[13] = ph.eem;
    c[14] = ph.dl;
    c[15] = (double)ph.phase;
    c[16] = (double)rng.draw;
    c[17] = (double)call0;
    c[18] = __longlong_as_double((long long)rng.id);
    c[19] = 0.0;
}

__device__ __forceinline__ void read_checkpoint(const double* c, Photon& ph, PhotonRng& rng,
                                                uint32_t& call0)
{
#pragma unroll
    for (int q = 0; q < 8; ++q) ph.s[q] = c[q];
#pragma unroll
    for (int q = 0; q < 4; ++q) ph.stokes[q] = c[8 + q];
    ph.freq = c[12];
    ph.eem = c[13];
    ph.dl = c[14];
    ph.phase = (uint8_t)c[15];
    ph.term = 0;
    rng.draw = (uint32_t)c[16];
    call0 = (uint32_t)c[17];
    rng.id = (uint64_t)__double_as_longlong(c[18]);
}

// Header update at the end of a functor call; returns the new header.
__device__ __forceinline__ uint64_t tape_end_call(uint64_t h, bool touched, bool term, int slots,
                                                  uint32_t* overflow)
{
    const uint32_t calls = (uint32_t)h + 1u;
    uint32_t committed = (uint32_t)(h >> 32) & 0x7FFFFFFFu;
    bool active = (h & kTapeActive) != 0 || touched;
    if ((calls % kSegmentCalls) == 0 || term) {
        if (active) {
            if (committed < (uint32_t)slots) ++committed;
            else atomicOr(overflow, 1u);
        }
        active = false;
    }
    return ((uint64_t)committed << 32) | calls | (active ? kTapeActive : 0ull);
}

// The persistent loop shared by the tracing kernels; Rhs and kMixed select
// the arithmetic of the step (see lane_iterate), kTape the geodesic
// checkpoints of the differentiable form.
template <int kIntegrator, int kSep, class Rhs, int kMixed, int kTape = 0>
__device__ __forceinline__ void trace_body(const TraceLaunch& a_in)
{
    extern __shared__ double lds[];
    // a_in stays in the argument segment (its constants are read from there,
    // see below); only the work list is resolved into registers
    const TraceLaunch& a = a_in;
    WorkList W{a_in.list_in, a_in.n_in, a_in.queue, a_in.grid_lanes};
    if (a_in.n_in_dev) {
        W.n_in = *a_in.n_in_dev;
        if (W.n_in == a_in.n_total) W.list_in = nullptr;
    }
    GridView G = a.grid;
    if (kSep) {
        // Stage the separable search tables in LDS (they always fit: larger
        // tables make the grid take the general path).  An unconditional
        // G.tab = lds lets the compiler address the searches as LDS.
        for (int t = threadIdx.x; t < a.lds_doubles; t += blockDim.x) lds[t] = G.tab[t];
        __syncthreads();
        G.tab = lds;
    }
    // The launch constants (step parameters, photon array pointers, grid) are read
    // through a constant-address-space pointer that the loop re-derives
    // opaquely every iteration: scalar loads at their use sites, instead of
    // some 130 scalar registers held across the whole loop and spilled to
    // vector lanes.
    // The constants are TraceLaunch::cst, at offset 0 of the kernel's
    // argument segment (every tracing kernel takes one TraceLaunch).
    typedef __attribute__((address_space(4))) const TraceConsts ConstTC;
    static_assert(offsetof(TraceLaunch, cst) == 0, "TraceLaunch::cst must come first");
    const ConstTC* c0 = (const ConstTC*)__builtin_amdgcn_kernarg_segment_ptr();
    const StepParams& P = *(const StepParams*)&c0->P;
    const PhotonPtrs& PH = *(const PhotonPtrs*)&c0->ph;

    // This wavefront's photon stage follows the tables.
    WaveStage* S = reinterpret_cast<WaveStage*>(lds + ((a.lds_doubles + 1) & ~1)) +
                   (threadIdx.x >> 6);
    StageState st{0u, 0u, 0u, 0u, 0u, 0u};

    LaneCounters n{0u, 0u, 0u, 0u, 0u, 0u};
    uint32_t last_term = 0;

    Lane L;
    PhotonRng rng;
    rng.seed = P.seed;
    rng.stream = P.stream << 1;
    uint32_t idx = 0;
    bool have = stage_take(S, st, W, PH, P.a, true, L.ph, rng, idx, L.r_open);
    uint64_t thdr = 0;
    if (kTape && have) thdr = a.tape_hdr[idx];
    lane_reset_calls(L);
    int calls_left = a.budget;
    int calls_done = 0;

    // The loop runs until no lane of the wavefront holds a photon; lanes
    // without one sit out the step and join the refills, which need the
    // whole wavefront.
    while (__ballot(have) != 0) {
        const ConstTC* cp = c0;
        asm volatile("" : "+s"(cp));
        const StepParams& P = *(const StepParams*)&cp->P;
        const PhotonPtrs& PH = *(const PhotonPtrs*)&cp->ph;
        GridView G = *(const GridView*)&cp->grid;
        if (kSep) G.tab = lds;
        bool need = false;
        if (have) {
            double au = 0.0, av = 0.0;
            if (kTape && !L.in_call && ((uint32_t)thdr % kSegmentCalls) == 0) {
                const uint32_t k = (uint32_t)(thdr >> 32) & 0x7FFFFFFFu;
                if (k < (uint32_t)a.tape_slots)
                    write_checkpoint(a.tape + ((size_t)idx * a.tape_slots + k) * kCkptDoubles,
                                     L.ph, rng, (uint32_t)thdr);
            }
            TapeFlag flag;
            bool call_done;
            if (kTape)
                call_done = lane_iterate<kIntegrator, kSep, Rhs, kMixed, TapeFlag>(
                    L, G, P, rng, n, au, av, flag);
            else
                call_done = lane_iterate<kIntegrator, kSep, Rhs, kMixed, 1>(L, G, P, rng, n, au,
                                                                            av);
            if (kTape && flag.touched) thdr |= kTapeActive;
            if (call_done) {
                ++n.calls;
                ++calls_done;
                --calls_left;
                if (kTape)
                    thdr = tape_end_call(thdr, false, L.ph.term != 0, a.tape_slots,
                                         a.tape_overflow);
                if (L.ph.term || calls_left == 0) {
                    store_photon(PH, idx, L.ph, rng, au, av);
                    if (kTape) a.tape_hdr[idx] = thdr;
                    if (L.ph.term) {
                        const uint32_t step = (uint32_t)(a.step0 + calls_done - 1);
                        last_term = step > last_term ? step : last_term;
                    } else {
                        const uint32_t o = wave_claim(a.n_out);
                        a.list_out[o] = idx;
                    }
                    need = true;
                }
            }
        }
        ++st.iters;
        const uint64_t need_mask = __ballot(need);
        if (need_mask != 0) {
            st.taken += (uint32_t)__popcll(need_mask);
            const bool got = stage_take(S, st, W, PH, P.a, need, L.ph, rng, idx, L.r_open);
            if (need) {
                have = got;
                if (got) {
                    if (kTape) thdr = a.tape_hdr[idx];
                    lane_reset_calls(L);
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
    const unsigned long long s_hits = wave_sum(n.hits);
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
        atomicAdd(&a.stats[kStatHits], s_hits);
        atomicMax(&a.stats[kStatLastTermStep], (unsigned long long)last_term);
    }
}

}  // namespace
}  // namespace grt
