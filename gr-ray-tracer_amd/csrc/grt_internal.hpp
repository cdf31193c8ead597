// grt_internal.hpp -- interface between the host driver (grt_api.cpp) and
// the HIP kernels (grt_kernels.hip).  Not part of the public C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "grt.h"
#include "grt_device.hpp"

namespace grt {

// Device copy of the photon SoA (utils.hpp:40-54) plus the per-photon RNG
// draw counter the library keeps alongside it.
struct PhotonPtrs {
    uint64_t* id;
    double* freq;
    double* eem;
    double* x[4];
    double* k[4];
    double* st[4];
    double* dl;
    uint8_t* term;
    uint8_t* phase;
    double* udisp;
    double* vdisp;
    uint32_t* rng_draw;
    uint8_t* fpend;     // library-owned: thermal frequency still to draw
};

enum StatSlot {
    kStatCalls = 0,
    kStatAccepted,
    kStatRejected,
    kStatFailed,
    kStatSamples,
    kStatLastTermStep,   // max over photons of the global step on which they terminated
    kStatHits,           // fluid samples that found a cell with valid fluid and tetrad
    kStatCount
};

// Launch constants.  A tracing launch carries them at the start of its
// kernel arguments (TraceLaunch::cst), which the kernel reads through a
// constant-address-space pointer to its argument segment (scalar loads, no
// copy before the launch, no wait on the kernel's own photon stores); the
// adjoint reads a device copy.
struct TraceConsts {
    PhotonPtrs ph;
    StepParams P;
    GridView grid;   // the launch's grid (the loop swaps in the LDS tables)
};

// Counters of the tracing driver: [0] work queue, [1] photons still live
// after a launch, [2] live count before the first launch, [3] tape
// overflow, [4] adjoint queue; then kStatCount u64 statistics (the word
// count keeps them 8-byte aligned), the block padded to a multiple of 16
// bytes (the runtime clears a block that is not with two fill kernels).
// Read back together in one copy.
constexpr size_t kCounterWords = 8;
constexpr size_t kReadbackBytes =
    (kCounterWords * sizeof(uint32_t) + kStatCount * sizeof(unsigned long long) + 15) & ~size_t(15);

// The device scratch of one trace call: the context's own allocations, or
// the fixed array of the no-allocation entry point (grt_noalloc.hip).
struct TraceScratch {
    uint32_t* lists;      // two work lists of list_cap entries
    size_t list_cap;
    uint32_t* rng;        // per-photon RNG draw counters
    uint8_t* fpend;       // per-photon deferred-frequency flags
    uint32_t* counters;   // kReadbackBytes: driver counters, then the statistics
    bool fill_kernel;     // clear with launch_zero rather than hipMemsetAsync
};

// grt_noalloc.hip.  noalloc_scratch_init looks up this device's copy of the
// fixed scratch array (call once per device, with that device current);
// noalloc_scratch carves it for a trace of at most GRT_NOALLOC_MAX_PHOTONS.
hipError_t noalloc_scratch_init(int device);
hipError_t noalloc_scratch(int device, TraceScratch* s);
// Zero-fills [p, p + bytes) with a kernel on `stream`.
hipError_t launch_zero(void* p, size_t bytes, hipStream_t stream);

// One wavefront's photon stage in LDS (grt_trace.hpp): the photons of the
// next (up to 64) work-list slots, loaded with one coalesced pass over the
// photon arrays, so a lane that needs a new photon reads it from LDS instead of
// waiting for HBM.  Placed after the search tables, one per wavefront.
// s[8], stokes[4], freq, eem, dl, and the Kerr-Schild radius of x (computed
// once per staged photon, with all 64 lanes, for the next call's radius tests)
constexpr int kStageDoubles = 16;
struct WaveStage {
    double v[kStageDoubles][64];
    uint64_t id[64];
    uint32_t draw[64];
    uint32_t idx[64];
    uint32_t phase[64];
};

// Bookkeeping of a wavefront's stage, wave-uniform (kept in scalar registers).
struct StageState {
    uint32_t base;      // work-list slot of stage position 0
    uint32_t next;      // next slot to hand out, base <= next <= end
    uint32_t end;       // one past the last staged slot, end <= base + 64
    uint32_t drained;   // the work counter has passed the end of the list
    uint32_t iters;     // loop iterations of the wavefront so far
    uint32_t taken;     // photons its lanes took after the first one each
};

// Dynamic LDS of a tracing launch: search tables, then one stage per wavefront.
inline size_t trace_lds_bytes(int lds_doubles, int block)
{
    const size_t tables = (((size_t)lds_doubles + 1) & ~(size_t)1) * sizeof(double);
    return tables + (size_t)(block / 64) * sizeof(WaveStage);
}

struct TraceLaunch {
    // first: the kernels find it at offset 0 of their argument segment
    TraceConsts cst;
    GridView grid;
    int integrator;
    const uint32_t* list_in;  // nullptr: the identity list (slot i is photon i)
    uint32_t n_in;
    // When set, the kernel reads n_in from here (the live count build_live
    // wrote just before, so the host need not wait for it) and walks the
    // identity list instead of list_in if all n_total photons are live.
    const uint32_t* n_in_dev;
    uint32_t n_total;
    uint32_t* queue;          // work counter over the slots of list_in from grid_lanes on,
                              // zeroed before launch
    uint32_t* list_out;       // photons still live after `budget` calls
    uint32_t* n_out;          // zeroed before launch
    unsigned long long* stats;  // kStatCount slots
    uint32_t grid_lanes;      // threads of the launch (blocks * block)
    int budget;               // functor calls per photon in this launch
    int step0;                // global step index of the first call of this launch
    int lds_doubles;          // separable search tables staged in LDS (doubles)
    int mixed;                // GRT_PRECISION_*: 1 runs the mixed FP32/FP64 kernel, 2 its
                              // FP64 control (both in grt_mixed.hip)
    // differentiable form: geodesic checkpoints (tape != nullptr selects the
    // checkpointing kernel)
    double* tape;             // [photon][tape_slots][kCkptDoubles]
    uint64_t* tape_hdr;       // per photon: calls | committed << 32 | touched << 63
    int tape_slots;
    uint32_t* tape_overflow;  // set when a photon needs more than tape_slots
};

// Reverse sweep of the differentiable form (grt_adjoint.hip).
struct AdjointLaunch {
    const TraceConsts* consts;
    GridView grid;
    int integrator;
    const double* tape;
    const uint64_t* tape_hdr;
    int tape_slots;
    const double* seed;       // dL/dI of every photon's final Stokes I
    uint32_t n;
    uint32_t* queue;          // zeroed before launch
    double* rec;              // per-lane replay records, interleaved by lane
    uint32_t rec_lanes;       // lanes the record buffer holds
    double* grad_rho;         // per cell, code units
    double* grad_T;
    double* grad_coeff;       // [absorption, emission]
    double density_scale, temperature_scale;
    int lds_doubles;
};

// Kernel launchers (grt_kernels.hip).  They only enqueue work on `stream`.
hipError_t launch_trace(const TraceLaunch& a, int blocks, int block, size_t lds_bytes,
                        hipStream_t stream);
// grt_mixed.hip
hipError_t launch_trace_mixed(const TraceLaunch& a, int blocks, int block, size_t lds_bytes,
                              hipStream_t stream);
// grt_adjoint.hip
hipError_t launch_trace_taped(const TraceLaunch& a, int blocks, int block, size_t lds_bytes,
                              hipStream_t stream);
hipError_t launch_tape_finalize(uint64_t* hdr, uint64_t n, int slots, uint32_t* overflow,
                                hipStream_t stream);
hipError_t launch_adjoint(const AdjointLaunch& a, int blocks, int block, size_t lds_bytes,
                          hipStream_t stream);
hipError_t launch_set_fluid_scalars(double* cells, double* rho_code, const double* rho,
                                    const double* tgas, uint64_t ncell, double density_scale,
                                    double temperature_scale, hipStream_t stream);
constexpr uint32_t kSegmentCalls = 32;   // functor calls per geodesic checkpoint
constexpr int kCkptDoubles = 20;   // s[8], stokes[4], freq, eem, dl, phase, draw, call0, id, pad
constexpr int kRecEntries = 64;    // replay records per lane: <= 2 per call
// cell (or -1: reset), I before, step, frequency; then, where the gradient
// companions read them (GRT_GRADIENT_READS_GEOMETRY), position, time and the
// magnetic field.  Density and temperature come back from the cell record.
constexpr int kRecDoubles = GRT_GRADIENT_READS_GEOMETRY ? 12 : 4;
// The live list and its count, and the call's draw counters and
// deferred-frequency flags set to zero (grt_kernels.hip).
hipError_t launch_build_live(const uint8_t* term, uint64_t n, uint32_t* list, uint32_t* count,
                             uint32_t* draw, uint8_t* fpend, hipStream_t stream);
// Draws the deferred thermal frequencies after a launch (grt_kernels.hip).
hipError_t launch_resolve_deferred(const PhotonPtrs& ph, uint64_t n, uint64_t seed,
                                   uint32_t stream_id, double a, double m, hipStream_t stream);
hipError_t launch_image_products(const PhotonPtrs& ph, uint64_t n, int64_t nx, int64_t ny,
                                 int64_t bins, double smin, double smax, double lmin, double lmax,
                                 double width, double height, double* image, double* spec,
                                 unsigned long long* arrived, hipStream_t stream);
hipError_t launch_init_image(const PhotonPtrs& ph, uint64_t n, uint64_t id0, const CameraInit& C,
                             double a, double m, double w, double h, uint64_t seed,
                             uint32_t stream, const double* su, const double* sv,
                             hipStream_t hs);
hipError_t disk_offsets(const double* rho_code, uint64_t ncell, double norm,
                        unsigned long long* counts, unsigned long long* off, void* temp,
                        size_t* temp_bytes, hipStream_t stream);
hipError_t launch_emit_disk(const PhotonPtrs& ph, uint64_t n, uint64_t id0, double initial_step,
                            const GridView& G, const StepParams& P, uint32_t stream,
                            const unsigned long long* off, uint64_t ncell, hipStream_t hs);
hipError_t trace_kernel_attributes(int* vgprs, int* sgprs, int* lds_static, int* max_threads);

}  // namespace grt
