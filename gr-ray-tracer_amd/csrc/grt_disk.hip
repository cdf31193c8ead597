// grt_disk.hip -- disk-mode packet emission on the device
// (initialize_photons.hpp:17-62 and :173-327).
//
// 1. count:  packets per cell = ComputeSuperphotonsPerCell(rho_code, norm),
//            negative counts clamp to 0 (initialize_photons.hpp:53-54);
// 2. scan:   exclusive prefix sum over the cells in C order (hipCUB);
// 3. emit:   photon i lives in the largest cell c with offset[c] <= i
//            (binary search, initialize_photons.hpp:38-51) and is emitted in
//            the fluid frame of that cell.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "grt_internal.hpp"

namespace grt {

namespace {

__global__ __launch_bounds__(256) void count_packets(const double* __restrict__ rho_code,
                                                     uint64_t ncell, double norm,
                                                     unsigned long long* __restrict__ counts)
{
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const int n = ComputeSuperphotonsPerCell(rho_code[c], norm);
    counts[c] = n > 0 ? (unsigned long long)n : 0ull;
}

// Cell-centre coordinates (r, theta, phi) of cell c.
__device__ __forceinline__ void cell_coords(const GridView& G, uint64_t c, double& r, double& th,
                                            double& ph)
{
    if (G.separable) {
        const uint64_t plane = (uint64_t)G.nth * G.nph;
        const int i = (int)(c / plane);
        const int j = (int)((c / G.nph) % G.nth);
        const int k = (int)(c % G.nph);
        r = G.axes[i];
        th = G.axes[G.nr + j];
        ph = G.axes[G.nr + G.nth + k];
    } else {
        r = G.r3[c];
        th = G.th3[c];
        ph = G.ph3[c];
    }
}

__global__ __launch_bounds__(256) void emit_disk(PhotonPtrs ph, uint64_t n, uint64_t id0,
                                                 double initial_step, GridView G, StepParams P,
                                                 uint32_t stream,
                                                 const unsigned long long* __restrict__ off,
                                                 uint64_t ncell)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // largest cell with off[cell] <= i
    uint64_t lo = 0, hi = ncell;
    while (lo + 1 < hi) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (off[mid] <= i) lo = mid;
        else hi = mid;
    }
    const uint64_t c = lo;
    const int packets = (int)(off[c + 1] - off[c]);
    double rc, thc, phc;
    cell_coords(G, c, rc, thc, phc);
    // spherical_kerr_schild_to_cartesian, fluid_frame.hpp:16-27
    const double st = sin(thc), ct = cos(thc), sp = sin(phc), cp = cos(phc);
    double s[8];
    s[0] = 0.0;
    s[1] = (rc * cp - P.a * sp) * st;
    s[2] = (rc * sp + P.a * cp) * st;
    s[3] = rc * ct;
    const double* rec = G.cells + c * kCellDoubles;
    FluidCellCGS fl;
    fl.position_cm[0] = s[1] * P.length_scale;
    fl.position_cm[1] = s[2] * P.length_scale;
    fl.position_cm[2] = s[3] * P.length_scale;
    fl.coordinate_time_s = s[0] * P.time_scale;
    fl.density_g_cm3 = rec[0];
    fl.temperature_k = rec[1];
#pragma unroll
    for (int q = 0; q < 4; ++q) fl.coordinate_four_velocity_cm_s[q] = rec[2 + q];
    const KSMetric g = ks_metric_at(s, P.a, P.m);
    double e[16];
    const bool valid = fluid_tetrad(g, fl.coordinate_four_velocity_cm_s, e);
    if (valid) {
        const double b[4] = {rec[6], rec[7], rec[8], rec[9]};
        fl.magnetic_gauss[0] = -inner(g, e, b);
        fl.magnetic_gauss[1] = inner(g, e + 4, b);
        fl.magnetic_gauss[2] = inner(g, e + 8, b);
        fl.magnetic_gauss[3] = inner(g, e + 12, b);
    }
    const uint64_t id = id0 + i;
    PhotonRng rng{id, P.seed, stream, 0u};
    double nu = 0.0, energy = 0.0;
    bool gen = false;
    if (valid) gen = generate_photon(P, fl, rng, packets, nu, energy);
    const double mu = 2.0 * rng.drand() - 1.0;
    const double dphi = kTwoPi * rng.drand();
    const double tr = sqrt(fmax(0.0, 1.0 - mu * mu));
    double p[4] = {0.0, 0.0, 0.0, 0.0};
    double eem = 0.0;
    if (gen) {
        // fluid_frame_direction_to_covariant_momentum, fluid_frame.hpp:83-104
        const double dx = tr * cos(dphi), dy = tr * sin(dphi), dz = mu;
        double w[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) w[q] = e[q] + dx * e[4 + q] + dy * e[8 + q] + dz * e[12 + q];
        const double lw = w[0] + g.l1 * w[1] + g.l2 * w[2] + g.l3 * w[3];
        p[0] = -w[0] + g.twoH * lw;
        p[1] = w[1] + g.twoH * g.l1 * lw;
        p[2] = w[2] + g.twoH * g.l2 * lw;
        p[3] = w[3] + g.twoH * g.l3 * lw;
#pragma unroll
        for (int q = 0; q < 4; ++q) eem -= p[q] * e[q];
    }
    ph.id[i] = id;
    ph.freq[i] = nu;
    ph.eem[i] = eem;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        ph.x[q][i] = s[q];
        ph.k[q][i] = p[q];
    }
    ph.st[0][i] = gen ? energy : 0.0;
    ph.st[1][i] = 0.0;
    ph.st[2][i] = 0.0;
    ph.st[3][i] = 0.0;
    ph.dl[i] = initial_step;
    ph.term[i] = gen ? 0 : 1;
    ph.phase[i] = gen ? kDisk : kRejected;
    ph.udisp[i] = 0.0;
    ph.vdisp[i] = 0.0;
}

}  // namespace

// counts and off hold ncell + 1 entries each; temp == nullptr queries the
// scan's scratch size into *temp_bytes.
hipError_t disk_offsets(const double* rho_code, uint64_t ncell, double norm,
                        unsigned long long* counts, unsigned long long* off, void* temp,
                        size_t* temp_bytes, hipStream_t stream)
{
    if (temp == nullptr)
        return hipcub::DeviceScan::ExclusiveSum(nullptr, *temp_bytes, counts, off, ncell + 1,
                                                stream);
    const int block = 256;
    if (ncell > 0) {
        hipLaunchKernelGGL(count_packets, dim3((unsigned)((ncell + block - 1) / block)), dim3(block),
                           0, stream, rho_code, ncell, norm, counts);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    // counts[ncell] = 0, so off[ncell] becomes the total.
    hipError_t e = hipMemsetAsync(counts + ncell, 0, sizeof(unsigned long long), stream);
    if (e != hipSuccess) return e;
    return hipcub::DeviceScan::ExclusiveSum(temp, *temp_bytes, counts, off, ncell + 1, stream);
}

hipError_t launch_emit_disk(const PhotonPtrs& ph, uint64_t n, uint64_t id0, double initial_step,
                            const GridView& G, const StepParams& P, uint32_t stream,
                            const unsigned long long* off, uint64_t ncell, hipStream_t hs)
{
    if (n == 0) return hipSuccess;
    const int block = 256;
    hipLaunchKernelGGL(emit_disk, dim3((unsigned)((n + block - 1) / block)), dim3(block), 0, hs, ph,
                       n, id0, initial_step, G, P, stream, off, ncell);
    return hipGetLastError();
}

}  // namespace grt
