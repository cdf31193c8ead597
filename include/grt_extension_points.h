/*
 * grt_extension_points.h -- the user-editable physics plug-ins of the traced
 * path, compiled into the HIP tracing kernel.
 *
 * Same names, parameters and FluidCellCGS field order as the reference's
 * compile-time plug-ins (edit the body, rebuild):
 *   FluidCellCGS                     photon_generation.hpp:87-97
 *   ComputeOpticalDepthIncrement     radiative_transfer.hpp:62-67
 *   ComputeEmissivityIncrement       radiative_transfer.hpp:70-75
 *   ComputeSuperphotonsPerCell       initialize_photons.hpp:17-19
 *   GeneratePhotonsCustom            photon_generation.hpp:149-156
 * The bodies below are the reference defaults.  Kokkos:: math becomes the
 * plain HIP device math functions (fmax, fabs, exp, ...).
 *
 * The differentiable form (grt_trace_taped / grt_adjoint) also needs the
 * partial derivatives of the two radiative-transfer increments; they come
 * from the two ...Gradient companions below, which have no counterpart in
 * the reference.  Whoever edits an increment edits its companion too; the
 * test suite checks each companion against central differences of its
 * increment.
 *
 * GRT_GRADIENT_READS_GEOMETRY declares whether the two companions read
 * fluid.position_cm, fluid.coordinate_time_s or fluid.magnetic_gauss.  The
 * adjoint keeps those fields of every transfer step for its reverse sweep
 * only when it is 1; with 0 (the default companions read only the density,
 * the temperature and the step) they reach the companions as NaN, so a
 * companion that reads them without the declaration gives NaN gradients,
 * not wrong ones.  The increments themselves always see every field.
 *
 * `real` is double in the FP64 path.  The RNG argument of
 * GeneratePhotonsCustom offers `double drand()` returning a uniform [0, 1)
 * draw from the per-photon counter-based stream.
 */
#ifndef GRT_EXTENSION_POINTS_H
#define GRT_EXTENSION_POINTS_H

#include <hip/hip_runtime.h>

namespace grt_ext {
using real = double;
}

#ifndef GRT_GRADIENT_READS_GEOMETRY
#define GRT_GRADIENT_READS_GEOMETRY 0
#endif

struct FluidCellCGS {
    double position_cm[3] = {0.0, 0.0, 0.0};
    double coordinate_time_s = 0.0;
    double density_g_cm3 = 0.0;
    double temperature_k = 0.0;
    /* orthonormal components in the fluid tetrad (G) */
    double magnetic_gauss[4] = {0.0, 0.0, 0.0, 0.0};
    /* Cartesian Kerr-Schild coordinate components (cm/s) */
    double coordinate_four_velocity_cm_s[4] = {1.0, 0.0, 0.0, 0.0};
};

/* Optical depth dtau >= 0 accumulated over one step of affine length
 * path_length_step (signed; negative while image rays run forward). */
__host__ __device__ inline double ComputeOpticalDepthIncrement(const FluidCellCGS& fluid,
                                                               double frequency_hz,
                                                               double path_length_step,
                                                               double absorption_coefficient)
{
    (void)frequency_hz;
    return fmax(0.0, absorption_coefficient * fluid.density_g_cm3 * fabs(path_length_step));
}

/* Intensity added to Stokes I over one step. */
__host__ __device__ inline double ComputeEmissivityIncrement(const FluidCellCGS& fluid,
                                                             double frequency_hz,
                                                             double path_length_step,
                                                             double emission_coefficient)
{
    (void)frequency_hz;
    return fmax(0.0, emission_coefficient * fluid.density_g_cm3 * fluid.temperature_k *
                         fabs(path_length_step));
}

/* dtau and its partial derivatives with respect to fluid.density_g_cm3,
 * fluid.temperature_k and absorption_coefficient (zero where the default
 * body's fmax(0, .) clamps). */
__host__ __device__ inline double ComputeOpticalDepthIncrementGradient(
    const FluidCellCGS& fluid, double frequency_hz, double path_length_step,
    double absorption_coefficient, double* d_density, double* d_temperature,
    double* d_coefficient)
{
    const double dtau = ComputeOpticalDepthIncrement(fluid, frequency_hz, path_length_step,
                                                     absorption_coefficient);
    const double len = fabs(path_length_step);
    const bool live = dtau > 0.0;
    *d_density = live ? absorption_coefficient * len : 0.0;
    *d_temperature = 0.0;
    *d_coefficient = live ? fluid.density_g_cm3 * len : 0.0;
    return dtau;
}

/* dI and its partial derivatives with respect to fluid.density_g_cm3,
 * fluid.temperature_k and emission_coefficient. */
__host__ __device__ inline double ComputeEmissivityIncrementGradient(
    const FluidCellCGS& fluid, double frequency_hz, double path_length_step,
    double emission_coefficient, double* d_density, double* d_temperature,
    double* d_coefficient)
{
    const double dI = ComputeEmissivityIncrement(fluid, frequency_hz, path_length_step,
                                                 emission_coefficient);
    const double len = fabs(path_length_step);
    const bool live = dI > 0.0;
    *d_density = live ? emission_coefficient * fluid.temperature_k * len : 0.0;
    *d_temperature = live ? emission_coefficient * fluid.density_g_cm3 * len : 0.0;
    *d_coefficient = live ? fluid.density_g_cm3 * fluid.temperature_k * len : 0.0;
    return dI;
}

/* Superphotons per cell from the density in code units.  The reference's
 * static_cast<int> is undefined outside the int range; this saturates. */
__host__ __device__ inline int ComputeSuperphotonsPerCell(double density_code_units,
                                                          double normalization)
{
    const double v = density_code_units * normalization;
    if (!(v == v)) return 0;
    if (v >= 2147483647.0) return 2147483647;
    if (v <= -2147483648.0) return -2147483647 - 1;
    return static_cast<int>(v);
}

/* Frequency and packet energy of the "custom" generator. */
template <typename RandomGenerator>
__host__ __device__ inline bool GeneratePhotonsCustom(const FluidCellCGS&, RandomGenerator&,
                                                      int packets_per_cell,
                                                      double configured_frequency_hz,
                                                      double energy_per_cell_erg,
                                                      double& sampled_frequency_hz,
                                                      double& packet_energy_erg)
{
    sampled_frequency_hz = configured_frequency_hz;
    packet_energy_erg = energy_per_cell_erg / static_cast<double>(packets_per_cell);
    return true;
}

#endif
