/*
 * grt_oracle.h -- CPU FP64 restatement of GR-ray-tracer's per-photon hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against (and, timed under OpenMP, the CPU baseline).  Nothing in
 * the product library links, loads or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference's src/ directory).  All arithmetic is IEEE binary64 in the
 * same operation order as the reference, built without automatic FMA
 * contraction; the integrators fuse products into sums through explicit fma
 * calls by default (orc_set_unfused_rounding selects the unfused rounding).
 *
 * RNG boundary: the reference draws from Kokkos::Random_XorShift64_Pool, whose
 * streams cannot be reproduced outside Kokkos.  The oracle plugs in the same
 * counter-based Philox4x32-10 stream as the HIP path (keyed by seed, photon id,
 * stream and draw index) so that GPU-vs-oracle comparisons of RNG-driven
 * branches are exact; parity of those branches with the reference's own
 * XorShift64 stream is statistical only.
 */
#ifndef GRT_ORACLE_H
#define GRT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Functor constructor arguments, integrate_cart_ks.hpp:35-61,63-94. */
typedef struct {
    double a_bh, m_bh;
    double r_term_min, r_term_max;
    double atol, rtol, min_step_scale, max_step_scale, safety;
    int32_t integrator;       /* 0 = RK4, 1 = RK45 (utils.hpp:169) */
    int32_t mode;             /* 0 = Image, 1 = Disk (simulation_options.hpp:11) */
    /* RadiativeTransferModel, radiative_transfer.hpp:54-59 */
    int32_t absorption_enabled, emission_enabled;
    double absorption_coefficient, emission_coefficient;
    /* ScatteringModel, scattering.hpp:10-14 */
    int32_t scattering_enabled;
    double albedo;
    uint64_t seed;
    uint32_t rng_stream;      /* Philox stream selector (no reference counterpart) */
    /* PhotonGenerationConfig, photon_generation.hpp:40-51 */
    int32_t generator;        /* 0 blackbody, 1 power-law, 2 custom */
    double superphoton_count_normalization, energy_per_cell_erg, power_law_slope;
    double nu_min_hz, nu_max_hz, custom_frequency_hz;
    /* camera, utils.hpp:149-153 */
    double camera_distance, camera_theta, camera_phi, camera_width, camera_height;
} orc_params;

/* ScatteringFluidGrid, scattering.hpp:22-39: C-order f8 arrays of shape
 * (nr, nth, nph) and (nr, nth, nph, 4), exactly the .npy layout. */
typedef struct {
    int32_t nr, nth, nph;
    const double *r, *theta, *phi, *rho, *tgas, *vel, *mag;
    double length_scale, time_scale, density_scale, temperature_scale;
    double velocity_scale, magnetic_scale;
    double spin, mass;
    int32_t available;
} orc_grid;

/* Photons SoA, utils.hpp:40-54. */
typedef struct {
    uint64_t *id;
    double *frequency, *emission_frame_energy;
    double *x0, *x1, *x2, *x3, *k0, *k1, *k2, *k3;
    double *I, *Q, *U, *V, *dlambda;
    uint8_t *terminate, *phase;
    double *theta_disp, *phi_disp;
} orc_photons;

/* FluidCellCGS, photon_generation.hpp:87-97 (same field order). */
typedef struct {
    double position_cm[3];
    double coordinate_time_s;
    double density_g_cm3;
    double temperature_k;
    double magnetic_gauss[4];
    double coordinate_four_velocity_cm_s[4];
} orc_fluid;

typedef struct {
    uint64_t functor_calls;   /* operator() invocations on live photons */
    uint64_t accepted_steps;  /* accepted RK45 steps or RK4 steps */
    uint64_t rejected_attempts;
    uint64_t failed_steps;    /* 12 consecutive rejections */
    uint64_t fluid_samples;   /* sample_scattering_fluid calls */
    uint64_t global_steps;    /* host loop iterations executed */
    uint64_t terminated;
    uint64_t fluid_hits;      /* samples that found a cell (the call returned true) */
} orc_stats;

/* Per-photon RNG cursor: Philox counter words (id_lo, id_hi, draw, stream). */
typedef struct { uint64_t id; uint32_t draw; uint32_t stream; uint64_t seed; } orc_rng;

/* ---- metric core, metrics/kerr_schild_core.hpp ---- */
double orc_compute_r(double x, double y, double z, double a);
void orc_compute_r_and_deriv(double x, double y, double z, double a, double *r, double dr[3]);
void orc_compute_metric(const double X[4], double a, double m, double g[16]);
void orc_compute_inverse_metric(const double X[4], double a, double m, double ginv[16]);
void orc_compute_inverse_metric_deriv(const double X[4], double a, double m, double dginv[64]);
void orc_hamiltonian_rhs(const double X[4], const double p[4], double a, double m, double rhs[8]);
void orc_hamiltonian_rhs_fused(const double state[8], double a, double m, double rhs[8]);
/* 0 (default): the fused rounding the tracing kernel reproduces; 1: unfused */
void orc_set_unfused_rounding(int unfused);
int orc_null_covariant_momentum(const double X[4], const double K[3], double a, double m, double p[4]);
double orc_inverse_norm(const double X[4], const double V[4], double a, double m);
void orc_geodesic_observables(const double state[8], double a, double m, double out[4]);

/* ---- RK, geodesic_integrate/{rk_integrators,integrate_cart_ks}.hpp ---- */
typedef void (*orc_rhs_fn)(const double *in, double *out, void *user);
void orc_rk4_generic(int n, double *state, double dt, orc_rhs_fn f, void *user);
int orc_rk45_generic(int n, double *state, double *dt, const double *k1, orc_rhs_fn f, void *user,
                     double atol, double rtol, double min_s, double max_s, double safety);
void orc_rk4_step_ks(double state[8], double dt, double a, double m);
int orc_rk45_step_ks(const orc_params *p, double state[8], double *dlambda, const double k1[8]);
double orc_quarter_root(double x);

/* ---- fluid sampling, radiative_transfer/{scattering,fluid_frame}.hpp ---- */
int orc_locate_cell(const orc_grid *g, double radius, double polar, double azimuth,
                    int32_t ijk[3]);
void orc_spherical_to_cartesian_vector(double r, double th, double ph, double spin,
                                       const double sph[4], double cart[4]);
int orc_build_fluid_tetrad(const double X[4], const double u[4], double spin, double mass,
                           double e[16]);
int orc_sample_fluid(const double state[8], const orc_grid *g, orc_fluid *fluid, double tetrad[16]);

/* ---- radiative transfer, radiative_transfer.hpp:54-96 ---- */
double orc_optical_depth_increment(const orc_fluid *f, double nu, double dl, double coeff);
double orc_emissivity_increment(const orc_fluid *f, double nu, double dl, double coeff);
double orc_optical_depth_increment_grad(const orc_fluid *f, double nu, double dl, double coeff,
                                        double d[3]);
double orc_emissivity_increment_grad(const orc_fluid *f, double nu, double dl, double coeff,
                                     double d[3]);
void orc_apply_radiative_transfer(double stokes[4], const orc_fluid *f, double nu, double dl,
                                  int absorb, int emit, double kabs, double kemit);

/* ---- RNG ---- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double orc_rng_drand(orc_rng *rng);

/* ---- scattering / generation ---- */
int orc_maybe_scatter(double state[8], double stokes[4], int enabled, double albedo,
                      double dlambda_abs, double length_scale, const orc_fluid *f,
                      const double tetrad[16], double spin, double mass, orc_rng *rng);
int orc_superphotons_per_cell(double density_code, double normalization);
int orc_generate_photon(const orc_params *p, const orc_fluid *cell, orc_rng *rng, int packets,
                        double *nu, double *energy);
/* Test helper: n draws of orc_generate_photon for one cell, draw i from the
 * Philox stream of photon id i (stream 1, seed p->seed); nu[i] = 0 where the
 * generator refused. */
void orc_sample_generator(const orc_params *p, const orc_fluid *cell, int packets, int64_t n,
                          double *nu, double *energy);

/* ---- camera, integrate_cart_ks.hpp:96-130 ---- */
int orc_camera_plane_crossing(const orc_params *p, const double state[8], double *u, double *v);
double orc_camera_observer_energy(const orc_params *p, const double state[8]);

/* ---- per-photon functor and host loop ---- */
void orc_step_photon(const orc_params *p, const orc_grid *g, orc_photons *ph, int64_t idx,
                     uint32_t *rng_draw, orc_stats *st);
void orc_trace(const orc_params *p, const orc_grid *g, orc_photons *ph, int64_t n,
               int32_t max_steps, double termination_percent, int64_t output_interval,
               int32_t nthreads, orc_stats *st);
void orc_trace_adjoint(const orc_params *p, const orc_grid *g, orc_photons *ph, int64_t n,
                       int32_t max_steps, double termination_percent, int64_t output_interval,
                       int32_t nthreads, const double *dL_dI, double *grad_rho, double *grad_T,
                       double grad_coeff[2], orc_stats *st);

/* ---- initialisation and observation products ---- */
void orc_init_image_photon(const orc_params *p, double screen_u, double screen_v,
                           double initial_step, double state[8], double *dlambda);
void orc_init_image(const orc_params *p, const double *screen_u, const double *screen_v,
                    int64_t n, uint64_t id_offset, double initial_step, orc_photons *ph);
void orc_init_disk(const orc_params *p, const orc_grid *g, int64_t n, uint64_t id_offset,
                   double initial_step, orc_photons *ph, uint64_t *total_out);
int64_t orc_observation_bin(double value, double lo, double hi, int64_t bins);
void orc_image_products(const orc_photons *ph, int64_t n, int64_t nx, int64_t ny,
                        int64_t spectrum_bins, double spec_min, double spec_max,
                        double width, double height, double *image /* 4*ny*nx */,
                        double *spectrum_nu, double *spectrum_I, uint64_t *arrived);

#ifdef __cplusplus
}
#endif
#endif
