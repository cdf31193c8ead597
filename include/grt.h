/*
 * grt.h -- C ABI of the MI355X-native GR-ray-tracer hot path (libgrt_amd.so).
 *
 * One context owns the device copy of the fluid grid and the scratch state of
 * the persistent tracing kernel.  Every entry point returns 0 on success and a
 * negative GRT_E* code on failure; grt_last_error() then returns the message.
 * No exception crosses this boundary.  Calls on one context are serialised by
 * the caller; separate contexts may be used from separate host threads.
 *
 * Per-photon failures stay in-band exactly as in the reference
 * (terminate / phase / zeroed Stokes), see integrate_cart_ks.hpp:231-419.
 */
#ifndef GRT_H
#define GRT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GRT_ABI_VERSION 2

enum {
    GRT_OK = 0,
    GRT_E_INVALID = -1,   /* bad argument, shape or configuration */
    GRT_E_NOGRID = -2,    /* a fluid-dependent feature with no grid set */
    GRT_E_DEVICE = -3,    /* HIP runtime error */
    GRT_E_NOMEM = -4,     /* device allocation failed */
    GRT_E_STATE = -5,     /* call out of sequence (e.g. adjoint without a taped trace) */
    GRT_E_RANGE = -6      /* a capacity given by the caller was exceeded */
};

enum { GRT_RK4 = 0, GRT_RK45 = 1 };                 /* utils.hpp:169 IntegratorType */
enum { GRT_MODE_IMAGE = 0, GRT_MODE_DISK = 1 };     /* simulation_options.hpp:11 SimulationMode */
enum { GRT_GEN_BLACKBODY = 0, GRT_GEN_POWER_LAW = 1, GRT_GEN_CUSTOM = 2 }; /* photon_generation.hpp:12-16 */
enum {                                              /* utils.hpp:135-141 PhotonPhase */
    GRT_PHASE_DISK = 0, GRT_PHASE_IMAGE_BACKWARD = 1, GRT_PHASE_IMAGE_FORWARD = 2,
    GRT_PHASE_IMAGE_ARRIVED = 3, GRT_PHASE_REJECTED = 4
};
/* GRT_PRECISION_FP64: the reference's FP64 arithmetic, bit for bit the
 * oracle's.  GRT_PRECISION_MIXED: FP64 geodesic with the reduced-algebra
 * right-hand side and relaxed FP64 rounding, FP32 fluid tetrad and
 * attenuation.  GRT_PRECISION_FP64_FAST_RHS: the mixed form's FP64 geodesic
 * (reduced-algebra right-hand side, relaxed rounding) with an FP64 tetrad
 * and attenuation -- the control that separates the algebra from the
 * precision in the mixed form's speed-up. */
enum { GRT_PRECISION_FP64 = 0, GRT_PRECISION_MIXED = 1, GRT_PRECISION_FP64_FAST_RHS = 2 };

/*
 * Everything Geodesic_cartesian_kerr_schild captures at construction
 * (integrate_cart_ks.hpp:35-61, :63-94) plus UnitConversions
 * (photon_generation.hpp:18-25).  Fill with grt_params_default() first.
 */
typedef struct {
    double a_bh, m_bh;                        /* spin, mass (utils.cpp:10-11) */
    double r_term_min, r_term_max;            /* termination radii (utils.hpp:164-167) */
    double atol, rtol;                        /* RK45 tolerances (utils.cpp:18-19) */
    double min_step_scale, max_step_scale, safety; /* utils.hpp:178-180 */
    int32_t integrator;                       /* GRT_RK4 / GRT_RK45 */
    int32_t mode;                             /* GRT_MODE_IMAGE / GRT_MODE_DISK */
    int32_t absorption_enabled, emission_enabled;   /* RadiativeTransferModel */
    double absorption_coefficient, emission_coefficient;
    int32_t scattering_enabled;               /* ScatteringModel (scattering.hpp:10-14) */
    double albedo;
    uint64_t seed;                            /* RNG key */
    uint32_t rng_stream;                      /* RNG stream selector, see grt_trace */
    int32_t generator;                        /* PhotonGenerationConfig */
    double superphoton_count_normalization, energy_per_cell_erg, power_law_slope;
    double nu_min_hz, nu_max_hz, custom_frequency_hz;
    double camera_distance, camera_theta, camera_phi, camera_width, camera_height;
    double length_cm_per_code, time_s_per_code, density_g_cm3_per_code;
    double temperature_k_per_code, four_velocity_cm_s_per_code, magnetic_gauss_per_code;
    int32_t precision;                        /* GRT_PRECISION_* */
} grt_params;

/*
 * The Photons structure of arrays, utils.hpp:40-54 (same fields, same types;
 * bool is one byte).  For grt_trace these are host pointers, for
 * grt_trace_device device pointers.  Every field must be non-NULL.
 */
typedef struct {
    uint64_t *id;
    double *frequency, *emission_frame_energy;
    double *x0, *x1, *x2, *x3;
    double *k0, *k1, *k2, *k3;                /* covariant p_mu */
    double *I, *Q, *U, *V, *dlambda;
    uint8_t *terminate, *phase;
    double *theta_disp, *phi_disp;
} grt_photons;

typedef struct {
    uint64_t functor_calls;      /* per-photon step calls on live photons */
    uint64_t accepted_steps;     /* accepted RK45 steps (or RK4 steps) */
    uint64_t rejected_attempts;  /* rejected RK45 attempts */
    uint64_t failed_steps;       /* steps that hit the 12-attempt limit */
    uint64_t fluid_samples;      /* sample_scattering_fluid calls */
    uint64_t fluid_hits;         /* of those, samples that found a cell with valid
                                    fluid and tetrad (the call returned true) */
    uint64_t global_steps;       /* host-loop iterations the reference would run */
    uint64_t terminated;         /* photons with terminate set at return */
    uint64_t launches;           /* tracing-kernel launches */
    double kernel_ms;            /* summed duration of the tracing-kernel launches */
} grt_stats;

/* Launch shape of the persistent tracing kernel (chosen by the autotuner). */
typedef struct {
    int32_t block;           /* work-items per workgroup (64..1024, multiple of 64) */
    int32_t blocks_per_cu;   /* resident workgroups requested per CU */
    int32_t sub_budget;      /* max steps per photon per launch before compaction */
} grt_tuning;

/* Fills *p with the reference's compiled-in defaults (utils.cpp:4-37,
 * simulation_options.hpp:13-32, photon_generation.hpp:40-51). */
void grt_params_default(grt_params *p);

typedef struct grt_ctx grt_ctx;

/* Replaces Geodesic_cartesian_kerr_schild's constructor
 * (integrate_cart_ks.hpp:63-94) together with the parameter validation of
 * toml_config.hpp:238-277 that concerns the traced path.  Creates a context
 * on the current HIP device. */
int grt_create(const grt_params *p, grt_ctx **out);
int grt_destroy(grt_ctx *ctx);

/* Replaces the functor's captured constants for later calls
 * (integrate_cart_ks.hpp:471-476 rebuilds the functor each step). */
int grt_set_params(grt_ctx *ctx, const grt_params *p);

/*
 * Replaces load_numpy_field_bundle (load_python_arrays.hpp:195-223) followed by
 * the ScatteringFluidGrid construction (integrate_cart_ks.hpp:453-459).
 * Host arrays in the .npy layout: r, theta, phi, rho, tgas are C-order
 * (nr, nth, nph) float64; vel and mag are (nr, nth, nph, 4).  The unit scales
 * of the current params are folded in here, so call again after a
 * grt_set_params that changes them.
 */
int grt_set_grid(grt_ctx *ctx, int32_t nr, int32_t nth, int32_t nph,
                 const double *r, const double *theta, const double *phi,
                 const double *rho, const double *tgas, const double *vel, const double *mag);

/*
 * Replaces integrate_geodesics (integrate_cart_ks.hpp:423-535) without its
 * output archives: advances every photon with the reference's global-step
 * semantics -- at most max_steps per-photon steps, the terminated fraction
 * re-counted on steps where step % output_interval == 0 or
 * step == max_steps - 1 (output_interval <= 0 counts as 1), stopping once it
 * reaches termination_percent.  Photons are updated in place.
 * Random draws (scattering, image-mode emission) come from a Philox4x32-10
 * stream keyed by (seed, photon id, rng_stream, draw number); the draw number
 * restarts at 0 on every call, so change rng_stream between calls that must
 * not reuse draws.
 * grt_trace takes host arrays (copied in and out); grt_trace_device takes
 * device arrays and runs on `stream` (a hipStream_t, NULL = default stream).
 * stats may be NULL.  At most 2^31 photons per call (GRT_E_INVALID beyond).
 */
int grt_trace(grt_ctx *ctx, const grt_photons *host, uint64_t n, int32_t max_steps,
              double termination_percent, int64_t output_interval, grt_stats *stats);
int grt_trace_device(grt_ctx *ctx, const grt_photons *dev, uint64_t n, int32_t max_steps,
                     double termination_percent, int64_t output_interval, void *stream,
                     grt_stats *stats);

/*
 * grt_trace_device without device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, semantics and results as grt_trace_device, but the work
 * lists, per-photon RNG state and counters live in one fixed device array of the library (about 13 MiB, held by every GPU the
 * library runs on until the process ends), so at most
 * GRT_NOALLOC_MAX_PHOTONS photons per call (GRT_E_RANGE beyond).  The call
 * makes no hipMalloc, hipFree or other allocating call (the grid, which
 * grt_set_grid placed, and the caller's photon arrays are its only other
 * device memory); it still synchronises `stream` after every launch.
 * Every no-allocation call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (from several threads or concurrently running streams)
 * race and must be serialised by the caller.
 */
#define GRT_NOALLOC_MAX_PHOTONS (1u << 20)
int grt_trace_device_noalloc(grt_ctx *ctx, const grt_photons *dev, uint64_t n, int32_t max_steps,
                             double termination_percent, int64_t output_interval, void *stream,
                             grt_stats *stats);

/*
 * Replaces initialize_photons_image (initialize_photons.hpp:82-170): fills n
 * image-mode rays on the camera plane of the current params with
 * id = id_offset + i, the null covariant momentum toward the hole,
 * dlambda = initial_step * camera_distance and phase IMAGE_BACKWARD.  Screen
 * positions (theta_disp, phi_disp) are taken from screen_u/screen_v, or, when
 * both are NULL, drawn uniformly over the plane from the photon's
 * initialisation stream (Philox, stream (rng_stream << 1) | 1) in place of the
 * reference's XorShift64 pool.  Host arrays for grt_init_image, device
 * arrays for grt_init_image_device.
 */
int grt_init_image(grt_ctx *ctx, const grt_photons *host, uint64_t n, uint64_t id_offset,
                   double initial_step, const double *screen_u, const double *screen_v);
int grt_init_image_device(grt_ctx *ctx, const grt_photons *dev, uint64_t n, uint64_t id_offset,
                          double initial_step, const double *screen_u, const double *screen_v,
                          void *stream);

/*
 * Replaces build_superphoton_cell_offsets + initialize_photons_disk
 * (initialize_photons.hpp:32-62, :173-327).  grt_disk_photon_count returns
 * the number of packets the grid emits, sum over cells of
 * max(0, ComputeSuperphotonsPerCell(rho_code, norm)); grt_init_disk* fill the
 * first n <= that many packets (packet i belongs to the cell whose prefix-sum
 * range holds i) with id = id_offset + i, a fluid-frame isotropic direction,
 * the configured spectrum and dlambda = initial_step.  Random draws come from
 * the packet's initialisation stream (Philox, stream (rng_stream << 1) | 1).
 * The grid must be set.
 */
int grt_disk_photon_count(grt_ctx *ctx, uint64_t *total);
int grt_init_disk(grt_ctx *ctx, const grt_photons *host, uint64_t n, uint64_t id_offset,
                  double initial_step);
int grt_init_disk_device(grt_ctx *ctx, const grt_photons *dev, uint64_t n, uint64_t id_offset,
                         double initial_step, void *stream);

/*
 * Replaces build_observation_products (observation.hpp:31-84): sums Stokes of
 * photons with phase == GRT_PHASE_IMAGE_ARRIVED into image[4][ny][nx]
 * (I, Q, U, V; pixel = y * nx + x, x from theta_disp, y from phi_disp, the
 * bin rule of observation.hpp:22-29) and spectrum_I[bins] on the log10
 * frequency grid spectrum_nu[bins].  Outputs are host arrays for
 * grt_image_products and device arrays for grt_image_products_device.
 */
int grt_image_products(grt_ctx *ctx, const grt_photons *host, uint64_t n, int64_t nx, int64_t ny,
                       int64_t spectrum_bins, double spectrum_min_hz, double spectrum_max_hz,
                       double *image, double *spectrum_nu, double *spectrum_I, uint64_t *arrived);
int grt_image_products_device(grt_ctx *ctx, const grt_photons *dev, uint64_t n, int64_t nx,
                              int64_t ny, int64_t spectrum_bins, double spectrum_min_hz,
                              double spectrum_max_hz, double *image, double *spectrum_nu,
                              double *spectrum_I, uint64_t *arrived, void *stream);

/*
 * The differentiable form (no counterpart in the reference, whose fluid
 * parameters enter only through the radiative-transfer update
 * apply_radiative_transfer_step, radiative_transfer.hpp:81-96).
 *
 * grt_trace_taped_device traces like grt_trace_device, always with the FP64
 * kernel, and keeps geodesic checkpoints: one per 32 functor calls of a
 * photon whose 32 calls touch Stokes I (a transfer update, the image-mode
 * emission event, or a zeroing), at most slots_per_photon per photon
 * (GRT_E_RANGE beyond that).  Scattering must be disabled.
 *
 * grt_adjoint_device then returns, for a loss L whose gradient with respect
 * to each photon's final Stokes I is dL_dI[n] (device array, the photons of
 * the taped trace in the same order), dL/drho and dL/dTgas per cell (device
 * arrays of nr*nth*nph doubles, code units, overwritten) and
 * grad_coeff[0..1] = dL/d absorption_coefficient, dL/d emission_coefficient
 * (device array, overwritten).  It replays each kept segment from its
 * checkpoint and sweeps the transfer recurrence backwards; the partial
 * derivatives of the increments come from the ...Gradient companions of the
 * extension points.  The coefficients may change between the taped trace
 * and the adjoint (the adjoint then differentiates at the new values along
 * the same geodesics); the geometry and the grid may not.  The image-mode
 * emission energy E / int(rho * norm) is piecewise constant in rho and is
 * treated as a constant.
 *
 * grt_set_fluid_scalars(_device) replaces the density and temperature of
 * every cell of the current grid (nr*nth*nph doubles each, host or device
 * arrays) without re-preparing the rest of the grid, for optimisation loops
 * and finite differences.
 */
int grt_trace_taped_device(grt_ctx *ctx, const grt_photons *dev, uint64_t n, int32_t max_steps,
                           double termination_percent, int64_t output_interval,
                           int32_t slots_per_photon, void *stream, grt_stats *stats);
int grt_adjoint_device(grt_ctx *ctx, const double *dL_dI, uint64_t n, double *grad_rho,
                       double *grad_tgas, double *grad_coeff, void *stream);
int grt_set_fluid_scalars(grt_ctx *ctx, const double *rho, const double *tgas);
int grt_set_fluid_scalars_device(grt_ctx *ctx, const double *rho, const double *tgas,
                                 void *stream);

/*
 * The file formats either side of the path.
 *
 * grt_load_config replaces toml_config::load_and_apply_toml_config
 * (toml_config.hpp:129-277): the same flat TOML subset, key set, defaults
 * (utils.cpp:4-37, simulation_options.hpp:13-32, photon_generation.hpp:40-50)
 * and error messages (available from grt_last_error(NULL)).  The physics,
 * camera, integration-tolerance and unit keys fill *params (starting from
 * grt_params_default); the run keys fill *run.  scattering.optical_depth is
 * not read, as in the reference.
 *
 * grt_read_npy reads an f8 C-order .npy file (cnpy.h:302-349; the messages of
 * load_python_arrays.hpp:82-112): with dst == NULL it only returns the shape.
 * grt_load_numpy_fields replaces discover_numpy_field_paths +
 * load_numpy_field_bundle (load_python_arrays.hpp:15-223): the seven files
 * r, theta, phi, rho, Tgas (nr, ntheta, nphi) and vel, mag (.., 4) of
 * `directory`, checked, then handed to grt_set_grid (with ctx == NULL the
 * bundle is only checked; messages then come from grt_last_error(NULL)).
 *
 * grt_write_image_products_npz writes <path> with the members of
 * write_image_products (write_output.hpp:107-142): image_I/Q/U/V (ny, nx),
 * spectrum_frequency_hz, spectrum_I (bins) and arrived_count (1, u8);
 * image is the [4][ny][nx] array of grt_image_products.
 * grt_write_photons_npz writes the selected photon arrays of host SoA `ph`,
 * every `stride`-th photon, as write_output_step does (write_output.hpp:
 * 26-60, 90-103); variables is an output.variables string (NULL: the
 * default selection).  Both write "<path>.tmp" and rename it over <path>
 * (write_output.hpp:66-86), as uncompressed, non-ZIP64 archives (cnpy.h).
 */
typedef struct {
    char name[256];               /* simulation.name */
    char numpy_dir[1024];         /* input.numpy_dir */
    char output_dir[1024];        /* output.directory */
    char output_variables[512];   /* output.variables, normalised comma list */
    char decomposition[128];      /* parallel.decomposition (one GPU: unused) */
    int32_t vacuum;               /* input.vacuum */
    int32_t max_steps;            /* integration.max_steps */
    uint64_t photon_count;        /* photons.count */
    double termination_percent;   /* integration.termination_percent */
    double initial_step;          /* integration.initial_step */
    uint64_t output_interval;     /* output.interval */
    uint64_t output_stride;       /* output.stride */
    uint64_t image_nx, image_ny;  /* output.image_nx / image_ny */
    uint64_t spectrum_bins;       /* output.spectrum_bins */
    double spectrum_min_hz, spectrum_max_hz;
} grt_run_config;

int grt_load_config(const char *path, grt_params *params, grt_run_config *run);

/*
 * The reference's Timers of integrate_geodesics (display.hpp:20-133,
 * integrate_cart_ks.hpp:466-521; one GPU, so no "Photon Exchange"):
 * accumulated wall seconds, plus the archives written.
 */
typedef struct {
    double geodesic_integration_s;  /* "Geodesic Integration": the tracing launches */
    double termination_count_s;     /* "Termination Count": recounting the terminated fraction */
    double output_s;                /* "Output": the step archives */
    uint64_t archives;              /* step archives written */
    int64_t last_archive_step;      /* global step of the last one (-1: none) */
} grt_timers;

/*
 * Replaces integrate_geodesics (integrate_cart_ks.hpp:423-535) with its
 * output: traces the n device photons as grt_trace_device does with
 * run->max_steps, run->termination_percent and run->output_interval, and on
 * every step s on which the terminated fraction is recounted
 * (s % output_interval == 0 or s == max_steps - 1) writes the archive
 * <output_dir>/<name>_step_<s>_rank0.npz: the output.variables selection of
 * every output.stride-th photon as the state after step s
 * (write_output.hpp:26-103).  In image mode it then writes
 * <output_dir>/<name>_image_products.npz (write_output.hpp:107-142) with the
 * run's image and spectrum settings.  stats and timers may be NULL.
 */
int grt_integrate_device(grt_ctx *ctx, const grt_photons *dev, uint64_t n,
                         const grt_run_config *run, void *stream, grt_stats *stats,
                         grt_timers *timers);
int grt_read_npy(const char *path, double *dst, uint64_t capacity, uint64_t shape[4],
                 int32_t *ndim);
int grt_load_numpy_fields(grt_ctx *ctx, const char *directory);
int grt_write_image_products_npz(const char *path, int64_t nx, int64_t ny, int64_t bins,
                                 const double *image, const double *spectrum_nu,
                                 const double *spectrum_I, uint64_t arrived);
int grt_write_photons_npz(const char *path, const grt_photons *ph, uint64_t n,
                          const char *variables, uint64_t stride);

/* Launch-shape control of the tracing kernel (no reference counterpart). */
int grt_get_tuning(grt_ctx *ctx, grt_tuning *t);
int grt_set_tuning(grt_ctx *ctx, const grt_tuning *t);

/* Message of the last failing call on ctx (or of the last failing
 * grt_create when ctx is NULL). */
const char *grt_last_error(const grt_ctx *ctx);
int grt_abi_version(void);

#ifdef __cplusplus
}
#endif
#endif
