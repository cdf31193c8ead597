// grt_api.cpp -- host side of libgrt_amd.so: the C ABI of include/grt.h.
//
// Owns the device copy of the fluid grid (cell records + search tables),
// the work lists of the persistent tracing kernel and the per-photon RNG
// counters, and reproduces integrate_geodesics' global-step loop
// (integrate_cart_ks.hpp:423-535) as a sequence of step-budgeted launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <functional>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "grt.h"
#include "grt_host_prep.hpp"
#include "grt_internal.hpp"

using grt::GridView;
using grt::PhotonPtrs;
using grt::StepParams;

// Device counters of the tracing driver, followed by the per-trace
// statistics in the same allocation, and their pinned host mirror.
using grt::kCounterWords;
using grt::kReadbackBytes;

struct grt_ctx {
    grt_params params{};
    StepParams sp{};
    grt_tuning tune{256, 2, 256};
    int device = 0;
    int num_cus = 0;
    size_t max_lds = 65536;   // dynamic LDS one workgroup may use
    std::string err;

    // grid
    bool has_grid = false;
    // The cell records fold in the spin (the spherical-to-Cartesian Jacobian)
    // and the unit scales of the params current at grt_set_grid; a later
    // grt_set_params that changes one of them leaves the records stale until
    // the grid is set again.
    bool grid_stale = false;
    int nr = 0, nth = 0, nph = 0;
    bool separable = false;
    double* d_tables = nullptr;   // separable search tables, then the three axes
    size_t table_doubles = 0;
    int key[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};   // bucket keys of the separable search
    double win[2] = {0.0, 0.0};   // radial window bounds of the separable search
    double* d_coords = nullptr;   // r3, th3, ph3 (general path only)
    double* d_cells = nullptr;    // kCellDoubles per cell
    double* d_rho_code = nullptr; // unscaled density, for disk-mode packet counts
    size_t ncell = 0;
    // disk-mode packet offsets (ncell + 1 counts and offsets, scan scratch)
    unsigned long long* d_disk = nullptr;
    void* d_disk_temp = nullptr;
    size_t disk_temp_bytes = 0;

    // scratch of the tracing driver
    uint32_t* d_lists = nullptr;  // two lists of list_cap entries
    size_t list_cap = 0;
    // per-photon scratch, one allocation: rng_cap u32 draw counters, then
    // rng_cap u8 deferred-frequency flags (cleared together)
    uint32_t* d_rng = nullptr;
    uint8_t* d_fpend = nullptr;
    size_t rng_cap = 0;
    // one allocation: kCounterWords u32 counters ([0] queue, [1] n_out,
    // [2] live count, [3] tape overflow, [4] adjoint queue), then the
    // statistics; read back together into the pinned h_readback
    uint32_t* d_counters = nullptr;
    unsigned long long* d_stats = nullptr;
    unsigned char* h_readback = nullptr;
    grt::TraceConsts* d_consts = nullptr;
    grt::TraceConsts h_consts{};   // host source of the copy to d_consts
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // result of looking up this device's fixed scratch (grt_noalloc.hip)
    hipError_t noalloc_status = hipErrorNotInitialized;

    // differentiable form: geodesic checkpoints of the last taped trace and
    // the replay records of the reverse sweep
    double* d_tape = nullptr;
    size_t tape_cap = 0;           // doubles
    uint64_t* d_thdr = nullptr;
    size_t thdr_cap = 0;
    uint64_t tape_n = 0;           // photons of the last taped trace (0: none, or
                                   // invalidated by a later change of grid,
                                   // fluid or geometry)
    int tape_slots = 0;
    int tape_integrator = 0;
    double* d_rec = nullptr;
    uint32_t rec_lanes = 0;

    // device staging for the host-pointer entry points
    void* d_stage = nullptr;
    size_t stage_cap = 0;
    double* d_img = nullptr;
    size_t img_cap = 0;
};

namespace {
std::string g_create_error;
}  // namespace

// The output.variables selection (grt_io.cpp).
int grt_output_selection(const char* spec, std::vector<std::string>* names);

// Message slots shared with grt_io.cpp.
void grt_set_global_error(const std::string& msg) { g_create_error = msg; }
int grt_set_error(grt_ctx* ctx, int code, const std::string& msg)
{
    if (ctx) ctx->err = msg;
    else g_create_error = msg;
    return code;
}

namespace {

int fail(grt_ctx* ctx, int code, const std::string& msg)
{
    if (ctx) ctx->err = msg;
    else g_create_error = msg;
    return code;
}

int hip_fail(grt_ctx* ctx, hipError_t e, const char* what)
{
    return fail(ctx, GRT_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define GRT_HIP(ctx, call)                                  \
    do {                                                    \
        const hipError_t e_ = (call);                       \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, #call); \
    } while (0)

bool finite(double v) { return std::isfinite(v); }

// Checks of toml_config.hpp:166-277 that concern the traced path.
int validate_params(grt_ctx* ctx, const grt_params* p)
{
    if (!p) return fail(ctx, GRT_E_INVALID, "params must not be NULL");
    if (!(p->m_bh >= 0.0) || !finite(p->m_bh))
        return fail(ctx, GRT_E_INVALID, "black_hole.mass must be finite and non-negative");
    if (!finite(p->a_bh) || std::fabs(p->a_bh) > p->m_bh)
        return fail(ctx, GRT_E_INVALID,
                    "black_hole.spin must satisfy |spin| <= mass (a naked singularity otherwise)");
    if (p->integrator != GRT_RK4 && p->integrator != GRT_RK45)
        return fail(ctx, GRT_E_INVALID, "integration.integrator must be \"rk4\" or \"rk45\"");
    if (p->mode != GRT_MODE_IMAGE && p->mode != GRT_MODE_DISK)
        return fail(ctx, GRT_E_INVALID, "simulation.mode must be 'image' or 'disk'");
    if (p->precision != GRT_PRECISION_FP64 && p->precision != GRT_PRECISION_MIXED &&
        p->precision != GRT_PRECISION_FP64_FAST_RHS)
        return fail(ctx, GRT_E_INVALID, "precision must be GRT_PRECISION_FP64, "
                                        "GRT_PRECISION_MIXED or GRT_PRECISION_FP64_FAST_RHS");
    if (!finite(p->atol) || !finite(p->rtol) || !finite(p->min_step_scale) ||
        !finite(p->max_step_scale) || !finite(p->safety))
        return fail(ctx, GRT_E_INVALID, "integration tolerances and step scales must be finite");
    if (!(p->camera_distance > 0.0) || !(p->camera_width > 0.0) || !(p->camera_height > 0.0))
        return fail(ctx, GRT_E_INVALID, "camera distance, width, and height must be positive");
    if (!finite(p->albedo) || p->albedo < 0.0 || p->albedo > 1.0)
        return fail(ctx, GRT_E_INVALID, "scattering.albedo must be finite and in [0, 1]");
    if (!finite(p->absorption_coefficient) || p->absorption_coefficient < 0.0)
        return fail(ctx, GRT_E_INVALID,
                    "radiative_transfer.absorption_coefficient must be finite and non-negative");
    if (!finite(p->emission_coefficient) || p->emission_coefficient < 0.0)
        return fail(ctx, GRT_E_INVALID,
                    "radiative_transfer.emission_coefficient must be finite and non-negative");
    // PhotonGenerationConfig::validate, photon_generation.hpp:52-74
    if (!(p->superphoton_count_normalization > 0.0) || !finite(p->superphoton_count_normalization))
        return fail(ctx, GRT_E_INVALID,
                    "photons.superphoton_count_normalization must be finite and positive");
    if (!(p->energy_per_cell_erg >= 0.0) || !finite(p->energy_per_cell_erg))
        return fail(ctx, GRT_E_INVALID, "photons.energy_per_cell_erg must be finite and non-negative");
    if (p->generator == GRT_GEN_POWER_LAW) {
        if (!finite(p->power_law_slope))
            return fail(ctx, GRT_E_INVALID, "photons.power_law_slope must be finite");
        if (!(p->nu_min_hz > 0.0) || !(p->nu_max_hz > p->nu_min_hz) || !finite(p->nu_min_hz) ||
            !finite(p->nu_max_hz))
            return fail(ctx, GRT_E_INVALID,
                        "power-law generation requires 0 < photons.nu_min_hz < photons.nu_max_hz");
    }
    if (p->generator == GRT_GEN_CUSTOM &&
        (!(p->custom_frequency_hz > 0.0) || !finite(p->custom_frequency_hz)))
        return fail(ctx, GRT_E_INVALID, "photons.custom_frequency_hz must be finite and positive");
    // UnitConversions::validate, photon_generation.hpp:26-37
    const double u[6] = {p->length_cm_per_code, p->time_s_per_code, p->density_g_cm3_per_code,
                         p->temperature_k_per_code, p->four_velocity_cm_s_per_code,
                         p->magnetic_gauss_per_code};
    for (double v : u)
        if (!(v > 0.0) || !finite(v))
            return fail(ctx, GRT_E_INVALID, "all [units] conversion factors must be finite and positive");
    return GRT_OK;
}

// Whether two parameter sets differ at most in the two transfer coefficients
// (field by field: the struct has padding).
bool same_except_coefficients(const grt_params& a, const grt_params& b)
{
#define GRT_SAME(f) (a.f == b.f)
    return GRT_SAME(a_bh) && GRT_SAME(m_bh) && GRT_SAME(r_term_min) && GRT_SAME(r_term_max) &&
           GRT_SAME(atol) && GRT_SAME(rtol) && GRT_SAME(min_step_scale) &&
           GRT_SAME(max_step_scale) && GRT_SAME(safety) && GRT_SAME(integrator) &&
           GRT_SAME(mode) && GRT_SAME(absorption_enabled) && GRT_SAME(emission_enabled) &&
           GRT_SAME(scattering_enabled) && GRT_SAME(albedo) && GRT_SAME(seed) &&
           GRT_SAME(rng_stream) && GRT_SAME(generator) &&
           GRT_SAME(superphoton_count_normalization) && GRT_SAME(energy_per_cell_erg) &&
           GRT_SAME(power_law_slope) && GRT_SAME(nu_min_hz) && GRT_SAME(nu_max_hz) &&
           GRT_SAME(custom_frequency_hz) && GRT_SAME(camera_distance) &&
           GRT_SAME(camera_theta) && GRT_SAME(camera_phi) && GRT_SAME(camera_width) &&
           GRT_SAME(camera_height) && GRT_SAME(length_cm_per_code) &&
           GRT_SAME(time_s_per_code) && GRT_SAME(density_g_cm3_per_code) &&
           GRT_SAME(temperature_k_per_code) && GRT_SAME(four_velocity_cm_s_per_code) &&
           GRT_SAME(magnetic_gauss_per_code) && GRT_SAME(precision);
#undef GRT_SAME
}

// Fluid-dependent entry points refuse a grid whose cell records no longer
// match the params.
int check_grid_fresh(grt_ctx* c)
{
    if (c->has_grid && c->grid_stale)
        return fail(c, GRT_E_STATE, "black_hole.spin or a [units] scale changed after "
                                    "grt_set_grid; call grt_set_grid again");
    return GRT_OK;
}

void free_grid(grt_ctx* c)
{
    if (c->d_tables) (void)hipFree(c->d_tables);
    if (c->d_coords) (void)hipFree(c->d_coords);
    if (c->d_cells) (void)hipFree(c->d_cells);
    if (c->d_rho_code) (void)hipFree(c->d_rho_code);
    if (c->d_disk) (void)hipFree(c->d_disk);
    if (c->d_disk_temp) (void)hipFree(c->d_disk_temp);
    c->d_tables = c->d_coords = c->d_cells = c->d_rho_code = nullptr;
    c->d_disk = nullptr;
    c->d_disk_temp = nullptr;
    c->disk_temp_bytes = 0;
    c->has_grid = false;
}

int ensure_scratch(grt_ctx* c, uint64_t n)
{
    if (n > c->list_cap) {
        if (c->d_lists) (void)hipFree(c->d_lists);
        c->d_lists = nullptr;
        c->list_cap = 0;
        const size_t cap = std::max<uint64_t>(n, 1024);
        GRT_HIP(c, hipMalloc(&c->d_lists, 2 * cap * sizeof(uint32_t)));
        c->list_cap = cap;
    }
    if (n > c->rng_cap) {
        if (c->d_rng) (void)hipFree(c->d_rng);
        c->d_rng = nullptr;
        c->d_fpend = nullptr;
        c->rng_cap = 0;
        const size_t cap = std::max<uint64_t>(n, 1024);
        GRT_HIP(c, hipMalloc(&c->d_rng, cap * (sizeof(uint32_t) + sizeof(uint8_t))));
        c->d_fpend = reinterpret_cast<uint8_t*>(c->d_rng + cap);
        c->rng_cap = cap;
    }
    return GRT_OK;
}

PhotonPtrs to_ptrs(const grt_photons& p, uint32_t* rng)
{
    PhotonPtrs d{};
    d.id = p.id;
    d.freq = p.frequency;
    d.eem = p.emission_frame_energy;
    d.x[0] = p.x0; d.x[1] = p.x1; d.x[2] = p.x2; d.x[3] = p.x3;
    d.k[0] = p.k0; d.k[1] = p.k1; d.k[2] = p.k2; d.k[3] = p.k3;
    d.st[0] = p.I; d.st[1] = p.Q; d.st[2] = p.U; d.st[3] = p.V;
    d.dl = p.dlambda;
    d.term = p.terminate;
    d.phase = p.phase;
    d.udisp = p.theta_disp;
    d.vdisp = p.phi_disp;
    d.rng_draw = rng;
    return d;
}

bool photons_complete(const grt_photons* p)
{
    const void* f[] = {p->id, p->frequency, p->emission_frame_energy, p->x0, p->x1, p->x2, p->x3,
                       p->k0, p->k1, p->k2, p->k3, p->I, p->Q, p->U, p->V, p->dlambda,
                       p->terminate, p->phase, p->theta_disp, p->phi_disp};
    for (const void* q : f)
        if (!q) return false;
    return true;
}

// Carve a device staging block into the 20 photon arrays.
int stage_photons(grt_ctx* c, uint64_t n, grt_photons* d)
{
    const size_t need = n * (18 * sizeof(double) + 2) + 64 * 20;
    if (need > c->stage_cap) {
        if (c->d_stage) (void)hipFree(c->d_stage);
        c->d_stage = nullptr;
        c->stage_cap = 0;
        GRT_HIP(c, hipMalloc(&c->d_stage, need));
        c->stage_cap = need;
    }
    char* base = static_cast<char*>(c->d_stage);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char* p = base + off;
        off += (bytes + 63) & ~size_t(63);
        return p;
    };
    d->id = reinterpret_cast<uint64_t*>(take(n * 8));
    double** dbl[] = {&d->frequency, &d->emission_frame_energy, &d->x0, &d->x1, &d->x2, &d->x3,
                      &d->k0, &d->k1, &d->k2, &d->k3, &d->I, &d->Q, &d->U, &d->V, &d->dlambda,
                      &d->theta_disp, &d->phi_disp};
    for (double** q : dbl) *q = reinterpret_cast<double*>(take(n * 8));
    d->terminate = reinterpret_cast<uint8_t*>(take(n));
    d->phase = reinterpret_cast<uint8_t*>(take(n));
    return GRT_OK;
}

int copy_photons(grt_ctx* c, const grt_photons* dst, const grt_photons* src, uint64_t n,
                 hipMemcpyKind kind)
{
    if (n == 0) return GRT_OK;
    GRT_HIP(c, hipMemcpy(dst->id, src->id, n * 8, kind));
    const std::pair<double*, double*> dd[] = {
        {dst->frequency, src->frequency}, {dst->emission_frame_energy, src->emission_frame_energy},
        {dst->x0, src->x0}, {dst->x1, src->x1}, {dst->x2, src->x2}, {dst->x3, src->x3},
        {dst->k0, src->k0}, {dst->k1, src->k1}, {dst->k2, src->k2}, {dst->k3, src->k3},
        {dst->I, src->I}, {dst->Q, src->Q}, {dst->U, src->U}, {dst->V, src->V},
        {dst->dlambda, src->dlambda}, {dst->theta_disp, src->theta_disp},
        {dst->phi_disp, src->phi_disp}};
    for (const auto& q : dd) GRT_HIP(c, hipMemcpy(q.first, q.second, n * 8, kind));
    GRT_HIP(c, hipMemcpy(dst->terminate, src->terminate, n, kind));
    GRT_HIP(c, hipMemcpy(dst->phase, src->phase, n, kind));
    return GRT_OK;
}

}  // namespace

extern "C" {

void grt_params_default(grt_params* p)
{
    if (!p) return;
    std::memset(p, 0, sizeof *p);
    p->m_bh = 1.0;                       // utils.cpp:10
    p->a_bh = 1.0;                       // utils.cpp:11
    p->r_term_min = 1.0;                 // R_HORIZON for a = M = 1 (utils.cpp:12,27)
    p->camera_distance = 10000.0;        // utils.cpp:21
    p->r_term_max = 1.5 * 10000.0;       // utils.cpp:28
    p->atol = 1e-8;                      // utils.cpp:30
    p->rtol = 1e-8;                      // utils.cpp:31
    p->min_step_scale = 0.01;            // utils.hpp:179
    p->max_step_scale = 10.0;            // utils.hpp:180
    p->safety = 0.9;                     // utils.hpp:178
    p->integrator = GRT_RK45;            // simulation_options.hpp:26
    p->mode = GRT_MODE_IMAGE;            // simulation_options.hpp:14
    p->albedo = 1.0;                     // simulation_options.hpp:20
    p->seed = 12345u;                    // simulation_options.hpp:21
    p->generator = GRT_GEN_BLACKBODY;    // photon_generation.hpp:41
    p->superphoton_count_normalization = 1.0;
    p->energy_per_cell_erg = 1.0;
    p->power_law_slope = 2.0;
    p->nu_min_hz = 1.0e9;
    p->nu_max_hz = 1.0e20;
    p->custom_frequency_hz = 2.30e11;
    const double pi = 3.14159265358979323846;
    p->camera_theta = pi / 4.0;          // utils.cpp:15
    p->camera_phi = pi / 4.0;            // utils.cpp:16
    p->camera_width = 20.0;              // utils.cpp:17
    p->camera_height = 20.0;             // utils.cpp:18
    p->length_cm_per_code = p->time_s_per_code = p->density_g_cm3_per_code = 1.0;
    p->temperature_k_per_code = p->four_velocity_cm_s_per_code = p->magnetic_gauss_per_code = 1.0;
    p->precision = GRT_PRECISION_FP64;
}

int grt_abi_version(void) { return GRT_ABI_VERSION; }

const char* grt_last_error(const grt_ctx* ctx)
{
    return ctx ? ctx->err.c_str() : g_create_error.c_str();
}

int grt_create(const grt_params* p, grt_ctx** out)
{
    if (!out) return fail(nullptr, GRT_E_INVALID, "out must not be NULL");
    *out = nullptr;
    const int v = validate_params(nullptr, p);
    if (v != GRT_OK) return v;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(nullptr, GRT_E_DEVICE, std::string("no HIP device: ") + hipGetErrorString(e));
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, dev);
    if (e != hipSuccess) return fail(nullptr, GRT_E_DEVICE, hipGetErrorString(e));
    grt_ctx* c = new grt_ctx();
    c->device = dev;
    c->num_cus = prop.multiProcessorCount;
    c->max_lds = prop.sharedMemPerBlock;
    c->params = *p;
    c->sp = grt::make_step_params(*p);
    if (hipMalloc(&c->d_counters, kReadbackBytes) != hipSuccess ||
        hipHostMalloc(&c->h_readback, kReadbackBytes) != hipSuccess ||
        hipMalloc(&c->d_consts, sizeof(grt::TraceConsts)) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        grt_destroy(c);
        return fail(nullptr, GRT_E_NOMEM, "device allocation failed in grt_create");
    }
    c->d_stats = reinterpret_cast<unsigned long long*>(c->d_counters + kCounterWords);
    // this device's copy of the fixed scratch of grt_trace_device_noalloc;
    // a failure here is reported by that entry point only
    c->noalloc_status = grt::noalloc_scratch_init(dev);
    *out = c;
    return GRT_OK;
}

int grt_destroy(grt_ctx* c)
{
    if (!c) return GRT_OK;
    free_grid(c);
    void* bufs[] = {c->d_lists, c->d_rng, c->d_counters, c->d_stage, c->d_img,
                    c->d_consts, c->d_tape, c->d_thdr, c->d_rec};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (c->h_readback) (void)hipHostFree(c->h_readback);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    delete c;
    return GRT_OK;
}

int grt_set_params(grt_ctx* c, const grt_params* p)
{
    if (!c) return fail(nullptr, GRT_E_INVALID, "ctx must not be NULL");
    const int v = validate_params(c, p);
    if (v != GRT_OK) return v;
    const grt_params& o = c->params;
    if (c->has_grid &&
        (p->a_bh != o.a_bh || p->density_g_cm3_per_code != o.density_g_cm3_per_code ||
         p->temperature_k_per_code != o.temperature_k_per_code ||
         p->four_velocity_cm_s_per_code != o.four_velocity_cm_s_per_code ||
         p->magnetic_gauss_per_code != o.magnetic_gauss_per_code))
        c->grid_stale = true;
    // The adjoint may differentiate at new transfer coefficients along the
    // taped geodesics; any other change makes the tape another problem's.
    if (!same_except_coefficients(*p, o)) c->tape_n = 0;
    c->params = *p;
    c->sp = grt::make_step_params(*p);
    return GRT_OK;
}

int grt_get_tuning(grt_ctx* c, grt_tuning* t)
{
    if (!c || !t) return fail(c, GRT_E_INVALID, "ctx and tuning must not be NULL");
    *t = c->tune;
    return GRT_OK;
}

int grt_set_tuning(grt_ctx* c, const grt_tuning* t)
{
    if (!c || !t) return fail(c, GRT_E_INVALID, "ctx and tuning must not be NULL");
    if (t->block < 64 || t->block > 256 || t->block % 64 != 0)
        return fail(c, GRT_E_INVALID, "tuning.block must be 64, 128, 192 or 256");
    if (t->blocks_per_cu < 1 || t->blocks_per_cu > 32)
        return fail(c, GRT_E_INVALID, "tuning.blocks_per_cu must be in [1, 32]");
    if (t->sub_budget < 1)
        return fail(c, GRT_E_INVALID, "tuning.sub_budget must be positive");
    c->tune = *t;
    return GRT_OK;
}

int grt_set_grid(grt_ctx* c, int32_t nr, int32_t nth, int32_t nph, const double* r,
                 const double* theta, const double* phi, const double* rho, const double* tgas,
                 const double* vel, const double* mag)
{
    if (!c) return fail(nullptr, GRT_E_INVALID, "ctx must not be NULL");
    if (nr <= 0 || nth <= 0 || nph <= 0)
        return fail(c, GRT_E_INVALID, "grid extents must be positive");
    if (!r || !theta || !phi || !rho || !tgas || !vel || !mag)
        return fail(c, GRT_E_INVALID, "every grid field pointer must be non-NULL");
    const size_t G = (size_t)nr * nth * nph;
    if (G > (size_t)1 << 31) return fail(c, GRT_E_INVALID, "grid has more than 2^31 cells");
    free_grid(c);
    c->grid_stale = false;
    c->tape_n = 0;
    grt::HostGrid hg;
    grt::prepare_grid(c->params, nr, nth, nph, r, theta, phi, rho, tgas, vel, mag, hg);
    c->nr = nr;
    c->nth = nth;
    c->nph = nph;
    c->ncell = G;
    c->separable = hg.separable;
    for (int q = 0; q < 9; ++q) c->key[q] = hg.key[q];
    hipError_t e;
    c->win[0] = c->win[1] = 0.0;
    if (hg.separable) {
        c->win[0] = hg.tables[2 * (size_t)nr];
        c->win[1] = hg.tables[4 * (size_t)nr - 1];
        std::vector<double> t = hg.tables;
        t.insert(t.end(), hg.axes.begin(), hg.axes.end());
        c->table_doubles = hg.tables.size();
        e = hipMalloc(&c->d_tables, t.size() * sizeof(double));
        if (e != hipSuccess) return hip_fail(c, e, "hipMalloc(tables)");
        e = hipMemcpy(c->d_tables, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_fail(c, e, "hipMemcpy(tables)");
    } else {
        e = hipMalloc(&c->d_coords, 3 * G * sizeof(double));
        if (e != hipSuccess) return hip_fail(c, e, "hipMalloc(coords)");
        const double* src[3] = {r, theta, phi};
        for (int q = 0; q < 3; ++q) {
            e = hipMemcpy(c->d_coords + q * G, src[q], G * sizeof(double), hipMemcpyHostToDevice);
            if (e != hipSuccess) return hip_fail(c, e, "hipMemcpy(coords)");
        }
    }
    e = hipMalloc(&c->d_rho_code, G * sizeof(double));
    if (e != hipSuccess) return hip_fail(c, e, "hipMalloc(rho)");
    e = hipMemcpy(c->d_rho_code, rho, G * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(c, e, "hipMemcpy(rho)");
    e = hipMalloc(&c->d_cells, hg.cells.size() * sizeof(double));
    if (e != hipSuccess) return hip_fail(c, e, "hipMalloc(cells)");
    e = hipMemcpy(c->d_cells, hg.cells.data(), hg.cells.size() * sizeof(double),
                  hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(c, e, "hipMemcpy(cells)");
    c->has_grid = true;
    return GRT_OK;
}

namespace {

GridView make_grid_view(const grt_ctx* c)
{
    return grt::grid_view(c->nr, c->nth, c->nph, c->has_grid, c->separable, c->d_tables,
                          c->d_tables ? c->d_tables + c->table_doubles : nullptr, c->d_coords,
                          c->d_cells, c->key, c->win);
}

}  // namespace

namespace {

// integrate_geodesics' loop on the device; tape_slots > 0 also writes the
// geodesic checkpoints of the differentiable form (FP64 kernel).
int trace_impl(grt_ctx* c, const grt_photons* dev, uint64_t n, int32_t max_steps,
               double termination_percent, int64_t output_interval, void* stream_v,
               grt_stats* stats, int tape_slots, bool fixed_scratch = false,
               const std::function<int(int64_t)>* on_check = nullptr, grt_timers* timers = nullptr)
{
    if (!c) return fail(nullptr, GRT_E_INVALID, "ctx must not be NULL");
    if (!dev || (n > 0 && !photons_complete(dev))) return fail(c, GRT_E_INVALID, "every photon array must be non-NULL");
    if (n > 0x80000000ull) return fail(c, GRT_E_INVALID, "more than 2^31 photons in one call");
    if (max_steps <= 0) return fail(c, GRT_E_INVALID, "integration.max_steps must be positive");
    if (!(termination_percent > 0.0) || termination_percent > 1.0)
        return fail(c, GRT_E_INVALID, "integration.termination_percent must be in (0, 1]");
    int rc = check_grid_fresh(c);
    if (rc != GRT_OK) return rc;
    hipStream_t stream = static_cast<hipStream_t>(stream_v);
    GRT_HIP(c, hipSetDevice(c->device));
    // Scratch: the context's own buffers (grown on demand), or the fixed
    // array of the no-allocation entry point.
    grt::TraceScratch S{};
    if (fixed_scratch) {
        if (c->noalloc_status != hipSuccess)
            return hip_fail(c, c->noalloc_status, "fixed scratch of grt_trace_device_noalloc");
        GRT_HIP(c, grt::noalloc_scratch(c->device, &S));
    } else {
        rc = ensure_scratch(c, n);
        if (rc != GRT_OK) return rc;
        S.lists = c->d_lists;
        S.list_cap = c->list_cap;
        S.rng = c->d_rng;
        S.fpend = c->d_fpend;
        S.counters = c->d_counters;
        S.fill_kernel = false;
    }
    unsigned long long* d_stats = reinterpret_cast<unsigned long long*>(S.counters + kCounterWords);
    auto clear = [&](void* p, size_t bytes) {
        return S.fill_kernel ? grt::launch_zero(p, bytes, stream) : hipMemsetAsync(p, 0, bytes, stream);
    };

    // Fluid-dependent features are active only with a grid, integrate_cart_ks.hpp:436-452.
    StepParams P = c->sp;
    const bool fluid = c->has_grid;
    P.absorb = P.absorb && fluid;
    P.emit = P.emit && fluid;
    P.scatter = P.scatter && fluid;

    grt_stats st{};
    // Counters and statistics; build_live also zeroes the draw counters and
    // deferred-frequency flags.
    GRT_HIP(c, clear(S.counters, kReadbackBytes));
    uint32_t* list[2] = {S.lists, S.lists + S.list_cap};
    GRT_HIP(c, grt::launch_build_live(dev->terminate, n, list[0], S.counters + 2, S.rng, S.fpend,
                                      stream));
    // The first launch reads the live count on the device; until its
    // readback, n bounds it.
    uint32_t live = (uint32_t)n;

    c->h_consts.ph = to_ptrs(*dev, S.rng);
    c->h_consts.ph.fpend = S.fpend;
    c->h_consts.P = P;
    c->h_consts.grid = make_grid_view(c);
    grt::TraceLaunch L{};
    L.cst = c->h_consts;   // carried in the kernel arguments
    L.integrator = P.integrator;
    L.grid = make_grid_view(c);
    L.queue = S.counters + 0;
    L.n_out = S.counters + 1;
    L.stats = d_stats;
    const int lds_doubles = (c->has_grid && c->separable) ? (int)c->table_doubles : 0;
    L.lds_doubles = lds_doubles;
    L.mixed = c->params.precision;   // 0 FP64, 1 mixed, 2 FP64 with the fast right-hand side
    if (tape_slots > 0) {
        const size_t need = (size_t)n * (size_t)tape_slots * grt::kCkptDoubles;
        if (need > c->tape_cap) {
            if (c->d_tape) (void)hipFree(c->d_tape);
            c->d_tape = nullptr;
            c->tape_cap = 0;
            GRT_HIP(c, hipMalloc(&c->d_tape, std::max<size_t>(need, 1) * sizeof(double)));
            c->tape_cap = need;
        }
        if (n > c->thdr_cap) {
            if (c->d_thdr) (void)hipFree(c->d_thdr);
            c->d_thdr = nullptr;
            c->thdr_cap = 0;
            GRT_HIP(c, hipMalloc(&c->d_thdr, std::max<uint64_t>(n, 1) * sizeof(uint64_t)));
            c->thdr_cap = n;
        }
        GRT_HIP(c, hipMemsetAsync(c->d_thdr, 0, std::max<uint64_t>(n, 1) * sizeof(uint64_t), stream));
        L.tape = c->d_tape;
        L.tape_hdr = c->d_thdr;
        L.tape_slots = tape_slots;
        L.tape_overflow = S.counters + 3;
        L.mixed = 0;
        c->tape_n = 0;
    }
    const int block = c->tune.block;
    const size_t lds_bytes = grt::trace_lds_bytes(lds_doubles, block);
    if (lds_bytes > c->max_lds)
        return fail(c, GRT_E_INVALID, "the grid's search tables and the photon stages of a " +
                                          std::to_string(block) + "-thread workgroup need " +
                                          std::to_string(lds_bytes) + " B of LDS, more than the " +
                                          std::to_string(c->max_lds) +
                                          " B a workgroup may use; lower tuning.block");
    const int persistent_blocks = std::max(1, c->num_cus * c->tune.blocks_per_cu);

    // Global-step schedule, integrate_cart_ks.hpp:468-527.  When the stop
    // test needs every photon terminated, it cannot end the loop while any
    // photon would still change, so one epoch of max_steps is equivalent --
    // unless something happens at every check step (the output archives).
    const int64_t interval = output_interval <= 0 ? 1 : output_interval;
    const bool exact_schedule =
        on_check != nullptr || !((double)n * termination_percent > (double)n - 1.0);
    using Clock = std::chrono::steady_clock;
    auto seconds_since = [](Clock::time_point t0) {
        return std::chrono::duration<double>(Clock::now() - t0).count();
    };
    double fraction = 0.0;
    int64_t g = 0;
    int cur = 0;
    float kernel_ms = 0.0f;
    uint64_t launches = 0;
    uint64_t global_steps = 0;
    while (g < max_steps) {
        int64_t check;
        if (!exact_schedule) {
            check = max_steps - 1;
        } else {
            check = ((g + interval - 1) / interval) * interval;
            if (check > max_steps - 1) check = max_steps - 1;
        }
        int64_t left = check - g + 1;
        int64_t step0 = g;
        const Clock::time_point t_epoch = Clock::now();
        while (left > 0 && live > 0) {
            const int budget = (int)std::min<int64_t>(left, c->tune.sub_budget);
            // With every photon live the list is a permutation of all
            // photons, and the order of processing does not change results:
            // the first launch then walks the photons directly (the kernel
            // decides, from the count build_live left on the device).
            L.list_in = list[cur];
            L.list_out = list[cur ^ 1];
            L.n_in = live;
            L.n_in_dev = launches == 0 ? S.counters + 2 : nullptr;
            L.n_total = (uint32_t)n;
            L.budget = budget;
            L.step0 = (int)step0;
            // (the first launch finds them zero from the clear above)
            if (launches > 0) GRT_HIP(c, clear(S.counters, 2 * sizeof(uint32_t)));
            const int need_blocks = (int)((live + block - 1) / block);
            const int blocks = std::max(1, std::min(persistent_blocks, need_blocks));
            L.grid_lanes = (uint32_t)blocks * (uint32_t)block;
            GRT_HIP(c, hipEventRecord(c->ev0, stream));
            GRT_HIP(c, grt::launch_trace(L, blocks, block, lds_bytes, stream));
            GRT_HIP(c, hipEventRecord(c->ev1, stream));
            // Thermal frequencies whose draw the tracing kernel deferred.  A
            // photon still in flight continues with its drawn frequency,
            // which the kernel would have drawn the same way later; queued
            // before the synchronisation, the last launch needs no second
            // round trip.  Counters and statistics come back in one copy.
            GRT_HIP(c, grt::launch_resolve_deferred(c->h_consts.ph, n, P.seed, P.stream << 1,
                                                    P.a, P.m, stream));
            GRT_HIP(c, hipMemcpyAsync(c->h_readback, S.counters, kReadbackBytes,
                                      hipMemcpyDeviceToHost, stream));
            GRT_HIP(c, hipStreamSynchronize(stream));
            live = reinterpret_cast<const uint32_t*>(c->h_readback)[1];
            float ms = 0.0f;
            GRT_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
            kernel_ms += ms;
            ++launches;
            cur ^= 1;
            left -= budget;
            step0 += budget;
        }
        if (timers) timers->geodesic_integration_s += seconds_since(t_epoch);
        // The terminated count is the kernel's own live count (read back
        // with the statistics after each launch), so the recount is free.
        const Clock::time_point t_count = Clock::now();
        global_steps = (uint64_t)check + 1;
        fraction = (double)(n - live) / (double)(n > 0 ? n : 1);
        g = check + 1;
        if (timers) timers->termination_count_s += seconds_since(t_count);
        if (on_check) {
            const Clock::time_point t_out = Clock::now();
            const int orc = (*on_check)(check);
            if (timers) timers->output_s += seconds_since(t_out);
            if (orc != GRT_OK) return orc;
        }
        if (fraction >= termination_percent) break;
        if (!exact_schedule && live == 0) break;
    }
    if (tape_slots > 0) {
        GRT_HIP(c, grt::launch_tape_finalize(c->d_thdr, n, tape_slots, S.counters + 3, stream));
        uint32_t overflow = 0;
        GRT_HIP(c, hipMemcpyAsync(&overflow, S.counters + 3, sizeof overflow,
                                  hipMemcpyDeviceToHost, stream));
        GRT_HIP(c, hipStreamSynchronize(stream));
        if (overflow)
            return fail(c, GRT_E_RANGE, "a photon needs more than slots_per_photon geodesic "
                                        "checkpoints; raise slots_per_photon");
        c->tape_n = n;
        c->tape_slots = tape_slots;
        c->tape_integrator = P.integrator;
    }
    if (launches == 0) {
        GRT_HIP(c, hipMemcpyAsync(c->h_readback, S.counters, kReadbackBytes,
                                  hipMemcpyDeviceToHost, stream));
        GRT_HIP(c, hipStreamSynchronize(stream));
    }
    const unsigned long long* hs =
        reinterpret_cast<const unsigned long long*>(c->h_readback + kCounterWords * sizeof(uint32_t));
    if (!exact_schedule && live == 0) {
        // The reference loop ends at the first output step at or after the
        // step on which the last photon terminated.
        const int64_t last = (n > 0 && hs[grt::kStatCalls] > 0) ? (int64_t)hs[grt::kStatLastTermStep] : 0;
        int64_t chk = ((last + interval - 1) / interval) * interval;
        if (chk > max_steps - 1) chk = max_steps - 1;
        global_steps = (uint64_t)chk + 1;
    }
    st.functor_calls = hs[grt::kStatCalls];
    st.accepted_steps = hs[grt::kStatAccepted];
    st.rejected_attempts = hs[grt::kStatRejected];
    st.failed_steps = hs[grt::kStatFailed];
    st.fluid_samples = hs[grt::kStatSamples];
    st.fluid_hits = hs[grt::kStatHits];
    st.global_steps = global_steps;
    st.terminated = n - live;
    st.launches = launches;
    st.kernel_ms = kernel_ms;
    if (stats) *stats = st;
    return GRT_OK;
}

}  // namespace

int grt_trace_device(grt_ctx* c, const grt_photons* dev, uint64_t n, int32_t max_steps,
                     double termination_percent, int64_t output_interval, void* stream_v,
                     grt_stats* stats)
{
    return trace_impl(c, dev, n, max_steps, termination_percent, output_interval, stream_v, stats,
                      0);
}

int grt_trace_device_noalloc(grt_ctx* c, const grt_photons* dev, uint64_t n, int32_t max_steps,
                             double termination_percent, int64_t output_interval, void* stream_v,
                             grt_stats* stats)
{
    if (!c) return fail(nullptr, GRT_E_INVALID, "ctx must not be NULL");
    if (n > GRT_NOALLOC_MAX_PHOTONS)
        return fail(c, GRT_E_RANGE, "grt_trace_device_noalloc takes at most " +
                                        std::to_string(GRT_NOALLOC_MAX_PHOTONS) +
                                        " photons per call (its fixed scratch); got " +
                                        std::to_string(n));
    return trace_impl(c, dev, n, max_steps, termination_percent, output_interval, stream_v, stats,
                      0, true);
}

int grt_trace_taped_device(grt_ctx* c, const grt_photons* dev, uint64_t n, int32_t max_steps,
                           double termination_percent, int64_t output_interval,
                           int32_t slots_per_photon, void* stream_v, grt_stats* stats)
{
    if (!c) return fail(nullptr, GRT_E_INVALID, "ctx must not be NULL");
    if (slots_per_photon <= 0) return fail(c, GRT_E_INVALID, "slots_per_photon must be positive");
    if (c->params.scattering_enabled)
        return fail(c, GRT_E_INVALID, "the differentiable form requires scattering.enabled = false");
    return trace_impl(c, dev, n, max_steps, termination_percent, output_interval, stream_v, stats,
                      slots_per_photon);
}

int grt_adjoint_device(grt_ctx* c, const double* dL_dI, uint64_t n, double* grad_rho,
                       double* grad_tgas, double* grad_coeff, void* stream_v)
{
    if (!c) return fail(nullptr, GRT_E_INVALID, "ctx must not be NULL");
    if (!dL_dI || !grad_rho || !grad_tgas || !grad_coeff)
        return fail(c, GRT_E_INVALID, "seed and gradient arrays must be non-NULL");
    if (!c->has_grid) return fail(c, GRT_E_NOGRID, "the adjoint needs a fluid grid (grt_set_grid)");
    if (check_grid_fresh(c) != GRT_OK) return GRT_E_STATE;
    if (c->tape_n == 0 || c->tape_n != n)
        return fail(c, GRT_E_STATE, "no taped trace of this many photons: call grt_trace_taped first");
    if (c->params.scattering_enabled)
        return fail(c, GRT_E_INVALID, "the differentiable form requires scattering.enabled = false");
    if (c->params.integrator != c->tape_integrator)
        return fail(c, GRT_E_STATE, "the integrator changed since the taped trace");
    hipStream_t stream = static_cast<hipStream_t>(stream_v);
    GRT_HIP(c, hipSetDevice(c->device));
    const int block = c->tune.block;
    const int blocks = std::max(1, c->num_cus * c->tune.blocks_per_cu);
    const uint32_t lanes = (uint32_t)blocks * (uint32_t)block;
    if (lanes > c->rec_lanes) {
        if (c->d_rec) (void)hipFree(c->d_rec);
        c->d_rec = nullptr;
        c->rec_lanes = 0;
        GRT_HIP(c, hipMalloc(&c->d_rec, (size_t)lanes * grt::kRecEntries * grt::kRecDoubles *
                                            sizeof(double)));
        c->rec_lanes = lanes;
    }
    // The checkpoint replay needs the launch constants of the taped trace,
    // which the last trace left in d_consts; refresh P from the current
    // parameters (the coefficients may have changed, the geometry may not).
    StepParams P = c->sp;
    P.absorb = P.absorb && c->has_grid;
    P.emit = P.emit && c->has_grid;
    P.scatter = 0;
    c->h_consts.P = P;
    GRT_HIP(c, hipMemcpyAsync(c->d_consts, &c->h_consts, sizeof c->h_consts,
                              hipMemcpyHostToDevice, stream));
    GRT_HIP(c, hipMemsetAsync(grad_rho, 0, c->ncell * sizeof(double), stream));
    GRT_HIP(c, hipMemsetAsync(grad_tgas, 0, c->ncell * sizeof(double), stream));
    GRT_HIP(c, hipMemsetAsync(grad_coeff, 0, 2 * sizeof(double), stream));
    GRT_HIP(c, hipMemsetAsync(c->d_counters + 4, 0, sizeof(uint32_t), stream));
    grt::AdjointLaunch A{};
    A.consts = c->d_consts;
    A.grid = make_grid_view(c);
    A.integrator = P.integrator;
    A.tape = c->d_tape;
    A.tape_hdr = c->d_thdr;
    A.tape_slots = c->tape_slots;
    A.seed = dL_dI;
    A.n = (uint32_t)n;
    A.queue = c->d_counters + 4;
    A.rec = c->d_rec;
    A.rec_lanes = c->rec_lanes;
    A.grad_rho = grad_rho;
    A.grad_T = grad_tgas;
    A.grad_coeff = grad_coeff;
    A.density_scale = c->params.density_g_cm3_per_code;
    A.temperature_scale = c->params.temperature_k_per_code;
    const int lds_doubles = c->separable ? (int)c->table_doubles : 0;
    A.lds_doubles = lds_doubles;
    GRT_HIP(c, grt::launch_adjoint(A, blocks, block, (size_t)lds_doubles * sizeof(double), stream));
    GRT_HIP(c, hipStreamSynchronize(stream));
    return GRT_OK;
}

int grt_set_fluid_scalars_device(grt_ctx* c, const double* rho, const double* tgas,
                                 void* stream_v)
{
    if (!c) return fail(nullptr, GRT_E_INVALID, "ctx must not be NULL");
    if (!rho || !tgas) return fail(c, GRT_E_INVALID, "rho and Tgas must be non-NULL");
    if (!c->has_grid) return fail(c, GRT_E_NOGRID, "no grid: call grt_set_grid first");
    if (check_grid_fresh(c) != GRT_OK) return GRT_E_STATE;
    c->tape_n = 0;   // the taped trace saw the old fluid
    GRT_HIP(c, hipSetDevice(c->device));
    GRT_HIP(c, grt::launch_set_fluid_scalars(c->d_cells, c->d_rho_code, rho, tgas, c->ncell,
                                             c->params.density_g_cm3_per_code,
                                             c->params.temperature_k_per_code,
                                             static_cast<hipStream_t>(stream_v)));
    return GRT_OK;
}

int grt_set_fluid_scalars(grt_ctx* c, const double* rho, const double* tgas)
{
    if (!c) return fail(nullptr, GRT_E_INVALID, "ctx must not be NULL");
    if (!rho || !tgas) return fail(c, GRT_E_INVALID, "rho and Tgas must be non-NULL");
    if (!c->has_grid) return fail(c, GRT_E_NOGRID, "no grid: call grt_set_grid first");
    GRT_HIP(c, hipSetDevice(c->device));
    double* d = nullptr;
    GRT_HIP(c, hipMalloc(&d, 2 * c->ncell * sizeof(double)));
    hipError_t e = hipMemcpy(d, rho, c->ncell * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(d + c->ncell, tgas, c->ncell * sizeof(double), hipMemcpyHostToDevice);
    int rc = GRT_OK;
    if (e != hipSuccess) rc = hip_fail(c, e, "hipMemcpy(fluid scalars)");
    if (rc == GRT_OK) rc = grt_set_fluid_scalars_device(c, d, d + c->ncell, nullptr);
    if (rc == GRT_OK) {
        e = hipDeviceSynchronize();
        if (e != hipSuccess) rc = hip_fail(c, e, "set_fluid_scalars");
    }
    (void)hipFree(d);
    return rc;
}

int grt_integrate_device(grt_ctx* c, const grt_photons* dev, uint64_t n,
                         const grt_run_config* run, void* stream_v, grt_stats* stats,
                         grt_timers* timers)
{
    if (!c) return fail(nullptr, GRT_E_INVALID, "ctx must not be NULL");
    if (!run) return fail(c, GRT_E_INVALID, "run must not be NULL");
    if (!dev || (n > 0 && !photons_complete(dev)))
        return fail(c, GRT_E_INVALID, "every photon array must be non-NULL");
    if (run->output_interval == 0) return fail(c, GRT_E_INVALID, "output.interval must be positive");
    if (run->output_stride == 0) return fail(c, GRT_E_INVALID, "output.stride must be positive");
    if (run->max_steps <= 0) return fail(c, GRT_E_INVALID, "integration.max_steps must be positive");
    std::vector<std::string> sel;
    if (grt_output_selection(run->output_variables[0] ? run->output_variables : nullptr, &sel) !=
        GRT_OK)
        return fail(c, GRT_E_INVALID, grt_last_error(nullptr));
    const bool image = c->params.mode == GRT_MODE_IMAGE;
    if (image && (run->image_nx == 0 || run->image_ny == 0 || run->spectrum_bins == 0 ||
                  !(run->spectrum_min_hz > 0.0) || !(run->spectrum_max_hz > run->spectrum_min_hz)))
        return fail(c, GRT_E_INVALID, "image mode needs output.image_nx, image_ny and "
                                      "spectrum_bins > 0 and 0 < spectrum_min_hz < spectrum_max_hz");
    hipStream_t stream = static_cast<hipStream_t>(stream_v);
    GRT_HIP(c, hipSetDevice(c->device));

    // write_output.hpp:90-103: <dir>/<name>_step_<s>_rank0.npz
    const std::string dir = run->output_dir;
    const std::string stem = (dir.empty() ? std::string() : dir + "/") + run->name;
    const uint64_t stride = run->output_stride;
    struct Field {
        const char* name;
        size_t member;   // offset of the pointer in grt_photons
        size_t elem;
    };
    static const Field kFields[] = {
#define GRT_FIELD(f, e) {#f, offsetof(grt_photons, f), e}
        GRT_FIELD(id, 8), GRT_FIELD(frequency, 8), GRT_FIELD(emission_frame_energy, 8),
        GRT_FIELD(x0, 8), GRT_FIELD(x1, 8), GRT_FIELD(x2, 8), GRT_FIELD(x3, 8),
        GRT_FIELD(k0, 8), GRT_FIELD(k1, 8), GRT_FIELD(k2, 8), GRT_FIELD(k3, 8),
        GRT_FIELD(I, 8), GRT_FIELD(Q, 8), GRT_FIELD(U, 8), GRT_FIELD(V, 8),
        GRT_FIELD(dlambda, 8), GRT_FIELD(terminate, 1), GRT_FIELD(phase, 1),
        GRT_FIELD(theta_disp, 8), GRT_FIELD(phi_disp, 8)
#undef GRT_FIELD
    };
    auto member = [](const grt_photons& p, size_t off) -> void*& {
        return *reinterpret_cast<void**>(reinterpret_cast<char*>(const_cast<grt_photons*>(&p)) + off);
    };
    std::vector<std::vector<unsigned char>> host(sel.size());
    grt_timers tm{};
    tm.last_archive_step = -1;
    const std::function<int(int64_t)> on_check = [&](int64_t step) -> int {
        grt_photons h{};
        for (size_t q = 0; q < sel.size(); ++q) {
            const Field* f = nullptr;
            for (const Field& k : kFields)
                if (sel[q] == k.name) f = &k;
            if (!f) return fail(c, GRT_E_INVALID, "unknown output variable: " + sel[q]);
            // one contiguous copy per array (a pitched copy of single
            // elements crawls); the writer takes every stride-th element
            host[q].resize(std::max<uint64_t>(n, 1) * f->elem);
            const void* src = member(*dev, f->member);
            if (n > 0)
                GRT_HIP(c, hipMemcpyAsync(host[q].data(), src, n * f->elem,
                                          hipMemcpyDeviceToHost, stream));
            member(h, f->member) = host[q].data();
        }
        GRT_HIP(c, hipStreamSynchronize(stream));
        const std::string path = stem + "_step_" + std::to_string(step) + "_rank0.npz";
        if (grt_write_photons_npz(path.c_str(), &h, n, run->output_variables[0]
                                                          ? run->output_variables : nullptr,
                                  stride) != GRT_OK)
            return fail(c, GRT_E_INVALID, grt_last_error(nullptr));
        ++tm.archives;
        tm.last_archive_step = step;
        return GRT_OK;
    };
    const int rc = trace_impl(c, dev, n, run->max_steps, run->termination_percent,
                              (int64_t)run->output_interval, stream_v, stats, 0, false, &on_check,
                              &tm);
    if (timers) *timers = tm;
    if (rc != GRT_OK || !image) return rc;

    // write_image_products (write_output.hpp:107-142)
    const int64_t nx = (int64_t)run->image_nx, ny = (int64_t)run->image_ny;
    const int64_t bins = (int64_t)run->spectrum_bins;
    const size_t need = (size_t)(4 * nx * ny + 2 * bins + 1);
    if (need > c->img_cap) {
        if (c->d_img) (void)hipFree(c->d_img);
        c->d_img = nullptr;
        c->img_cap = 0;
        GRT_HIP(c, hipMalloc(&c->d_img, need * sizeof(double)));
        c->img_cap = need;
    }
    double* dimg = c->d_img;
    double* dnu = dimg + 4 * nx * ny;
    double* dsp = dnu + bins;
    uint64_t* darr = reinterpret_cast<uint64_t*>(dsp + bins);
    int prc = grt_image_products_device(c, dev, n, nx, ny, bins, run->spectrum_min_hz,
                                        run->spectrum_max_hz, dimg, dnu, dsp, darr, stream_v);
    if (prc != GRT_OK) return prc;
    std::vector<double> img((size_t)(4 * nx * ny)), nu((size_t)bins), sp((size_t)bins);
    uint64_t arrived = 0;
    GRT_HIP(c, hipMemcpy(img.data(), dimg, img.size() * sizeof(double), hipMemcpyDeviceToHost));
    GRT_HIP(c, hipMemcpy(nu.data(), dnu, nu.size() * sizeof(double), hipMemcpyDeviceToHost));
    GRT_HIP(c, hipMemcpy(sp.data(), dsp, sp.size() * sizeof(double), hipMemcpyDeviceToHost));
    GRT_HIP(c, hipMemcpy(&arrived, darr, sizeof arrived, hipMemcpyDeviceToHost));
    const std::string path = stem + "_image_products.npz";
    if (grt_write_image_products_npz(path.c_str(), nx, ny, bins, img.data(), nu.data(), sp.data(),
                                     arrived) != GRT_OK)
        return fail(c, GRT_E_INVALID, grt_last_error(nullptr));
    return GRT_OK;
}

int grt_trace(grt_ctx* c, const grt_photons* host, uint64_t n, int32_t max_steps,
              double termination_percent, int64_t output_interval, grt_stats* stats)
{
    if (!c) return fail(nullptr, GRT_E_INVALID, "ctx must not be NULL");
    if (!host || (n > 0 && !photons_complete(host))) return fail(c, GRT_E_INVALID, "every photon array must be non-NULL");
    GRT_HIP(c, hipSetDevice(c->device));
    grt_photons d{};
    int rc = stage_photons(c, n, &d);
    if (rc != GRT_OK) return rc;
    rc = copy_photons(c, &d, host, n, hipMemcpyHostToDevice);
    if (rc != GRT_OK) return rc;
    rc = grt_trace_device(c, &d, n, max_steps, termination_percent, output_interval, nullptr, stats);
    if (rc != GRT_OK) return rc;
    return copy_photons(c, host, &d, n, hipMemcpyDeviceToHost);
}

namespace {

grt::CameraInit camera_init(const grt_params& p, double initial_step)
{
    // initialize_photons.hpp:94-108 and :130-136, :160
    grt::CameraInit C{};
    const double d = p.camera_distance, th = p.camera_theta, ph = p.camera_phi;
    C.c[0] = d * std::sin(th) * std::cos(ph);
    C.c[1] = d * std::sin(th) * std::sin(ph);
    C.c[2] = d * std::cos(th);
    C.t[0] = std::cos(th) * std::cos(ph);
    C.t[1] = std::cos(th) * std::sin(ph);
    C.t[2] = -std::sin(th);
    C.f[0] = -std::sin(ph);
    C.f[1] = std::cos(ph);
    C.f[2] = 0.0;
    double kx = -C.c[0], ky = -C.c[1], kz = -C.c[2];
    const double kn = std::sqrt(kx * kx + ky * ky + kz * kz);
    kx /= kn;
    ky /= kn;
    kz /= kn;
    C.k[0] = kx;
    C.k[1] = ky;
    C.k[2] = kz;
    C.dlambda0 = initial_step * d;
    return C;
}

}  // namespace

int grt_init_image_device(grt_ctx* c, const grt_photons* dev, uint64_t n, uint64_t id_offset,
                          double initial_step, const double* screen_u, const double* screen_v,
                          void* stream_v)
{
    if (!c) return fail(nullptr, GRT_E_INVALID, "ctx must not be NULL");
    if (!dev || (n > 0 && !photons_complete(dev))) return fail(c, GRT_E_INVALID, "every photon array must be non-NULL");
    if ((screen_u == nullptr) != (screen_v == nullptr))
        return fail(c, GRT_E_INVALID, "screen_u and screen_v must both be given or both be NULL");
    if (!std::isfinite(initial_step)) return fail(c, GRT_E_INVALID, "initial_step must be finite");
    GRT_HIP(c, hipSetDevice(c->device));
    const grt::CameraInit C = camera_init(c->params, initial_step);
    const PhotonPtrs p = to_ptrs(*dev, nullptr);
    const grt_params& q = c->params;
    GRT_HIP(c, grt::launch_init_image(p, n, id_offset, C, q.a_bh, q.m_bh, q.camera_width,
                                      q.camera_height, q.seed, (q.rng_stream << 1) | 1u, screen_u,
                                      screen_v, static_cast<hipStream_t>(stream_v)));
    return GRT_OK;
}

int grt_init_image(grt_ctx* c, const grt_photons* host, uint64_t n, uint64_t id_offset,
                   double initial_step, const double* screen_u, const double* screen_v)
{
    if (!c) return fail(nullptr, GRT_E_INVALID, "ctx must not be NULL");
    if (!host || (n > 0 && !photons_complete(host))) return fail(c, GRT_E_INVALID, "every photon array must be non-NULL");
    GRT_HIP(c, hipSetDevice(c->device));
    grt_photons d{};
    int rc = stage_photons(c, n, &d);
    if (rc != GRT_OK) return rc;
    double* dsu = nullptr;
    if (screen_u) {
        GRT_HIP(c, hipMalloc(&dsu, 2 * std::max<uint64_t>(n, 1) * sizeof(double)));
        hipError_t e = hipMemcpy(dsu, screen_u, n * sizeof(double), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(dsu + n, screen_v, n * sizeof(double), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipFree(dsu);
            return hip_fail(c, e, "hipMemcpy(screen)");
        }
    }
    rc = grt_init_image_device(c, &d, n, id_offset, initial_step, dsu, dsu ? dsu + n : nullptr,
                               nullptr);
    if (rc == GRT_OK) {
        const hipError_t e = hipDeviceSynchronize();
        if (e != hipSuccess) rc = hip_fail(c, e, "init_image");
    }
    if (dsu) (void)hipFree(dsu);
    if (rc != GRT_OK) return rc;
    return copy_photons(c, host, &d, n, hipMemcpyDeviceToHost);
}

namespace {

// Packet offsets of every cell (initialize_photons.hpp:32-62) into c->d_disk
// (offsets start at d_disk + ncell + 1); returns the total through *total.
int disk_offsets(grt_ctx* c, uint64_t* total, hipStream_t stream)
{
    if (!c->has_grid) return fail(c, GRT_E_NOGRID, "disk mode needs the fluid grid (grt_set_grid)");
    if (check_grid_fresh(c) != GRT_OK) return GRT_E_STATE;
    const uint64_t nc = c->ncell;
    if (!c->d_disk) {
        GRT_HIP(c, hipMalloc(&c->d_disk, 2 * (nc + 1) * sizeof(unsigned long long)));
        size_t bytes = 0;
        GRT_HIP(c, grt::disk_offsets(nullptr, nc, 0.0, nullptr, nullptr, nullptr, &bytes, stream));
        GRT_HIP(c, hipMalloc(&c->d_disk_temp, std::max<size_t>(bytes, 16)));
        c->disk_temp_bytes = bytes;
    }
    unsigned long long* counts = c->d_disk;
    unsigned long long* off = c->d_disk + nc + 1;
    GRT_HIP(c, grt::disk_offsets(c->d_rho_code, nc, c->params.superphoton_count_normalization,
                                 counts, off, c->d_disk_temp, &c->disk_temp_bytes, stream));
    unsigned long long t = 0;
    GRT_HIP(c, hipMemcpyAsync(&t, off + nc, sizeof t, hipMemcpyDeviceToHost, stream));
    GRT_HIP(c, hipStreamSynchronize(stream));
    *total = t;
    return GRT_OK;
}

}  // namespace

int grt_disk_photon_count(grt_ctx* c, uint64_t* total)
{
    if (!c) return fail(nullptr, GRT_E_INVALID, "ctx must not be NULL");
    if (!total) return fail(c, GRT_E_INVALID, "total must not be NULL");
    GRT_HIP(c, hipSetDevice(c->device));
    return disk_offsets(c, total, nullptr);
}

int grt_init_disk_device(grt_ctx* c, const grt_photons* dev, uint64_t n, uint64_t id_offset,
                         double initial_step, void* stream_v)
{
    if (!c) return fail(nullptr, GRT_E_INVALID, "ctx must not be NULL");
    if (!dev || (n > 0 && !photons_complete(dev))) return fail(c, GRT_E_INVALID, "every photon array must be non-NULL");
    if (!std::isfinite(initial_step)) return fail(c, GRT_E_INVALID, "initial_step must be finite");
    hipStream_t stream = static_cast<hipStream_t>(stream_v);
    GRT_HIP(c, hipSetDevice(c->device));
    uint64_t total = 0;
    const int rc = disk_offsets(c, &total, stream);
    if (rc != GRT_OK) return rc;
    if (n > total)
        return fail(c, GRT_E_INVALID, "more photons requested than the grid emits (" +
                                          std::to_string(total) + ")");
    const PhotonPtrs p = to_ptrs(*dev, nullptr);
    GRT_HIP(c, grt::launch_emit_disk(p, n, id_offset, initial_step, make_grid_view(c), c->sp,
                                     (c->params.rng_stream << 1) | 1u,
                                     c->d_disk + c->ncell + 1, c->ncell, stream));
    return GRT_OK;
}

int grt_init_disk(grt_ctx* c, const grt_photons* host, uint64_t n, uint64_t id_offset,
                  double initial_step)
{
    if (!c) return fail(nullptr, GRT_E_INVALID, "ctx must not be NULL");
    if (!host || (n > 0 && !photons_complete(host))) return fail(c, GRT_E_INVALID, "every photon array must be non-NULL");
    GRT_HIP(c, hipSetDevice(c->device));
    grt_photons d{};
    int rc = stage_photons(c, n, &d);
    if (rc != GRT_OK) return rc;
    rc = grt_init_disk_device(c, &d, n, id_offset, initial_step, nullptr);
    if (rc != GRT_OK) return rc;
    GRT_HIP(c, hipDeviceSynchronize());
    return copy_photons(c, host, &d, n, hipMemcpyDeviceToHost);
}

int grt_image_products_device(grt_ctx* c, const grt_photons* dev, uint64_t n, int64_t nx,
                              int64_t ny, int64_t bins, double smin, double smax, double* image,
                              double* spec_nu, double* spec_I, uint64_t* arrived, void* stream_v)
{
    if (!c) return fail(nullptr, GRT_E_INVALID, "ctx must not be NULL");
    if (!dev || (n > 0 && !photons_complete(dev))) return fail(c, GRT_E_INVALID, "every photon array must be non-NULL");
    if (nx <= 0 || ny <= 0 || bins <= 0)
        return fail(c, GRT_E_INVALID, "output image dimensions and spectrum_bins must be positive");
    if (!(smin > 0.0) || !(smax > smin))
        return fail(c, GRT_E_INVALID, "output requires 0 < spectrum_min_hz < spectrum_max_hz");
    if (!image || !spec_nu || !spec_I || !arrived)
        return fail(c, GRT_E_INVALID, "output pointers must be non-NULL");
    hipStream_t stream = static_cast<hipStream_t>(stream_v);
    GRT_HIP(c, hipSetDevice(c->device));
    // spectrum frequency grid, observation.hpp:48-56
    const double lmin = std::log10(smin), lmax = std::log10(smax);
    std::vector<double> nu((size_t)bins);
    for (int64_t b = 0; b < bins; ++b) {
        const double fr = bins > 1 ? (double)b / (double)(bins - 1) : 0.0;
        nu[(size_t)b] = std::pow(10.0, lmin + fr * (lmax - lmin));
    }
    GRT_HIP(c, hipMemcpyAsync(spec_nu, nu.data(), nu.size() * sizeof(double),
                              hipMemcpyHostToDevice, stream));
    GRT_HIP(c, hipMemsetAsync(image, 0, (size_t)(4 * nx * ny) * sizeof(double), stream));
    GRT_HIP(c, hipMemsetAsync(spec_I, 0, (size_t)bins * sizeof(double), stream));
    GRT_HIP(c, hipMemsetAsync(arrived, 0, sizeof(uint64_t), stream));
    const PhotonPtrs p = to_ptrs(*dev, nullptr);
    GRT_HIP(c, grt::launch_image_products(p, n, nx, ny, bins, smin, smax, lmin, lmax,
                                          c->params.camera_width, c->params.camera_height, image,
                                          spec_I, reinterpret_cast<unsigned long long*>(arrived),
                                          stream));
    GRT_HIP(c, hipStreamSynchronize(stream));
    return GRT_OK;
}

int grt_image_products(grt_ctx* c, const grt_photons* host, uint64_t n, int64_t nx, int64_t ny,
                       int64_t bins, double smin, double smax, double* image, double* spec_nu,
                       double* spec_I, uint64_t* arrived)
{
    if (!c) return fail(nullptr, GRT_E_INVALID, "ctx must not be NULL");
    if (!host || (n > 0 && !photons_complete(host))) return fail(c, GRT_E_INVALID, "every photon array must be non-NULL");
    if (nx <= 0 || ny <= 0 || bins <= 0)
        return fail(c, GRT_E_INVALID, "output image dimensions and spectrum_bins must be positive");
    if (!image || !spec_nu || !spec_I || !arrived)
        return fail(c, GRT_E_INVALID, "output pointers must be non-NULL");
    GRT_HIP(c, hipSetDevice(c->device));
    grt_photons d{};
    int rc = stage_photons(c, n, &d);
    if (rc != GRT_OK) return rc;
    rc = copy_photons(c, &d, host, n, hipMemcpyHostToDevice);
    if (rc != GRT_OK) return rc;
    const size_t need = (size_t)(4 * nx * ny + 2 * bins + 1);
    if (need > c->img_cap) {
        if (c->d_img) (void)hipFree(c->d_img);
        c->d_img = nullptr;
        c->img_cap = 0;
        GRT_HIP(c, hipMalloc(&c->d_img, need * sizeof(double)));
        c->img_cap = need;
    }
    double* dimg = c->d_img;
    double* dnu = dimg + 4 * nx * ny;
    double* dsp = dnu + bins;
    uint64_t* darr = reinterpret_cast<uint64_t*>(dsp + bins);
    rc = grt_image_products_device(c, &d, n, nx, ny, bins, smin, smax, dimg, dnu, dsp, darr, nullptr);
    if (rc != GRT_OK) return rc;
    GRT_HIP(c, hipMemcpy(image, dimg, (size_t)(4 * nx * ny) * sizeof(double), hipMemcpyDeviceToHost));
    GRT_HIP(c, hipMemcpy(spec_nu, dnu, (size_t)bins * sizeof(double), hipMemcpyDeviceToHost));
    GRT_HIP(c, hipMemcpy(spec_I, dsp, (size_t)bins * sizeof(double), hipMemcpyDeviceToHost));
    GRT_HIP(c, hipMemcpy(arrived, darr, sizeof(uint64_t), hipMemcpyDeviceToHost));
    return GRT_OK;
}

}  // extern "C"
